// Learned initializer for the Rys rule: a 2x64 tanh MLP mapping
// (log x, nroots) to initial roots and weights, trained on the device by the
// kernels below (forward + backward + Adam, FP32), followed by FP64 Newton
// steps on the moment equations sum_i w_i s_i^m = F_m(x), m = 0..2n-1.  The
// Chebyshev-table rule (the plain solver) judges the polished rule: where any
// root u_i or weight w_i differs from it by more than 1e-14 relative, the
// table rule is returned instead.
//
// Network: z = ((ln x + 3.45)/10.35, (n - 7)/6); h1 = tanh(W1 z + b1) (64);
// h2 = tanh(W2 h1 + b2) (64); y = W3 h2 + b3 (26): y[r] ~ ln u_r and
// y[13 + r] ~ ln w_r for r < n (u = s/(1-s)).  Parameters are one float array
// laid out W1[64][2] b1[64] W2[64][64] b2[64] W3[26][64] b3[26].
#include <hip/hip_runtime.h>

#include <vector>

#include "eri_device.h"

namespace cx {

constexpr int kH = 64;
constexpr int kIn = 2;
constexpr int kOut = 26;
constexpr int kOW1 = 0;
constexpr int kOB1 = kOW1 + kH * kIn;
constexpr int kOW2 = kOB1 + kH;
constexpr int kOB2 = kOW2 + kH * kH;
constexpr int kOW3 = kOB2 + kH;
constexpr int kOB3 = kOW3 + kOut * kH;
constexpr int kNParam = kOB3 + kOut;

__host__ __device__ inline float mlp_z0(double x) { return (float)((log(x) + 3.45) / 10.35); }
__host__ __device__ inline float mlp_z1(int n) { return (float)((n - 7) / 6.0); }

// Per-sample forward and backward pass of the squared-error loss over the
// n valid root and weight outputs; stores the activations and deltas the
// weight-gradient kernel needs.  One thread per sample, parameters in LDS.
__global__ void __launch_bounds__(256) mlp_fwd_bwd(const float* __restrict__ prm,
                                                   const float* __restrict__ z,
                                                   const int* __restrict__ nr,
                                                   const float* __restrict__ tgt, int B,
                                                   float* __restrict__ H1, float* __restrict__ H2,
                                                   float* __restrict__ D1, float* __restrict__ D2,
                                                   float* __restrict__ DY,
                                                   float* __restrict__ loss) {
  __shared__ float sp[kNParam];
  for (int i = threadIdx.x; i < kNParam; i += blockDim.x) sp[i] = prm[i];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float z0 = z[2 * b], z1 = z[2 * b + 1];
  const int n = nr[b];
  float h1[kH], h2[kH];
  for (int i = 0; i < kH; ++i)
    h1[i] = tanhf(sp[kOW1 + 2 * i] * z0 + sp[kOW1 + 2 * i + 1] * z1 + sp[kOB1 + i]);
  for (int i = 0; i < kH; ++i) {
    float acc = sp[kOB2 + i];
    for (int j = 0; j < kH; ++j) acc += sp[kOW2 + i * kH + j] * h1[j];
    h2[i] = tanhf(acc);
  }
  float dy[kOut];
  float l = 0;
  const float inv = 1.0f / (2.0f * n);
  for (int o = 0; o < kOut; ++o) {
    const int r = o < 13 ? o : o - 13;
    float y = sp[kOB3 + o];
    for (int j = 0; j < kH; ++j) y += sp[kOW3 + o * kH + j] * h2[j];
    const float d = (r < n) ? (y - tgt[(size_t)b * kOut + o]) : 0.0f;
    l += d * d * inv;
    dy[o] = 2.0f * d * inv;
  }
  loss[b] = l;
  float d2[kH];
  for (int j = 0; j < kH; ++j) {
    float acc = 0;
    for (int o = 0; o < kOut; ++o) acc += sp[kOW3 + o * kH + j] * dy[o];
    d2[j] = acc * (1.0f - h2[j] * h2[j]);
  }
  for (int j = 0; j < kH; ++j) {
    float acc = 0;
    for (int i = 0; i < kH; ++i) acc += sp[kOW2 + i * kH + j] * d2[i];
    D1[(size_t)j * B + b] = acc * (1.0f - h1[j] * h1[j]);
  }
  for (int j = 0; j < kH; ++j) {
    H1[(size_t)j * B + b] = h1[j];
    H2[(size_t)j * B + b] = h2[j];
    D2[(size_t)j * B + b] = d2[j];
  }
  for (int o = 0; o < kOut; ++o) DY[(size_t)o * B + b] = dy[o];
}

// Batch-mean gradients: one thread per parameter, a dot product over the
// batch of the stored deltas and activations (columns are contiguous per
// feature, so a wavefront reads coalesced rows).
__global__ void __launch_bounds__(256) mlp_grad(const float* __restrict__ z, int B,
                                                const float* __restrict__ H1,
                                                const float* __restrict__ H2,
                                                const float* __restrict__ D1,
                                                const float* __restrict__ D2,
                                                const float* __restrict__ DY,
                                                float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  const float* a = nullptr;  // delta column
  const float* c = nullptr;  // activation column (nullptr: bias)
  int zcol = -1;
  if (p < kOB1) {
    a = D1 + (size_t)(p / 2) * B;
    zcol = p % 2;
  } else if (p < kOW2) {
    a = D1 + (size_t)(p - kOB1) * B;
  } else if (p < kOB2) {
    a = D2 + (size_t)((p - kOW2) / kH) * B;
    c = H1 + (size_t)((p - kOW2) % kH) * B;
  } else if (p < kOW3) {
    a = D2 + (size_t)(p - kOB2) * B;
  } else if (p < kOB3) {
    a = DY + (size_t)((p - kOW3) / kH) * B;
    c = H2 + (size_t)((p - kOW3) % kH) * B;
  } else {
    a = DY + (size_t)(p - kOB3) * B;
  }
  double acc = 0;
  for (int b = 0; b < B; ++b) {
    float f = c ? c[b] : (zcol >= 0 ? z[2 * b + zcol] : 1.0f);
    acc += (double)a[b] * f;
  }
  grad[p] = (float)(acc / B);
}

__global__ void mlp_adam(float* __restrict__ prm, const float* __restrict__ grad,
                         float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                         float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  const float g = grad[p];
  const float mm = b1 * m[p] + (1 - b1) * g;
  const float vv = b2 * v[p] + (1 - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  prm[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// Training samples: x log-uniform in [1e-6, 1e3], n uniform in 1..13
// (counter-based hash, reproducible), targets from the table rule.
__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void mlp_make_samples(RysTable tab, uint64_t seed, int64_t first, int B,
                                 float* __restrict__ z, int* __restrict__ nr,
                                 float* __restrict__ tgt, double* __restrict__ xs) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const uint64_t h1 = mix64(seed ^ (uint64_t)(first + b) * 2ull);
  const uint64_t h2 = mix64(seed ^ ((uint64_t)(first + b) * 2ull + 1ull));
  const double u01 = (h1 >> 11) * (1.0 / 9007199254740992.0);
  const double x = exp(log(1e-6) + u01 * (log(1e3) - log(1e-6)));
  const int n = 1 + (int)(h2 % 13ull);
  z[2 * b] = mlp_z0(x);
  z[2 * b + 1] = mlp_z1(n);
  nr[b] = n;
  if (xs) xs[b] = x;
  for (int o = 0; o < kOut; ++o) tgt[(size_t)b * kOut + o] = 0.0f;
  for (int r = 0; r < n; ++r) {
    double s, q, w;
    rys_eval(tab, n, r, x, s, q, w);
    tgt[(size_t)b * kOut + r] = (float)log(s / q);
    tgt[(size_t)b * kOut + 13 + r] = (float)log(w);
  }
}

// FP64 Boys function F_m(x), m = 0..mmax (series + downward recursion below
// x = 30 + m, erf start + upward recursion above; fmt.c:186-226 regimes).
__device__ void boys_fm(int mmax, double x, double* f) {
  if (x < 30.0 + mmax) {
    const double e = exp(-x);
    double b = mmax + 0.5, term = 1.0 / b, s = term;
    for (int k = 0; k < 400; ++k) {
      b += 1.0;
      term *= x / b;
      s += term;
      if (term < s * 1e-17) break;
    }
    f[mmax] = 0.5 * e * s;
    for (int i = mmax; i > 0; --i) f[i - 1] = (2.0 * x * f[i] + e) / (2 * i - 1);
  } else {
    const double sx = sqrt(x), e = exp(-x);
    f[0] = 0.886226925452758014 / sx * erf(sx);
    for (int i = 1; i <= mmax; ++i) f[i] = ((2 * i - 1) * f[i - 1] - e) / (2.0 * x);
  }
}

// Network inference (FP32) -> (s, w) initial rule, or the asymptotic
// Hermite rule when mode == 1 (the initializer without the network).
__device__ void mlp_initial(const float* __restrict__ sp, const RysTable& tab, int n, double x,
                            int mode, double* s, double* w) {
  if (mode == 1) {
    const double* L = tab.large + n * (n - 1);
    for (int r = 0; r < n; ++r) {
      double u = L[r] / x;
      s[r] = u / (1.0 + u);
      w[r] = L[n + r] / sqrt(x);
    }
    return;
  }
  const float z0 = mlp_z0(x), z1 = mlp_z1(n);
  float h1[kH], h2[kH];
  for (int i = 0; i < kH; ++i)
    h1[i] = tanhf(sp[kOW1 + 2 * i] * z0 + sp[kOW1 + 2 * i + 1] * z1 + sp[kOB1 + i]);
  for (int i = 0; i < kH; ++i) {
    float acc = sp[kOB2 + i];
    for (int j = 0; j < kH; ++j) acc += sp[kOW2 + i * kH + j] * h1[j];
    h2[i] = tanhf(acc);
  }
  for (int r = 0; r < n; ++r) {
    float yu = sp[kOB3 + r], yw = sp[kOB3 + 13 + r];
    for (int j = 0; j < kH; ++j) {
      yu += sp[kOW3 + r * kH + j] * h2[j];
      yw += sp[kOW3 + (13 + r) * kH + j] * h2[j];
    }
    const double u = exp((double)yu);
    s[r] = u / (1.0 + u);
    w[r] = exp((double)yw);
  }
}

// Newton steps on the moment equations; returns false when an update is not
// finite or leaves the domain (0 < s < 1, w > 0).
__device__ bool newton_moments(int n, double x, double* s, double* w, int steps) {
  double F[2 * kMaxRoots];
  boys_fm(2 * n - 1, x, F);
  double J[2 * kMaxRoots][2 * kMaxRoots + 1];
  for (int it = 0; it < steps; ++it) {
    const int N = 2 * n;
    for (int m = 0; m < N; ++m) {
      double res = -F[m];
      for (int i = 0; i < n; ++i) {
        const double sm1 = m > 0 ? pow(s[i], m - 1) : 0.0;
        const double sm = m > 0 ? sm1 * s[i] : 1.0;
        res += w[i] * sm;
        J[m][i] = sm / F[m];
        J[m][n + i] = m * w[i] * sm1 / F[m];
      }
      J[m][N] = -res / F[m];
    }
    // Gaussian elimination with partial pivoting
    for (int c = 0; c < N; ++c) {
      int piv = c;
      for (int r = c + 1; r < N; ++r)
        if (fabs(J[r][c]) > fabs(J[piv][c])) piv = r;
      if (piv != c)
        for (int k = c; k <= N; ++k) {
          double tmp = J[c][k];
          J[c][k] = J[piv][k];
          J[piv][k] = tmp;
        }
      const double d = J[c][c];
      if (d == 0.0) return false;
      for (int r = c + 1; r < N; ++r) {
        const double f = J[r][c] / d;
        for (int k = c; k <= N; ++k) J[r][k] -= f * J[c][k];
      }
    }
    double dv[2 * kMaxRoots];
    for (int r = N - 1; r >= 0; --r) {
      double acc = J[r][N];
      for (int k = r + 1; k < N; ++k) acc -= J[r][k] * dv[k];
      dv[r] = acc / J[r][r];
    }
    for (int i = 0; i < n; ++i) {
      w[i] += dv[i];
      s[i] += dv[n + i];
    }
  }
  for (int i = 0; i < n; ++i)
    if (!(s[i] > 0.0 && s[i] < 1.0 && w[i] > 0.0)) return false;
  return true;
}

__global__ void mlp_eval_kernel(const float* __restrict__ prm, RysTable tab, int n,
                                const double* __restrict__ x, int64_t nx, int steps, int mode,
                                double* __restrict__ u, double* __restrict__ w,
                                int* __restrict__ status, double* __restrict__ perr) {
  __shared__ float sp[kNParam];
  for (int i = threadIdx.x; i < kNParam; i += blockDim.x) sp[i] = prm[i];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const double xi = x[i];
  // the plain solver: table rule, u = s/q with q = 1 - s to full accuracy
  double ut[kMaxRoots], wt[kMaxRoots];
  for (int r = 0; r < n; ++r) {
    double sr, qr;
    rys_eval(tab, n, r, xi, sr, qr, wt[r]);
    ut[r] = sr / qr;
  }
  bool ok = false;
  double err = -1.0;  // mode 2: not polished
  double s[kMaxRoots], ww[kMaxRoots];
  if (mode != 2) {
    if (mode == 3) {  // start from the table rule itself: the polish's own floor
      for (int r = 0; r < n; ++r) {
        s[r] = ut[r] / (1.0 + ut[r]);
        ww[r] = wt[r];
      }
    } else {
      mlp_initial(sp, tab, n, xi, mode, s, ww);
    }
    err = INFINITY;
    if (newton_moments(n, xi, s, ww, steps)) {
      err = 0;
      for (int r = 0; r < n; ++r) {
        const double ur = s[r] / (1.0 - s[r]);
        err = fmax(err, fabs(ur - ut[r]) / ut[r]);
        err = fmax(err, fabs(ww[r] - wt[r]) / wt[r]);
      }
      if (!(err == err)) err = INFINITY;
      ok = err <= 1e-14;
    }
  }
  for (int r = 0; r < n; ++r) {
    u[i * n + r] = ok ? s[r] / (1.0 - s[r]) : ut[r];
    w[i * n + r] = ok ? ww[r] : wt[r];
  }
  status[i] = ok ? 1 : 0;
  if (perr) perr[i] = err;
}

hipError_t mlp_train_device(const RysTable& tab, float* prm, int64_t nsamples, int epochs,
                            int batch, float lr, uint64_t seed, double* final_loss,
                            hipStream_t st) {
  float *z, *tgt, *H1, *H2, *D1, *D2, *DY, *loss, *grad, *m, *v;
  int* nr;
  const size_t B = (size_t)batch;
  hipError_t e;
#define MLP_ALLOC(ptr, bytes) \
  if ((e = hipMalloc(&ptr, bytes)) != hipSuccess) return e
  MLP_ALLOC(z, B * 2 * 4);
  MLP_ALLOC(tgt, B * kOut * 4);
  MLP_ALLOC(H1, B * kH * 4);
  MLP_ALLOC(H2, B * kH * 4);
  MLP_ALLOC(D1, B * kH * 4);
  MLP_ALLOC(D2, B * kH * 4);
  MLP_ALLOC(DY, B * kOut * 4);
  MLP_ALLOC(loss, B * 4);
  MLP_ALLOC(grad, kNParam * 4);
  MLP_ALLOC(m, kNParam * 4);
  MLP_ALLOC(v, kNParam * 4);
  MLP_ALLOC(nr, B * 4);
#undef MLP_ALLOC
  hipMemsetAsync(m, 0, kNParam * 4, st);
  hipMemsetAsync(v, 0, kNParam * 4, st);
  const int64_t steps_per_epoch = (nsamples + batch - 1) / batch;
  const dim3 gb((unsigned)((B + 255) / 256)), gp((kNParam + 255) / 256);
  float b1 = 0.9f, b2 = 0.999f, c1 = 1, c2 = 1;
  int64_t step = 0;
  std::vector<float> hl(B);
  double last = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t k = 0; k < steps_per_epoch; ++k, ++step) {
      hipLaunchKernelGGL(mlp_make_samples, gb, dim3(256), 0, st, tab, seed,
                         (int64_t)(k * batch), batch, z, nr, tgt, (double*)nullptr);
      hipLaunchKernelGGL(mlp_fwd_bwd, gb, dim3(256), 0, st, prm, z, nr, tgt, batch, H1, H2, D1,
                         D2, DY, loss);
      hipLaunchKernelGGL(mlp_grad, gp, dim3(256), 0, st, z, batch, H1, H2, D1, D2, DY, grad);
      c1 *= b1;
      c2 *= b2;
      // cosine decay of the learning rate over the run
      const float frac = (float)step / (float)(epochs * steps_per_epoch);
      const float lrs = lr * 0.5f * (1.0f + cosf(3.14159265f * frac));
      hipLaunchKernelGGL(mlp_adam, gp, dim3(256), 0, st, prm, grad, m, v, lrs, b1, b2, 1e-8f,
                         1.0f - c1, 1.0f - c2);
    }
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(hl.data(), loss, B * 4, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  for (float f : hl) last += f;
  if (final_loss) *final_loss = last / B;
  for (void* p : {(void*)z, (void*)tgt, (void*)H1, (void*)H2, (void*)D1, (void*)D2, (void*)DY,
                  (void*)loss, (void*)grad, (void*)m, (void*)v, (void*)nr})
    (void)hipFree(p);
  return hipSuccess;
}

hipError_t mlp_grad_device(const float* prm, const float* z, const int* nr, const float* tgt,
                           int batch, float* grad, float* loss, hipStream_t st) {
  float *H1, *H2, *D1, *D2, *DY;
  const size_t B = (size_t)batch;
  hipError_t e;
  if ((e = hipMalloc(&H1, B * kH * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(&H2, B * kH * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(&D1, B * kH * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(&D2, B * kH * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(&DY, B * kOut * 4)) != hipSuccess) return e;
  const dim3 gb((unsigned)((B + 255) / 256)), gp((kNParam + 255) / 256);
  hipLaunchKernelGGL(mlp_fwd_bwd, gb, dim3(256), 0, st, prm, z, nr, tgt, batch, H1, H2, D1, D2,
                     DY, loss);
  hipLaunchKernelGGL(mlp_grad, gp, dim3(256), 0, st, z, batch, H1, H2, D1, D2, DY, grad);
  e = hipStreamSynchronize(st);
  for (void* p : {(void*)H1, (void*)H2, (void*)D1, (void*)D2, (void*)DY}) (void)hipFree(p);
  return e;
}

hipError_t mlp_eval_device(const float* prm, const RysTable& tab, int n, const double* x,
                           int64_t nx, int steps, int mode, double* u, double* w, int* status,
                           double* perr, hipStream_t st) {
  if (nx <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_eval_kernel, dim3((unsigned)((nx + 127) / 128)), dim3(128), 0, st, prm,
                     tab, n, x, nx, steps, mode, u, w, status, perr);
  return hipGetLastError();
}

int mlp_nparam() { return kNParam; }

}  // namespace cx
