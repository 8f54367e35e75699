// Stand-alone evaluation of the device Rys rule (the same rys_eval the class
// kernels inline), used to check the table against the oracle and as the
// plain solver that the learned initializer is compared with.
#include "eri_device.h"

namespace cx {

__global__ void rys_eval_kernel(RysTable tab, int n, const double* __restrict__ x, int64_t nx,
                                double* __restrict__ u, double* __restrict__ w) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  double xi = x[i];
  for (int r = 0; r < n; ++r) {
    double sr, qr, wr;
    rys_eval(tab, n, r, xi, sr, qr, wr);
    u[i * n + r] = sr / qr;  // libcint's u = t^2/(1-t^2)
    w[i * n + r] = wr;
  }
}

hipError_t launch_rys_eval(const RysTable& tab, int n, const double* x, int64_t nx, double* u,
                           double* w, hipStream_t s) {
  if (nx <= 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(rys_eval_kernel, dim3((unsigned)((nx + bs - 1) / bs)), dim3(bs), 0, s, tab, n,
                     x, nx, u, w);
  return hipGetLastError();
}

}  // namespace cx

namespace cx {

// Fill n doubles with one value (a kernel, not hipMemsetAsync, so that it
// can sit in a captured graph and write any bit pattern, NaN included).
__global__ void fill_kernel(double* __restrict__ dst, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = v;
}

hipError_t launch_fill(double* dst, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t want = (n + 255) / 256;
  const unsigned grid = (unsigned)(want < 4096 ? want : 4096);
  hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, s, dst, n, v);
  return hipGetLastError();
}

}  // namespace cx
