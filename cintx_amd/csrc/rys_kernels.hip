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
