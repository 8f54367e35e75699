// Coulomb/exchange digestion of unique shell quartets written by the class
// kernels (the bins whose kernel has no fused J/K epilogue: the
// runtime-class kernel), the density symmetrisation and the allocation-free
// J/K's static scratch.  The reductions themselves are in jk_device.h.
#include <hip/hip_runtime.h>

#include "../../include/cintx_hip.h"
#include "common.h"
#include "jk_device.h"

namespace cx {

// The digestion of quartet t by `nl` cooperating lanes (lane index `ln`).
__device__ __forceinline__ void jk_digest_one(
    const double* __restrict__ eri, int64_t out_base, int64_t block, const int2* __restrict__ q,
    int64_t t, int ln, int nl, const int2* __restrict__ pair_sh, const int* __restrict__ ao_off,
    const int* __restrict__ shl_l, const double* __restrict__ D, double* __restrict__ J,
    double* __restrict__ K, int nao, bool skip_bra_j = false) {
  const int2 pq = q[t];
  const int sh[4] = {pair_sh[pq.x].x, pair_sh[pq.x].y, pair_sh[pq.y].x, pair_sh[pq.y].y};
  const int ns[4] = {2 * shl_l[sh[0]] + 1, 2 * shl_l[sh[1]] + 1, 2 * shl_l[sh[2]] + 1,
                     2 * shl_l[sh[3]] + 1};
  const int off[4] = {ao_off[sh[0]], ao_off[sh[1]], ao_off[sh[2]], ao_off[sh[3]]};
  jk_digest_block(eri + out_base + t * block, ns, sh, off, ln, nl, D, J, K, nao, skip_bra_j);
}

// Large blocks: one 64-lane workgroup per quartet.
__global__ void __launch_bounds__(64) jk_digest_kernel(
    const double* __restrict__ eri, int64_t out_base, int64_t block, const int2* __restrict__ q,
    int64_t nq, const int2* __restrict__ pair_sh, const int* __restrict__ ao_off,
    const int* __restrict__ shl_l, const double* __restrict__ D, double* __restrict__ J,
    double* __restrict__ K, int nao) {
  for (int64_t t = blockIdx.x; t < nq; t += gridDim.x)
    jk_digest_one(eri, out_base, block, q, t, threadIdx.x, blockDim.x, pair_sh, ao_off, shl_l, D,
                  J, K, nao);
}

// Small blocks: one lane per quartet.
__global__ void __launch_bounds__(256) jk_digest_lane_kernel(
    const double* __restrict__ eri, int64_t out_base, int64_t block, const int2* __restrict__ q,
    int64_t nq, const int2* __restrict__ pair_sh, const int* __restrict__ ao_off,
    const int* __restrict__ shl_l, const double* __restrict__ D, double* __restrict__ J,
    double* __restrict__ K, int nao) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < nq;
  // consecutive quartets of a bin mostly share their bra pair; when every
  // live lane of the wavefront does, the J[I,J] sums are added across the
  // wavefront first and issued as one atomic per element instead of up to
  // 64 atomics on the same address (lane 0 is live whenever any lane is)
  const int bra = live ? q[t].x : -1;
  const int bra0 = __shfl(bra, 0);
  const bool shared = __all(!live || bra == bra0) && bra0 >= 0;
  if (live)
    jk_digest_one(eri, out_base, block, q, t, 0, 1, pair_sh, ao_off, shl_l, D, J, K, nao, shared);
  if (shared) {
    const int si = pair_sh[bra0].x, sj = pair_sh[bra0].y;
    const int ni = 2 * shl_l[si] + 1, nj = 2 * shl_l[sj] + 1;
    const int oi = ao_off[si], oj = ao_off[sj];
    int nk = 0, nl = 0, ok = 0, ol = 0;
    double wgt = 0.0;
    const double* blk = eri;
    if (live) {
      const int2 pq = q[t];
      const int sk = pair_sh[pq.y].x, sl = pair_sh[pq.y].y;
      nk = 2 * shl_l[sk] + 1;
      nl = 2 * shl_l[sl] + 1;
      ok = ao_off[sk];
      ol = ao_off[sl];
      wgt = 2.0 * jk_weight(si, sj, sk, sl);
      blk = eri + out_base + t * block;
    }
    const int st2 = ni * nj, st3 = ni * nj * nk;
    for (int o = 0; o < ni * nj; ++o) {  // uniform over the wavefront
      double sum = 0.0;
      for (int ml = 0; ml < nl; ++ml)
        for (int mk = 0; mk < nk; ++mk)
          sum += blk[o + mk * st2 + ml * st3] * D[(size_t)(ok + mk) * nao + ol + ml];
      sum *= wgt;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d);
      if ((threadIdx.x & 63) == 0 && sum != 0.0)
        atomicAdd(&J[(size_t)(oi + o % ni) * nao + oj + o / ni], sum);
    }
  }
}

hipError_t launch_jk_digest(const double* eri, int64_t out_base, int64_t block, const int2* q,
                            int64_t nq, const int2* pair_sh, const int* ao_off, const int* shl_l,
                            const double* D, double* J, double* K, int nao, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  if (block <= 45) {  // up to (d p|p s): a lane per quartet
    const int64_t grid = (nq + 255) / 256;
    hipLaunchKernelGGL(jk_digest_lane_kernel, dim3((unsigned)grid), dim3(256), 0, s, eri,
                       out_base, block, q, nq, pair_sh, ao_off, shl_l, D, J, K, nao);
  } else {
    const int grid = (int)std::min<int64_t>(nq, 65536);
    hipLaunchKernelGGL(jk_digest_kernel, dim3(grid), dim3(64), 0, s, eri, out_base, block, q,
                       nq, pair_sh, ao_off, shl_l, D, J, K, nao);
  }
  return hipGetLastError();
}

// A <- scale (A + A^T), one thread per element pair p <= q (in place).
__global__ void __launch_bounds__(256) jk_symmetrize_kernel(double* __restrict__ A, int n,
                                                            double scale) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int p = (int)(e / n), q = (int)(e % n);
  if (p > q) return;
  const double s = scale * (A[(size_t)p * n + q] + A[(size_t)q * n + p]);
  A[(size_t)p * n + q] = s;
  A[(size_t)q * n + p] = s;
}

// dst <- (src + src^T) / 2 (out of place)
__global__ void __launch_bounds__(256) jk_sym_copy_kernel(const double* __restrict__ src,
                                                          double* __restrict__ dst, int n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int p = (int)(e / n), q = (int)(e % n);
  dst[e] = 0.5 * (src[(size_t)p * n + q] + src[(size_t)q * n + p]);
}

hipError_t launch_jk_sym_copy(const double* src, double* dst, int n, hipStream_t s) {
  const int64_t nn = (int64_t)n * n;
  if (nn == 0) return hipSuccess;
  hipLaunchKernelGGL(jk_sym_copy_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, src,
                     dst, n);
  return hipGetLastError();
}

// Workspace of the allocation-free J/K entry point (cintx_plan_jk_noalloc):
// the symmetrised density lives here instead of in a per-call allocation.
// Each device holds its own copy from library load until the process ends;
// host code reaches it only through the address hipGetSymbolAddress returns
// (jk_scratch_address), never through the symbol itself.
constexpr size_t kJkScratchBytes = CINTX_JK_NOALLOC_MAX_BYTES;
static __device__ __align__(256) double jk_scratch[kJkScratchBytes / sizeof(double)];

hipError_t jk_scratch_address(double** ptr) {
  return hipGetSymbolAddress((void**)ptr, HIP_SYMBOL(jk_scratch));
}

size_t jk_scratch_bytes() { return kJkScratchBytes; }

hipError_t launch_jk_symmetrize(double* A, int n, double scale, hipStream_t s) {
  const int64_t nn = (int64_t)n * n;
  if (nn == 0) return hipSuccess;
  hipLaunchKernelGGL(jk_symmetrize_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, A,
                     n, scale);
  return hipGetLastError();
}

}  // namespace cx
