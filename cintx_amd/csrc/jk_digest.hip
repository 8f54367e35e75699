// Coulomb/exchange digestion of unique shell quartets:
//   J[p][q] += (pq|rs) D[r][s],  K[p][r] += (pq|rs) D[q][s]
// over every distinct index image of each stored quartet block (the 8-fold
// permutational symmetry (ij|kl) = (ji|kl) = (ij|lk) = (kl|ij), deduplicated at
// shell level, so a block whose bra shells coincide is not counted twice).
// One workgroup per quartet; per image the lanes reduce over the contracted
// pair and issue one FP64 atomic per J or K element of the image (instead of
// one per integral).
// The consumer side of SURVEY's J/K build (a Fock-like contraction against a
// symmetric density); the block layout is the plan's canonical one
// ([l][k][j][i], i fastest).
#include <hip/hip_runtime.h>

#include "common.h"

namespace cx {

// The digestion of quartet t by `nl` cooperating lanes (lane index `ln`).
__device__ __forceinline__ void jk_digest_one(
    const double* __restrict__ eri, int64_t out_base, int64_t block, const int2* __restrict__ q,
    int64_t t, int ln, int nl, const int2* __restrict__ pair_sh, const int* __restrict__ ao_off,
    const int* __restrict__ shl_l, const double* __restrict__ D, double* __restrict__ J,
    double* __restrict__ K, int nao) {
  const int2 pq = q[t];
  const int sh[4] = {pair_sh[pq.x].x, pair_sh[pq.x].y, pair_sh[pq.y].x, pair_sh[pq.y].y};
  const int ns[4] = {2 * shl_l[sh[0]] + 1, 2 * shl_l[sh[1]] + 1, 2 * shl_l[sh[2]] + 1,
                     2 * shl_l[sh[3]] + 1};
  // distinct shell-level images: position order of (p, q, r, s)
  int img[8][4];
  int nimg = 0;
  const int perms[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                           {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};
  for (int a = 0; a < 8; ++a) {
    bool dup = false;
    for (int b = 0; b < nimg && !dup; ++b) {
      bool same = true;
      for (int c = 0; c < 4; ++c) same = same && sh[perms[a][c]] == sh[img[b][c]];
      dup = same;
    }
    if (!dup) {
      for (int c = 0; c < 4; ++c) img[nimg][c] = perms[a][c];
      ++nimg;
    }
  }
  const double* blk = eri + out_base + t * block;
  // per image, reduce over the summed pair first, then one atomic per output
  // element: J(p,q) = sum_rs (pq|rs) D(r,s), K(p,r) = sum_qs (pq|rs) D(q,s)
  const int st[4] = {1, ns[0], ns[0] * ns[1], ns[0] * ns[1] * ns[2]};
  for (int a = 0; a < nimg; ++a) {
    const int P = img[a][0], Q = img[a][1], R = img[a][2], S = img[a][3];
    const int np = ns[P], nq2 = ns[Q], nr = ns[R], nsv = ns[S];
    const int op = ao_off[sh[P]], oq = ao_off[sh[Q]], orr = ao_off[sh[R]], os = ao_off[sh[S]];
    for (int o = ln; o < np * nq2; o += nl) {
      const int mp = o % np, mq = o / np;
      double sum = 0;
      for (int ms = 0; ms < nsv; ++ms)
        for (int mr = 0; mr < nr; ++mr)
          sum += blk[mp * st[P] + mq * st[Q] + mr * st[R] + ms * st[S]] *
                 D[(size_t)(orr + mr) * nao + os + ms];
      if (sum != 0.0) atomicAdd(&J[(size_t)(op + mp) * nao + oq + mq], sum);
    }
    for (int o = ln; o < np * nr; o += nl) {
      const int mp = o % np, mr = o / np;
      double sum = 0;
      for (int ms = 0; ms < nsv; ++ms)
        for (int mq = 0; mq < nq2; ++mq)
          sum += blk[mp * st[P] + mq * st[Q] + mr * st[R] + ms * st[S]] *
                 D[(size_t)(oq + mq) * nao + os + ms];
      if (sum != 0.0) atomicAdd(&K[(size_t)(op + mp) * nao + orr + mr], sum);
    }
  }
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
  if (t < nq) jk_digest_one(eri, out_base, block, q, t, 0, 1, pair_sh, ao_off, shl_l, D, J, K, nao);
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

}  // namespace cx
