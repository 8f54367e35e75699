// Coulomb/exchange digestion of unique shell quartets:
//   J[p][q] += (pq|rs) D[r][s],  K[p][r] += (pq|rs) D[q][s]
// over every distinct index image of each stored quartet block (the 8-fold
// permutational symmetry (ij|kl) = (ji|kl) = (ij|lk) = (kl|ij), deduplicated at
// shell level, so a block whose bra shells coincide is not counted twice).
// One workgroup per quartet, lanes over the block's elements, FP64 atomics.
// The consumer side of SURVEY's J/K build (a Fock-like contraction against a
// symmetric density); the block layout is the plan's canonical one
// ([l][k][j][i], i fastest).
#include <hip/hip_runtime.h>

#include "common.h"

namespace cx {

__global__ void __launch_bounds__(256) jk_digest_kernel(
    const double* __restrict__ eri, int64_t out_base, int64_t block, const int2* __restrict__ q,
    int64_t nq, const int2* __restrict__ pair_sh, const int* __restrict__ ao_off,
    const int* __restrict__ shl_l, const double* __restrict__ D, double* __restrict__ J,
    double* __restrict__ K, int nao) {
  for (int64_t t = blockIdx.x; t < nq; t += gridDim.x) {
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
    const int n = ns[0] * ns[1] * ns[2] * ns[3];
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      const double v = blk[e];
      if (v == 0.0) continue;
      int m[4];
      m[0] = e % ns[0];
      int e1 = e / ns[0];
      m[1] = e1 % ns[1];
      int e2 = e1 / ns[1];
      m[2] = e2 % ns[2];
      m[3] = e2 / ns[2];
      int ao[4];
      for (int c = 0; c < 4; ++c) ao[c] = ao_off[sh[c]] + m[c];
      for (int a = 0; a < nimg; ++a) {
        const int p = ao[img[a][0]], qq = ao[img[a][1]], r = ao[img[a][2]], s = ao[img[a][3]];
        atomicAdd(&J[(size_t)p * nao + qq], v * D[(size_t)r * nao + s]);
        atomicAdd(&K[(size_t)p * nao + r], v * D[(size_t)qq * nao + s]);
      }
    }
  }
}

hipError_t launch_jk_digest(const double* eri, int64_t out_base, int64_t block, const int2* q,
                            int64_t nq, const int2* pair_sh, const int* ao_off, const int* shl_l,
                            const double* D, double* J, double* K, int nao, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  const int grid = (int)std::min<int64_t>(nq, 65536);
  hipLaunchKernelGGL(jk_digest_kernel, dim3(grid), dim3(64), 0, s, eri, out_base, block, q, nq,
                     pair_sh, ao_off, shl_l, D, J, K, nao);
  return hipGetLastError();
}

}  // namespace cx
