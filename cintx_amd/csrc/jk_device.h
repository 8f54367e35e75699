// Coulomb/exchange digestion of one unique shell quartet (device side),
// shared by the stand-alone digest kernels (jk_digest.hip) and the class
// kernels' fused epilogues (eri_device.h).
//
//   J[p][q] = sum_rs (pq|rs) D[r][s],  K[p][r] = sum_qs (pq|rs) D[q][s]
// for a symmetric D.  A stored quartet (IJ|KL) stands for its orbit under the
// 8-fold permutational symmetry; summed over all 8 permutations with weight
// 1/|stabiliser| (= distinct images / 8) every ordered quartet counts once.
// With D symmetric the 8 permutations give only six distinct reductions and
// their transposes:
//   A[I,J] += 2w sum_kl (ij|kl) D_kl,   A[K,L] += 2w sum_ij (ij|kl) D_ij,
//   B[I,K] += w sum_jl (ij|kl) D_jl,    B[J,K] += w sum_il (ij|kl) D_il,
//   B[I,L] += w sum_jk (ij|kl) D_jk,    B[J,L] += w sum_ik (ij|kl) D_ik,
// and J = A + A^T, K = B + B^T (jk_symmetrize).
#pragma once
#include <hip/hip_runtime.h>

namespace cx {

// 1/|stabiliser| of the shell quartet (I J|K L) under the 8-fold symmetry
__device__ __forceinline__ double jk_weight(int si, int sj, int sk, int sl) {
  const bool braket = (si == sk && sj == sl) || (si == sl && sj == sk);
  const int stab = (si == sj ? 2 : 1) * (sk == sl ? 2 : 1) * (braket ? 2 : 1);
  return 1.0 / stab;
}

// out[(o[a]+ma), (o[b]+mb)] += wgt sum_{mc, md} blk[...] D[(o[c]+mc), (o[d]+md)]
// over the output elements ln, ln + nl, ... (one FP64 atomic each)
__device__ __forceinline__ void jk_reduce(const double* blk, const int* ns, const int* st,
                                          const int* off, int a, int b, int c, int d, double wgt,
                                          const double* __restrict__ D, double* __restrict__ out,
                                          int nao, int ln, int nl) {
  const int na = ns[a], nb = ns[b], nc = ns[c], nd = ns[d];
  constexpr int kMaxN = 9;  // 2l+1 for l <= 4
  for (int o = ln; o < na * nb; o += nl) {
    const int ma = o % na, mb = o / na;
    const double* e = blk + ma * st[a] + mb * st[b];
    double sum = 0;
    for (int md = 0; md < nd; ++md) {
      // the column's density values and block values are loaded together
      // (one memory latency per column instead of one per element)
      const double* drow = D + (size_t)(off[d] + md) + (size_t)off[c] * nao;
      double dv[kMaxN], ev[kMaxN];
#pragma unroll
      for (int mc = 0; mc < kMaxN; ++mc)
        if (mc < nc) {
          dv[mc] = drow[(size_t)mc * nao];
          ev[mc] = e[mc * st[c] + md * st[d]];
        }
#pragma unroll
      for (int mc = 0; mc < kMaxN; ++mc)
        if (mc < nc) sum += ev[mc] * dv[mc];
    }
    if (sum != 0.0) atomicAdd(&out[(size_t)(off[a] + ma) * nao + off[b] + mb], wgt * sum);
  }
}

// The six reductions of one spherical block blk ([l][k][j][i], i fastest,
// extents ns, element strides st) of shells sh at AO offsets off, by `nl`
// cooperating lanes (lane index `ln`); skip_bra_j leaves A[I,J] to the
// caller.
__device__ __forceinline__ void jk_digest_block_st(const double* blk, const int* ns,
                                                   const int* st, const int* sh, const int* off,
                                                   int ln, int nl, const double* __restrict__ D,
                                                   double* __restrict__ J,
                                                   double* __restrict__ K, int nao,
                                                   bool skip_bra_j = false) {
  const double w = jk_weight(sh[0], sh[1], sh[2], sh[3]);
  if (!skip_bra_j) jk_reduce(blk, ns, st, off, 0, 1, 2, 3, 2 * w, D, J, nao, ln, nl);
  jk_reduce(blk, ns, st, off, 2, 3, 0, 1, 2 * w, D, J, nao, ln, nl);
  jk_reduce(blk, ns, st, off, 0, 2, 1, 3, w, D, K, nao, ln, nl);
  jk_reduce(blk, ns, st, off, 1, 2, 0, 3, w, D, K, nao, ln, nl);
  jk_reduce(blk, ns, st, off, 0, 3, 1, 2, w, D, K, nao, ln, nl);
  jk_reduce(blk, ns, st, off, 1, 3, 0, 2, w, D, K, nao, ln, nl);
}

// The same for a dense block (i fastest).
__device__ __forceinline__ void jk_digest_block(const double* blk, const int* ns, const int* sh,
                                                const int* off, int ln, int nl,
                                                const double* __restrict__ D,
                                                double* __restrict__ J, double* __restrict__ K,
                                                int nao, bool skip_bra_j = false) {
  const int st[4] = {1, ns[0], ns[0] * ns[1], ns[0] * ns[1] * ns[2]};
  jk_digest_block_st(blk, ns, st, sh, off, ln, nl, D, J, K, nao, skip_bra_j);
}

// Runs of equal keys over the lanes of a wavefront (every lane of the
// wavefront calls): `start` is the first lane of the lane's run, `tail`
// marks its last lane, `single` that every run is one lane long.
struct WaveRun {
  int start;
  bool tail;
  bool single;
};
__device__ __forceinline__ WaveRun wave_run(unsigned long long key) {
  const int l = __lane_id();
  const unsigned long long prev = __shfl_up(key, 1);
  const unsigned long long next = __shfl_down(key, 1);
  const unsigned long long heads = __ballot(l == 0 || prev != key);
  const unsigned long long below = heads & ((2ull << l) - 1);  // l = 63: all
  WaveRun r;
  r.start = 63 - __clzll(below);
  r.tail = l == 63 || next != key;
  r.single = heads == ~0ull;
  return r;
}
// Inclusive sum of v over the lanes [start, lane] of the lane's run (a
// segmented Hillis-Steele scan; the run's tail holds the run's total).
__device__ __forceinline__ double wave_run_sum(double v, WaveRun r) {
  if (r.single) return v;
  const int l = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double up = __shfl_up(v, d);
    if (l - d >= r.start) v += up;
  }
  return v;
}

// out[idx] += the run's total of v (one atomic, by the run's tail)
__device__ __forceinline__ void jk_put(bool live, double* __restrict__ out, size_t idx, double v,
                                       WaveRun r) {
  v = wave_run_sum(v, r);
  if (live && r.tail && v != 0.0) atomicAdd(&out[idx], v);
}

// The six reductions restricted to the (k, l) output pairs kl = k + NK l in
// [r0, r0 + nr) of a block staged in the kernel's dynamic LDS: row r holds
// the NI NJ (i, j) values of pair r0 + r at stage[base + (r NI NJ + i + NI j)
// ls] (ls: lane stride).  One lane per quartet; every lane of the wavefront
// calls (`live` false for a lane without a quartet), r0 and nr are uniform.
// Row by row: the 2 NI + 2 NJ + 1 density values a row needs are loaded
// together (independent loads, one latency per row); A[I,J] is summed over
// the rows in registers.  Lanes whose quartets share an output block (the
// bra pair for A[I,J], the ket pair for A[K,L], the shell pair for each K
// block) and sit side by side in the wavefront add their contributions
// first (wave_run_sum), and the run's last lane issues one FP64 atomic per
// element: tuples are ordered bra-major, so runs are long, and the atomics
// they replace would all hit one address.  It reads LDS through the
// dynamic-LDS symbol, not a pointer.  A kernel with one digest site inlines
// it (as a call it raised the small classes' register counts, and so
// lowered their occupancy, to what the callee and the caller's live state
// need together); one with several (row chunks, passes) calls the
// out-of-line copy below, so its code is not replicated per site.
template <int NI, int NJ, int NK, int NL>
__device__ __forceinline__ void jk_digest_rows(int base, int ls, int r0, int nr, bool live, int s0,
                                            int s1, int s2, int s3, int o0, int o1, int o2,
                                            int o3, const double* __restrict__ D,
                                            double* __restrict__ J, double* __restrict__ K,
                                            int nao) {
  extern __shared__ double jk_stage[];
  constexpr int NB = NI * NJ;
  const double w = jk_weight(s0, s1, s2, s3);
  const size_t n = (size_t)nao;
  const size_t oI = o0, oJ = o1, oK = o2, oL = o3;
  // one key per output block: its first element's index (a lane without a
  // quartet gets a key no index takes, unique per lane)
  const unsigned long long dead = ~0ull - __lane_id();
  const WaveRun rIJ = wave_run(live ? oI * n + oJ : dead);
  const WaveRun rKL = wave_run(live ? oK * n + oL : dead);
  const WaveRun rIK = wave_run(live ? oI * n + oK : dead);
  const WaveRun rIL = wave_run(live ? oI * n + oL : dead);
  const WaveRun rJK = wave_run(live ? oJ * n + oK : dead);
  const WaveRun rJL = wave_run(live ? oJ * n + oL : dead);
  double dij[NB], aij[NB];
#pragma unroll
  for (int mj = 0; mj < NJ; ++mj)
#pragma unroll
    for (int mi = 0; mi < NI; ++mi) {
      dij[mi + NI * mj] = D[(oI + mi) * n + oJ + mj];
      aij[mi + NI * mj] = 0.0;
    }
  for (int r = 0; r < nr; ++r) {
    const int kl = r0 + r, mk = kl % NK, ml = kl / NK;
    const double dkl = D[(oK + mk) * n + oL + ml];
    double djl[NJ], djk[NJ], dil[NI], dik[NI];
#pragma unroll
    for (int mj = 0; mj < NJ; ++mj) {
      djl[mj] = D[(oJ + mj) * n + oL + ml];
      djk[mj] = D[(oJ + mj) * n + oK + mk];
    }
#pragma unroll
    for (int mi = 0; mi < NI; ++mi) {
      dil[mi] = D[(oI + mi) * n + oL + ml];
      dik[mi] = D[(oI + mi) * n + oK + mk];
    }
    double akl = 0.0, bik[NI], bil[NI], bjk[NJ], bjl[NJ];
#pragma unroll
    for (int mi = 0; mi < NI; ++mi) bik[mi] = bil[mi] = 0.0;
#pragma unroll
    for (int mj = 0; mj < NJ; ++mj) bjk[mj] = bjl[mj] = 0.0;
#pragma unroll
    for (int mj = 0; mj < NJ; ++mj)
#pragma unroll
      for (int mi = 0; mi < NI; ++mi) {
        const double v = live ? jk_stage[base + (r * NB + mi + NI * mj) * ls] : 0.0;
        aij[mi + NI * mj] += v * dkl;
        akl += v * dij[mi + NI * mj];
        bik[mi] += v * djl[mj];
        bjk[mj] += v * dil[mi];
        bil[mi] += v * djk[mj];
        bjl[mj] += v * dik[mi];
      }
    jk_put(live, J, (oK + mk) * n + oL + ml, 2.0 * w * akl, rKL);
#pragma unroll
    for (int mi = 0; mi < NI; ++mi) {
      jk_put(live, K, (oI + mi) * n + oK + mk, w * bik[mi], rIK);
      jk_put(live, K, (oI + mi) * n + oL + ml, w * bil[mi], rIL);
    }
#pragma unroll
    for (int mj = 0; mj < NJ; ++mj) {
      jk_put(live, K, (oJ + mj) * n + oK + mk, w * bjk[mj], rJK);
      jk_put(live, K, (oJ + mj) * n + oL + ml, w * bjl[mj], rJL);
    }
  }
#pragma unroll
  for (int mj = 0; mj < NJ; ++mj)
#pragma unroll
    for (int mi = 0; mi < NI; ++mi)
      jk_put(live, J, (oI + mi) * n + oJ + mj, 2.0 * w * aij[mi + NI * mj], rIJ);
}

// jk_digest_rows for the compile-time rows [0, NR) (a kernel with one digest
// site, inlined): the K contributions of all rows are summed in registers
// per output element first, so each K element gets one segmented sum and
// one atomic per tuple instead of one per row that reaches it.
template <int NI, int NJ, int NK, int NL, int NR>
__device__ __forceinline__ void jk_digest_rows_fixed(int base, int ls, bool live, int s0, int s1,
                                                     int s2, int s3, int o0, int o1, int o2,
                                                     int o3, const double* __restrict__ D,
                                                     double* __restrict__ J,
                                                     double* __restrict__ K, int nao) {
  extern __shared__ double jk_stage[];
  constexpr int NB = NI * NJ;
  const double w = jk_weight(s0, s1, s2, s3);
  const size_t n = (size_t)nao;
  const size_t oI = o0, oJ = o1, oK = o2, oL = o3;
  const unsigned long long dead = ~0ull - __lane_id();
  const WaveRun rIJ = wave_run(live ? oI * n + oJ : dead);
  const WaveRun rKL = wave_run(live ? oK * n + oL : dead);
  const WaveRun rIK = wave_run(live ? oI * n + oK : dead);
  const WaveRun rIL = wave_run(live ? oI * n + oL : dead);
  const WaveRun rJK = wave_run(live ? oJ * n + oK : dead);
  const WaveRun rJL = wave_run(live ? oJ * n + oL : dead);
  double dij[NB], aij[NB];
#pragma unroll
  for (int mj = 0; mj < NJ; ++mj)
#pragma unroll
    for (int mi = 0; mi < NI; ++mi) {
      dij[mi + NI * mj] = D[(oI + mi) * n + oJ + mj];
      aij[mi + NI * mj] = 0.0;
    }
  double bik[NK][NI], bil[NL][NI], bjk[NK][NJ], bjl[NL][NJ];
#pragma unroll
  for (int mi = 0; mi < NI; ++mi) {
#pragma unroll
    for (int mk = 0; mk < NK; ++mk) bik[mk][mi] = 0.0;
#pragma unroll
    for (int ml = 0; ml < NL; ++ml) bil[ml][mi] = 0.0;
  }
#pragma unroll
  for (int mj = 0; mj < NJ; ++mj) {
#pragma unroll
    for (int mk = 0; mk < NK; ++mk) bjk[mk][mj] = 0.0;
#pragma unroll
    for (int ml = 0; ml < NL; ++ml) bjl[ml][mj] = 0.0;
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int mk = r % NK, ml = r / NK;
    const double dkl = D[(oK + mk) * n + oL + ml];
    double djl[NJ], djk[NJ], dil[NI], dik[NI];
#pragma unroll
    for (int mj = 0; mj < NJ; ++mj) {
      djl[mj] = D[(oJ + mj) * n + oL + ml];
      djk[mj] = D[(oJ + mj) * n + oK + mk];
    }
#pragma unroll
    for (int mi = 0; mi < NI; ++mi) {
      dil[mi] = D[(oI + mi) * n + oL + ml];
      dik[mi] = D[(oI + mi) * n + oK + mk];
    }
    double akl = 0.0;
#pragma unroll
    for (int mj = 0; mj < NJ; ++mj)
#pragma unroll
      for (int mi = 0; mi < NI; ++mi) {
        const double v = live ? jk_stage[base + (r * NB + mi + NI * mj) * ls] : 0.0;
        aij[mi + NI * mj] += v * dkl;
        akl += v * dij[mi + NI * mj];
        bik[mk][mi] += v * djl[mj];
        bjk[mk][mj] += v * dil[mi];
        bil[ml][mi] += v * djk[mj];
        bjl[ml][mj] += v * dik[mi];
      }
    jk_put(live, J, (oK + mk) * n + oL + ml, 2.0 * w * akl, rKL);
  }
  // the k and l components the rows reached
  constexpr int K1 = NR < NK ? NR : NK, L1 = (NR - 1) / NK + 1;
#pragma unroll
  for (int mk = 0; mk < K1; ++mk) {
#pragma unroll
    for (int mi = 0; mi < NI; ++mi) jk_put(live, K, (oI + mi) * n + oK + mk, w * bik[mk][mi], rIK);
#pragma unroll
    for (int mj = 0; mj < NJ; ++mj) jk_put(live, K, (oJ + mj) * n + oK + mk, w * bjk[mk][mj], rJK);
  }
#pragma unroll
  for (int ml = 0; ml < L1; ++ml) {
#pragma unroll
    for (int mi = 0; mi < NI; ++mi) jk_put(live, K, (oI + mi) * n + oL + ml, w * bil[ml][mi], rIL);
#pragma unroll
    for (int mj = 0; mj < NJ; ++mj) jk_put(live, K, (oJ + mj) * n + oL + ml, w * bjl[ml][mj], rJL);
  }
#pragma unroll
  for (int mj = 0; mj < NJ; ++mj)
#pragma unroll
    for (int mi = 0; mi < NI; ++mi)
      jk_put(live, J, (oI + mi) * n + oJ + mj, 2.0 * w * aij[mi + NI * mj], rIJ);
}

template <int NI, int NJ, int NK, int NL>
__device__ __noinline__ void jk_digest_rows_call(int base, int ls, int r0, int nr, bool live,
                                                 int s0, int s1, int s2, int s3, int o0, int o1,
                                                 int o2, int o3, const double* __restrict__ D,
                                                 double* __restrict__ J,
                                                 double* __restrict__ K, int nao) {
  jk_digest_rows<NI, NJ, NK, NL>(base, ls, r0, nr, live, s0, s1, s2, s3, o0, o1, o2, o3, D, J, K,
                                 nao);
}

}  // namespace cx
