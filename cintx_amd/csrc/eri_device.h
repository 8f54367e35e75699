// Device code of the Rys two-electron class kernels (gfx950 / CDNA4).
//
// One template covers int2e, int2e_ip1, int3c2e and int3c2e_ip1 for a
// compile-time angular-momentum class (LI, LJ, LK, LL):
//   * a workgroup holds Q shell tuples, T lanes per tuple (T | 64, so a tuple
//     never straddles a wavefront);
//   * per primitive tuple: the T lanes evaluate the NR Rys roots from the
//     Chebyshev table (lane r -> root r), then the 3*NR (axis, root) recurrence
//     chains (2D VRR + bra/ket HRR, registers only) and store the 4D g tensor
//     in LDS as g[axis][i][j][k][l][root], then every lane accumulates its
//     Cartesian elements sum_r gx*gy*gz in registers;
//   * after the primitive loops the Cartesian block is moved to LDS, turned
//     into real spherical harmonics axis by axis, and written with caller
//     strides (so a canonicalised tuple lands in the requested orientation).
// Arithmetic follows libcint: per-root coefficients g2e.c:4523-4543, VRR
// g2e.c:272-411, HRR g2e.c:428-690 (ibase/kbase choice g2e.c:102-150), gout
// cint2e.c:961-1180, nabla_i g2e.c:4574-4620, screening cint2e.c:199-236
// (a bra primitive pair with cce >= expcutoff has a zero prefactor in
// libcint's pair data, optimizer.c:317-341, and is skipped here).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "jk_device.h"
#include "launch.h"

namespace cx {

__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

// libcint Cartesian order (cint_bas.c:166-180): lx from l down, then ly from
// l-lx down; within an lx group the position equals lz.
__host__ __device__ constexpr void cart_decode(int l, int c, int& lx, int& ly, int& lz) {
  int start = 0;
  for (int x = l; x >= 0; --x) {
    int cnt = l - x + 1;
    if (c < start + cnt) {
      lx = x;
      lz = c - start;
      ly = l - x - lz;
      return;
    }
    start += cnt;
  }
  lx = ly = lz = 0;
}

// Rys rule for n roots, root r, argument x: s = t^2 = u/(1+u), q = 1 - s and
// the weight (rys_roots.c:57-123 regimes: Chebyshev table below 35+5n,
// asymptotic Hermite rule above).  The table holds, per root and interval,
// the fit of the smaller of s and q; a fitted q is stored negated.
__device__ __forceinline__ void rys_eval_at(const double* __restrict__ cn,
                                            const double* __restrict__ large, int n, int r,
                                            double x, double& s, double& q, double& w) {
  if (!(x < rys_largex(n))) {
    const double* L = large + n * (n - 1);
    s = L[r] / x;
    q = 1.0 - s;
    w = L[n + r] / sqrt(x);
    return;
  }
  int iv = (int)x;
  double t = 2.0 * (x - (double)iv) - 1.0;
  double t2 = t + t;
  const double* cu = cn + ((size_t)iv * 2 * n + r) * kRysNc;
  const double* cw = cu + (size_t)n * kRysNc;
  double bu1 = 0, bu2 = 0, bw1 = 0, bw2 = 0;
#pragma unroll
  for (int k = kRysDeg; k >= 1; --k) {
    double bu0 = fma(t2, bu1, cu[k] - bu2);
    double bw0 = fma(t2, bw1, cw[k] - bw2);
    bu2 = bu1;
    bu1 = bu0;
    bw2 = bw1;
    bw1 = bw0;
  }
  const double v = fma(t, bu1, cu[0] - bu2);
  const double small = fabs(v);
  s = v < 0 ? 1.0 - small : small;
  q = v < 0 ? small : 1.0 - small;
  w = fma(t, bw1, cw[0] - bw2);
}

__device__ __forceinline__ void rys_eval(const RysTable& tab, int n, int r, double x, double& s,
                                         double& q, double& w) {
  rys_eval_at(tab.coef + tab.off[n], tab.large, n, r, x, s, q, w);
}

// Doubles of the table block of n roots (intervals x 2n series x coefficients).
__host__ __device__ constexpr int rys_block_doubles(int n) {
  return (int)rys_largex(n) * 2 * n * kRysNc;
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; remap
// so that XCD x runs a contiguous range of block indices (tuples that are
// adjacent in a bin share bra pairs, whose primitive data then stays in that
// XCD's L2).  A bijection on [0, gridDim.x).
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int x = b & 7, local = b >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + local;
}

// Per-tuple scalars of one primitive tuple.
struct PrimTuple {
  double aij, akl, a0, a1;
  double inv_sum;        // 1/(a_ij + a_kl)
  double hij, hkl;       // 0.5/a_ij, 0.5/a_kl
  double dx[3];   // P - Q
  double prb[3];  // P - R(bra base)
  double qrk[3];  // Q - R(ket base)
  double ai;      // exponent on shell i (nabla_i)
};

// One (axis, root) recurrence chain: 2D VRR then bra and ket HRR; writes
// g[((axis*DI + i)*DJ + j)*DK + k)*DL + l)*RS + r] (root stride RS, NR by
// default; 1 when a lane builds one root at a time in registers).
//
// Register budget: the VRR is streamed over the ket power m (columns m-1, m,
// m+1 live), each finished column goes straight through the bra transfer
// (in place, one extra column) into g at the (i, j) block's first DK*DL
// slots, and the ket transfer then runs block by block in place.  This keeps
// about 4*(NMAX+1) + (MMAX+1) values live instead of the whole 2D and 3D
// intermediates.
// FLIP: the bra transfer starts from the other centre than libcint's rule
// (prb and rij_ax are then the caller's for that centre).
template <int LIc, int LJ, int LK, int LL, typename R, int RS = -1, bool FLIP = false>
__device__ __forceinline__ void build_chain(R* __restrict__ g, int axis, int r, double s,
                                            double q, double wz, double inv_sum, double hij, double hkl,
                                            double aij, double akl, double dx, double prb,
                                            double qrk, double unit, double rij_ax,
                                            double rkl_ax) {
  constexpr int NR = (LIc + LJ + LK + LL) / 2 + 1;
  constexpr int NMAX = LIc + LJ, MMAX = LK + LL;
  constexpr bool IBASE = (LIc > LJ) != FLIP, KBASE = LK > LL;
  constexpr int DI = LIc + 1, DJ = LJ + 1, DK = LK + 1, DL = LL + 1;
  constexpr int A = IBASE ? LIc : LJ, B = IBASE ? LJ : LIc;
  constexpr int C = KBASE ? LK : LL, D = KBASE ? LL : LK;
  constexpr int BLK = DK * DL;  // positions per (i, j) block
  constexpr int NRS = RS < 0 ? NR : RS;

  // libcint's per-root coefficients (g2e.c:4523-4543) written in s = u/(1+u):
  // tmp4 = .5/(a1 (1+u)) = .5 (1-s)/a1 and tmp5 = u2 tmp4 = .5 s/(aij+akl).
  const double sa = s * inv_sum;
  const double oms = q;  // 1 - s with full relative accuracy
  const R b00 = (R)(0.5 * sa);
  const R b10 = (R)(0.5 * sa + oms * hij);
  const R b01 = (R)(0.5 * sa + oms * hkl);
  const R c00 = (R)(prb - sa * akl * dx);
  const R c0p = (R)(qrk + sa * aij * dx);
  const R drb = (R)rij_ax;
  const R drk = (R)rkl_ax;
  R* __restrict__ ga = g + (size_t)axis * DI * DJ * BLK * NRS + r;

  // bra transfer of one finished VRR column m: h1[i][j][m] -> block(i,j)+m
  auto bra = [&](const R* col, int m) {
    R t[NMAX + 1];
#pragma unroll
    for (int a = 0; a <= NMAX; ++a) t[a] = col[a];
#pragma unroll
    for (int b = 0; b <= B; ++b) {
      if (b > 0) {
#pragma unroll
        for (int a = 0; a <= A + B - b; ++a) t[a] = t[a + 1] + drb * t[a];
      }
#pragma unroll
      for (int a = 0; a <= A; ++a) {
        const int i = IBASE ? a : b;
        const int j = IBASE ? b : a;
        ga[((i * DJ + j) * BLK + m) * NRS] = t[a];
      }
    }
  };

  R cm1[NMAX + 1], c0[NMAX + 1], c1[NMAX + 1];
  c0[0] = (R)(unit + (1.0 - unit) * wz);  // 1 for x and y, the weight for z
#pragma unroll
  for (int n = 0; n < NMAX; ++n) {
    R v = c00 * c0[n];
    if (n > 0) v += (R)n * b10 * c0[n - 1];
    c0[n + 1] = v;
  }
  bra(c0, 0);
#pragma unroll
  for (int m = 0; m < MMAX; ++m) {
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) {
      R v = c0p * c0[n];
      if (m > 0) v += (R)m * b01 * cm1[n];
      if (n > 0) v += (R)n * b00 * c0[n - 1];
      c1[n] = v;
    }
    bra(c1, m + 1);
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) {
      cm1[n] = c0[n];
      c0[n] = c1[n];
    }
  }
  // ket transfer, block by block in place
  if constexpr (D > 0 || !KBASE) {
#pragma unroll
    for (int ij = 0; ij < DI * DJ; ++ij) {
      R* blk = ga + (size_t)ij * BLK * NRS;
      R t[MMAX + 1];
#pragma unroll
      for (int c = 0; c <= MMAX; ++c) t[c] = blk[c * NRS];
#pragma unroll
      for (int d = 0; d <= D; ++d) {
        if (d > 0) {
#pragma unroll
          for (int c = 0; c <= C + D - d; ++c) t[c] = t[c + 1] + drk * t[c];
        }
#pragma unroll
        for (int c = 0; c <= C; ++c) {
          const int k = KBASE ? c : d;
          const int l = KBASE ? d : c;
          blk[(k * DL + l) * NRS] = t[c];
        }
      }
    }
  }
}

// Apply the real-spherical transform along one axis of a [l][k][j][i] block
// (i fastest) held in LDS: src dims (n0,n1,n2,n3), axis `ax` goes from
// ncart(lax) to nsph(lax) entries.
__device__ __forceinline__ void c2s_pass(double* __restrict__ dst, const double* __restrict__ src,
                                         const int* dims, int ax, int lax,
                                         const double* __restrict__ c2s, int t, int T) {
  int inner = 1, outer = 1;
  for (int b = 0; b < ax; ++b) inner *= dims[b];
  for (int b = ax + 1; b < 4; ++b) outer *= dims[b];
  const int nc = dims[ax], ns = nsph(lax);
  const int total = outer * ns * inner;
  const double* tab = c2s + (size_t)lax * kC2sM * kC2sC;
  for (int o = t; o < total; o += T) {
    int in = o % inner;
    int rest = o / inner;
    int s = rest % ns;
    int ou = rest / ns;
    const double* row = tab + s * kC2sC;
    const double* sp = src + (size_t)ou * nc * inner + in;
    double acc = 0;
    for (int c = 0; c < nc; ++c) {
      double cf = row[c];
      if (cf != 0.0) acc += cf * sp[(size_t)c * inner];
    }
    dst[o] = acc;
  }
}

// The same pass as small dense products on the matrix cores, for lane
// groups of whole wavefronts (t, T multiples of 64 apart): dst = T_l src with
// T_l the nsph x ncart table, as D(16 x 16) += A(16 x 4) B(4 x 16) steps of
// v_mfma_f64_16x16x4f64 over 16-column tiles of the flattened other axes.
// Operand layout (checked on gfx950): lane L holds A[L % 16][L / 16] and
// B[L / 16][L % 16]; element j of its result is D[L / 16 + 4 j][L % 16].
typedef double c2s_v4d __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void c2s_pass_mfma(double* __restrict__ dst,
                                              const double* __restrict__ src, const int* dims,
                                              int ax, int lax, const double* __restrict__ c2s,
                                              int t, int T) {
  int inner = 1, outer = 1;
  for (int b = 0; b < ax; ++b) inner *= dims[b];
  for (int b = ax + 1; b < 4; ++b) outer *= dims[b];
  const int nc = dims[ax], ns = nsph(lax);
  const int ncol = outer * inner;
  const double* tab = c2s + (size_t)lax * kC2sM * kC2sC;
  const int lane = t & 63, wave = t >> 6, nwave = T >> 6;
  const int col = lane & 15, kq = lane >> 4;
  for (int n0 = wave * 16; n0 < ncol; n0 += nwave * 16) {  // uniform per wavefront
    const int n = n0 + col;
    const int ou = n / inner, in = n - ou * inner;
    c2s_v4d d = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < nc; k0 += 4) {
      const int k = k0 + kq;
      const double av = (col < ns && k < nc) ? tab[col * kC2sC + k] : 0.0;
      double bv = 0.0;
      if (n < ncol && k < nc) bv = src[((size_t)ou * nc + k) * inner + in];
      d = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, d, 0, 0, 0);
    }
    if (n < ncol) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int m = kq + 4 * j;
        if (m < ns) dst[((size_t)ou * ns + m) * inner + in] = d[j];
      }
    }
  }
}

// Synchronisation of the T lanes of one tuple.  A tuple of at most 64 lanes
// lies inside one wavefront (T divides 64), whose LDS operations complete in
// issue order, so a wavefront barrier with LDS fences is enough; wider tuples
// need the workgroup barrier.  Trip counts of the loops that contain these
// calls are uniform over the workgroup.
template <int T>
__device__ __forceinline__ void tuple_sync() {
  if constexpr (T > 64) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

template <int T>
__device__ __forceinline__ bool tuple_any(bool v) {
  if constexpr (T > 64) {
    return __syncthreads_or(v);
  } else {
    tuple_sync<T>();
    return __any(v);
  }
}

__host__ __device__ constexpr int pow2ceil(int v) {
  int p = 1;
  while (p < v) p *= 2;
  return p;
}

// The same transform on a block held in registers (one lane per tuple):
// in[OUTER][ncart(L)][INNER] -> out[OUTER][nsph(L)][INNER], indices
// compile-time once unrolled; the coefficient reads are wave-uniform.
template <int L, int INNER, int OUTER>
__device__ __forceinline__ void c2s_reg(const double* in, double* out,
                                        const double* __restrict__ c2s) {
  constexpr int NC = ncart(L), NSL = nsph(L);
  if constexpr (L < 2) {
#pragma unroll
    for (int i = 0; i < OUTER * NC * INNER; ++i) out[i] = in[i];
  } else {
    const double* tab = c2s + (size_t)L * kC2sM * kC2sC;
#pragma unroll
    for (int o = 0; o < OUTER; ++o)
#pragma unroll
      for (int m = 0; m < NSL; ++m)
#pragma unroll
        for (int i = 0; i < INNER; ++i) {
          double v = 0;
#pragma unroll
          for (int c = 0; c < NC; ++c) v += tab[m * kC2sC + c] * in[(o * NC + c) * INNER + i];
          out[(o * NSL + m) * INNER + i] = v;
        }
  }
}

// One real-solid-harmonic coefficient (s and p: identity).
__device__ __forceinline__ double c2s_coef(const double* __restrict__ c2s, int l, int m, int c) {
  if (l < 2) return m == c ? 1.0 : 0.0;
  return c2s[((size_t)l * kC2sM + m) * kC2sC + c];
}

// Lanes per tuple (TT), primitive tuples per step (PP) and tuples per
// workgroup of one class-kernel variant.  The candidate (TT, PP) pairs of a
// class are generated by the build and chosen per plan bin by timing them
// (cintx_plan_autotune); the first candidate is the untuned default.
//
// Gout blocking of the wide (T > 1) path: a lane owns a row of NB elements
// that share their (k, l) components (all (i, j) components; nabla_i: all i
// components of one (j, k, l)), so per root it loads the (LI+1)(LJ+1) g
// values of each axis once and reuses them across the row instead of loading
// three values per element.  With more lanes than rows, G lane groups split
// the P primitive tuples of a step and their sums are added at the end.
template <int LI, int LJ, int LK, int LL, int DERIV, int TT, int PP>
struct ClassShape {
  static constexpr int LIc = LI + DERIV;
  static constexpr int NR = (LIc + LJ + LK + LL) / 2 + 1;
  static constexpr int DI = LIc + 1, DJ = LJ + 1, DK = LK + 1, DL = LL + 1;
  static constexpr int G4 = DI * DJ * DK * DL;
  static constexpr int GS = 3 * G4 * NR;
  static constexpr int NF = ncart(LI) * ncart(LJ) * ncart(LK) * ncart(LL);
  static constexpr int NCOMP = DERIV ? 3 : 1;
  static constexpr int NE = NF * NCOMP;
  // IROW (PP bit 8, LDS path only): rows of the i components alone, as for
  // nabla_i, so more lanes share a large class
  static constexpr bool IROW = !DERIV && TT > 1 && ((PP >> 8) & 1);
  static constexpr int NB = (DERIV || IROW) ? ncart(LI) : ncart(LI) * ncart(LJ);
  static constexpr int NROW = NF / NB;
  static constexpr int T = TT;
  static constexpr int P = TT == 1 ? 1 : (PP & 255);
  // one-lane path: the rows of the block are done in NCH chunks (PP), each a
  // full pass over the primitive tuples, so that a chunk's accumulators and
  // the g tensor of one root fit in registers
  static constexpr int NCH = TT == 1 ? (PP & 255) : 1;
  // PP bits 9-10: minimum waves per SIMD the compiler must fit (register
  // budget 512 / MINW per lane), 1 by default
  static constexpr int MINW = ((PP >> 9) & 3) + 1;
  // PP bit 11: Cartesian output (the _cart families), no cart -> sph
  static constexpr bool CART = (PP >> 11) & 1;
  // PP bit 12: fused J/K epilogue (int2e_sph only): blocks are digested into
  // J and K instead of written
  static constexpr bool JK = (PP >> 12) & 1;
  static_assert(!JK || (!DERIV && !CART), "fused J/K serves int2e_sph only");
  // PP bit 13 (wide path only): the bra transfer starts, per primitive
  // tuple, from the centre of the larger exponent, as the runtime-class
  // kernel does (for f and g shells paired with a shell of l >= 2, where
  // libcint's fixed base loses up to 1e-10 in FP64)
  static constexpr bool PB = (PP >> 13) & 1;
  static_assert(!PB || (TT > 1 && !JK), "per-primitive transfer base: wide path, no J/K");
  static constexpr int nout(int l) { return CART ? ncart(l) : nsph(l); }
  static constexpr int RPC = (NROW + NCH - 1) / NCH;  // rows per chunk
  static constexpr int G = T >= NROW ? T / NROW : 1;               // groups per row
  static constexpr int RPL = T >= NROW ? 1 : (NROW + T - 1) / T;   // rows per lane
  static constexpr int NFT = (NF + T - 1) / T;                     // T == 1 path
  static constexpr int NSC = 16;  // per-primitive scalars shared in LDS
  // one-lane path, fused J/K: (k, l) output pairs staged per pass for the
  // 256 lanes of a workgroup within 128 KB of dynamic LDS (the static LDS
  // of these kernels, the Rys table block, stays below 30 KB)
  // (never more than the (k, l) pairs a chunk reaches, so that small
  // classes keep the occupancy of their plain kernels)
  static constexpr int jk_rows_per_pass() {
    const int rpp = 64 / (nout(LI) * nout(LJ));
    const int reach = (LK < 2 && LL < 2) ? RPC : nout(LK) * nout(LL);
    return rpp < 1 ? 1 : (rpp < reach ? rpp : reach);
  }
  static constexpr size_t jk_lds_bytes() {
    return (size_t)jk_rows_per_pass() * nout(LI) * nout(LJ) * 256 * sizeof(double);
  }
  // the Cartesian/spherical staging buffers of the epilogue alias the g slots
  template <typename R>
  static constexpr int union_doubles() {  // per tuple: g slots | staging
    return cmax((P * GS * (int)sizeof(R) + 7) / 8, 2 * NF);
  }
  template <typename R>
  static constexpr int lds_per_tuple() {
    return T == 1 ? 0 : (union_doubles<R>() + P * (3 * NR + NSC)) * 8;
  }
  // at most 256 lanes per workgroup and about 32 KB of LDS (finer occupancy
  // steps on the 160 KB of a CU), one tuple at least
  template <typename R>
  static constexpr int Q() {
    int q = 256 / T;
    if (q < 1) q = 1;
    while (q > 1 && q * lds_per_tuple<R>() > 32 * 1024) q /= 2;
    return q;
  }
  static_assert(NF % NB == 0, "row blocking");
  static_assert(T == 1 || T >= 4, "lanes per tuple");
};

template <int LI, int LJ, int LK, int LL, int DERIV, typename R, int TT, int PP>
__global__ void __launch_bounds__((ClassShape<LI, LJ, LK, LL, DERIV, TT, PP>::T *
                                   ClassShape<LI, LJ, LK, LL, DERIV, TT, PP>::template Q<R>()),
                                  (ClassShape<LI, LJ, LK, LL, DERIV, TT, PP>::MINW))
    eri_class_kernel(EriArgs a) {
  using S = ClassShape<LI, LJ, LK, LL, DERIV, TT, PP>;
  static_assert(S::template Q<R>() * S::template lds_per_tuple<R>() <= 64 * 1024, "LDS");
  constexpr int T = S::T;
  constexpr int Q = S::template Q<R>();
  constexpr int LIc = S::LIc;
  constexpr int NR = S::NR;
  constexpr int DI = S::DI, DJ = S::DJ, DK = S::DK, DL = S::DL;
  constexpr int G4 = S::G4;
  constexpr int NFI = ncart(LI), NFJ = ncart(LJ), NFK = ncart(LK), NFL = ncart(LL);
  constexpr int NF = S::NF;
  constexpr int NCOMP = S::NCOMP;
  constexpr int NFT = S::NFT;
  constexpr bool IBASE = LIc > LJ, KBASE = LK > LL;
  constexpr int GS = S::GS;
  constexpr int NB = S::NB, NROW = S::NROW, G = S::G, RPL = S::RPL;
  constexpr bool IROWS = DERIV || S::IROW;  // rows are (j, k, l) components

  // One lane per tuple keeps g and the roots in registers (every index is a
  // compile-time constant once the loops are unrolled); wider tuples share
  // them through LDS.
  constexpr bool REG = (T == 1);
  constexpr int UD = REG ? 0 : S::template union_doubles<R>();
  constexpr int LDS_D = REG ? 1 : Q * (UD + S::P * (3 * NR + S::NSC));
  __shared__ double lds[LDS_D];
  constexpr int ACC = REG ? S::RPC * NB * NCOMP : RPL * NB * NCOMP;
  double acc[ACC];
#pragma unroll
  for (int kk = 0; kk < ACC; ++kk) acc[kk] = 0.0;

  const int slot = threadIdx.x / T;
  const int t = threadIdx.x - slot * T;
  const int64_t q = (int64_t)xcd_block() * Q + slot;
  const bool active = q < a.nq;
  const int2 pq = a.quartets[active ? q : 0];
  const int boff = a.pair_off[pq.x];
  const int koff = a.pair_off[pq.y];
  const double* gb = a.pair_geo + 8 * (size_t)pq.x;
  const double* gk = a.pair_geo + 8 * (size_t)pq.y;
  const double ri[3] = {gb[0], gb[1], gb[2]};
  const double rj[3] = {gb[3], gb[4], gb[5]};
  const double rk[3] = {gk[0], gk[1], gk[2]};
  const double rl[3] = {gk[3], gk[4], gk[5]};
  const double logd_ij = gb[6], logd_kl = gk[6];
  // coordinate differences: every position the recurrences need is formed
  // from these and the primitive weights (P - R_i = w_j R_ji), never from
  // an absolute product centre
  double rij[3], rkl[3], rji[3], rlk[3], rik[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    rji[d] = rj[d] - ri[d];
    rlk[d] = rl[d] - rk[d];
    rik[d] = ri[d] - rk[d];
    rij[d] = IBASE ? -rji[d] : rji[d];
    rkl[d] = KBASE ? -rlk[d] : rlk[d];
  }
  unsigned int nsurv = 0;

  // Primitive counts are per tuple; a workgroup of several tuples loops to
  // its largest counts (tuples are sorted by (npij, npkl) so that is rarely
  // more) and masks the excess iterations.
  const int np_ij = a.pair_np[pq.x], np_kl = a.pair_np[pq.y];
  int mij = np_ij, mkl = np_kl;
  if constexpr (T > 1) {
    __shared__ int smax[2];
    if (threadIdx.x == 0) smax[0] = smax[1] = 0;
    __syncthreads();
    if (t == 0 && active) {
      atomicMax(&smax[0], np_ij);
      atomicMax(&smax[1], np_kl);
    }
    __syncthreads();
    mij = smax[0];
    mkl = smax[1];
  }
  double* b0 = nullptr;  // wide path: epilogue staging (Cartesian in, spherical out)
  double* b1 = nullptr;
  if constexpr (REG) {
  // One lane per tuple, everything in registers: per primitive tuple and
  // root, the three axis chains (3*G4 values) and then every element of the
  // current chunk of rows; the chunk's block is transformed and written (or
  // added to the output, for chunks after the first) before the next chunk.
  (void)lds;
  constexpr bool F64 = sizeof(R) == sizeof(double);
  constexpr int NCH = S::NCH, RPC = S::RPC, CE = RPC * NB;  // elements per chunk
  // up to 3 roots the table block of this n (<= 29 KB) is copied to LDS once
  // per workgroup: the coefficient reads of every root then hit LDS instead
  // of scattered L2 lines
  constexpr bool LTAB = NR <= 3;
  __shared__ double stab[LTAB ? rys_block_doubles(NR) : 1];
  const double* coefn = a.rys.coef + a.rys.off[NR];
  if constexpr (LTAB) {
    for (int i = threadIdx.x; i < rys_block_doubles(NR); i += blockDim.x) stab[i] = coefn[i];
    __syncthreads();
    coefn = stab;
  }
  R g[3 * G4];
  R part[F64 ? 1 : ACC];  // FP32: this primitive tuple's sums over roots
  auto tgt = [&](int i) -> R& {
    if constexpr (F64)
      return acc[i];
    else
      return part[i];
  };
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
  if (ch > 0) {
#pragma unroll
    for (int kk = 0; kk < ACC; ++kk) acc[kk] = 0.0;
  }
  // Each lane walks only its own surviving primitive tuples.  The primitive
  // pairs of every pair are stored in ascending cce, so the bra primitives
  // that pass the screen against ket primitive kp are a prefix of the bra
  // list, and a ket primitive whose first bra primitive fails ends the
  // tuple (later ket primitives have larger cce).  A wavefront then runs
  // as many iterations as its lane with the most survivors (tuples are
  // sorted by survivor count), not the product of the primitive counts.
  // The bra record of the next candidate is loaded one iteration ahead (its
  // latency then hides behind the current primitive tuple's work).
  (void)mij;
  (void)mkl;
  int kp = 0, ip = 0;
  double akl = 0, wk = 0, wl = 0, kkl = 0, hkl = 0, eijcut = 0;
  double qk[3] = {0, 0, 0}, qb[3] = {0, 0, 0};  // Q - R_k, Q - R(ket base)
  bool klok = false;
  auto set_ket = [&](int k) {
    const double* pk = a.prim + 8 * (size_t)(koff + k);
    akl = pk[0];
    wk = pk[1];
    wl = pk[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      qk[d] = wl * rlk[d];
      qb[d] = KBASE ? qk[d] : -wk * rlk[d];
    }
    kkl = pk[4];
    hkl = pk[5];
    const double ccekl = a.three_center ? 0.0 : pk[6] - a.lkl_ceil * logd_kl;
    klok = ccekl <= a.expcutoff;
    eijcut = a.expcutoff - ccekl;
  };
  double nb[8];  // record of bra candidate ip (prefetched)
  auto load_bra = [&](int i) {
    const double* q = a.prim + 8 * (size_t)(boff + i);
#pragma unroll
    for (int d = 0; d < 8; ++d) nb[d] = q[d];
  };
  if (active && np_kl > 0) {
    set_ket(0);
    load_bra(0);
  }
  for (;;) {
    bool have = false;
    while (active && kp < np_kl) {
      if (klok && ip < np_ij) {
        const double cceij = nb[6] - a.lij_ceil * logd_ij;
        // (libcint zeroes the prefactor of a bra pair with cce >= expcutoff,
        // optimizer.c:317-341, so it contributes nothing)
        if (cceij <= eijcut && cceij < a.expcutoff) {
          have = true;
          break;
        }
      }
      if (ip == 0) {  // nothing survives from this ket primitive on
        kp = np_kl;
        break;
      }
      ++kp;
      ip = 0;
      if (kp < np_kl) {
        set_ket(kp);
        load_bra(0);
      }
    }
    if (!have) break;
    double pi[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) pi[d] = nb[d];
    ++ip;
    if (ip < np_ij) load_bra(ip);
    {
      PrimTuple p;
      p.aij = pi[0];
      p.akl = akl;
      const double wi = pi[1], wj = pi[2];
      double rr = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double pr = wj * rji[d];  // P - R_i
        p.dx[d] = (rik[d] + pr) - qk[d];
        rr += p.dx[d] * p.dx[d];
        p.prb[d] = IBASE ? pr : -wi * rji[d];
        p.qrk[d] = qb[d];
      }
      p.inv_sum = 1.0 / (p.aij + akl);
      p.hij = pi[5];
      p.hkl = hkl;
      p.a1 = p.aij * akl;
      p.a0 = p.a1 * p.inv_sum;
      // sqrt(a0/a1^3) = 1/(a1 sqrt(aij+akl)) = 4 hij hkl sqrt(inv_sum)
      const double fac = a.common * pi[4] * kkl;
      const double fac1 = 4.0 * p.hij * hkl * sqrt(p.inv_sum) * fac;
      const double x = p.a0 * rr;
      const R ai2 = DERIV ? (R)(-2.0 * pi[7]) : (R)0;
      if (ch == 0) ++nsurv;
      if constexpr (!F64) {
#pragma unroll
        for (int kk = 0; kk < ACC; ++kk) part[kk] = 0;
      }
      // small blocks: all roots first (their Clenshaw recurrences interleave),
      // then one root at a time through the chains and the elements; larger
      // blocks evaluate each root in turn to keep registers for g
      constexpr bool ROOTS_FIRST = 3 * G4 * (int)sizeof(R) / 8 + ACC <= 64;
      double rs[ROOTS_FIRST ? NR : 1], rq[ROOTS_FIRST ? NR : 1], rw[ROOTS_FIRST ? NR : 1];
      if constexpr (ROOTS_FIRST) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
          rys_eval_at(coefn, a.rys.large, NR, r, x, rs[r], rq[r], rw[r]);
      }
#pragma unroll 1
      for (int r = 0; r < NR; ++r) {
        double sr, qr, wr;
        if constexpr (ROOTS_FIRST) {
          sr = rs[r];
          qr = rq[r];
          wr = rw[r];
        } else {
          rys_eval_at(coefn, a.rys.large, NR, r, x, sr, qr, wr);
        }
        const double wz = wr * fac1;
#pragma unroll
        for (int axis = 0; axis < 3; ++axis)
          build_chain<LIc, LJ, LK, LL, R, 1>(g, axis, 0, sr, qr, wz, p.inv_sum, p.hij, p.hkl,
                                             p.aij, p.akl, p.dx[axis], p.prb[axis], p.qrk[axis],
                                             axis == 2 ? 0.0 : 1.0, rij[axis], rkl[axis]);
        // the chunk's elements; all g indices are compile-time constants
#pragma unroll
        for (int el = 0; el < CE; ++el) {
          const int row = ch * RPC + el / NB;
          if (row >= NROW) continue;
          const int e = row * NB + el % NB;
          int ix, iy, iz, jx, jy, jz, kx, ky, kz, lx, ly, lz;
          cart_decode(LI, e % NFI, ix, iy, iz);
          cart_decode(LJ, (e / NFI) % NFJ, jx, jy, jz);
          cart_decode(LK, (e / (NFI * NFJ)) % NFK, kx, ky, kz);
          cart_decode(LL, e / (NFI * NFJ * NFK), lx, ly, lz);
          const int px = ((ix * DJ + jx) * DK + kx) * DL + lx;
          const int py = ((iy * DJ + jy) * DK + ky) * DL + ly;
          const int pz = ((iz * DJ + jz) * DK + kz) * DL + lz;
          const R* gx = g;
          const R* gy = g + G4;
          const R* gz = g + 2 * G4;
          if constexpr (!DERIV) {
            tgt(el) += gx[px] * gy[py] * gz[pz];
          } else {
            constexpr int SI = DJ * DK * DL;  // stride of the i power
            R fx = ai2 * gx[px + SI];
            R fy = ai2 * gy[py + SI];
            R fz = ai2 * gz[pz + SI];
            if (ix > 0) fx += (R)ix * gx[px - SI];
            if (iy > 0) fy += (R)iy * gy[py - SI];
            if (iz > 0) fz += (R)iz * gz[pz - SI];
            tgt(el) += fx * gy[py] * gz[pz];
            tgt(CE + el) += gx[px] * fy * gz[pz];
            tgt(2 * CE + el) += gx[px] * gy[py] * fz;
          }
        }
      }
      if constexpr (!F64) {
#pragma unroll
        for (int kk = 0; kk < ACC; ++kk) acc[kk] += (double)part[kk];
      }
    }
  }
  // cart -> sph of the chunk, in registers: the row's own axes (i, j; i only
  // for nabla_i rows) per row, then the row axes as a weighted sum over the
  // chunk's rows, added to what earlier chunks wrote (or, for a fused J/K,
  // staged in LDS and digested)
  {
    constexpr bool CART = S::CART;
    constexpr int NSI = S::nout(LI), NSJ = S::nout(LJ), NSK = S::nout(LK), NSL = S::nout(LL);
    constexpr int NSB = DERIV ? NSI : NSI * NSJ;  // output size of a row
#pragma unroll
    for (int comp = 0; comp < NCOMP; ++comp) {
      double d[RPC * NSB];
#pragma unroll
      for (int rl = 0; rl < RPC; ++rl) {
        const double* src = acc + comp * CE + rl * NB;
        if constexpr (CART) {
#pragma unroll
          for (int e = 0; e < NSB; ++e) d[rl * NSB + e] = src[e];
        } else if constexpr (DERIV) {
          c2s_reg<LI, 1, 1>(src, d + rl * NSB, a.c2s);
        } else {
          double t1[NSI * NFJ];
          c2s_reg<LI, 1, NFJ>(src, t1, a.c2s);
          c2s_reg<LJ, NSI, 1>(t1, d + rl * NSB, a.c2s);
        }
      }
      // The chunk's share of output (ml, mk, mj, mb).  s and p row axes map
      // rows to outputs one to one: an output only takes rows whose
      // components match there, so a chunk reaches just some outputs (all
      // compile-time): `mine` when any of its rows reach the output,
      // `earlier` when a row of an earlier chunk did.
      auto out_value = [&](int ml, int mk, int mj, int mb, bool& mine, bool& earlier) -> double {
        auto reaches = [&](int row) {
          int fj = 0, fk, fl;
          if constexpr (DERIV) {
            fj = row % NFJ;
            fk = (row / NFJ) % NFK;
            fl = row / (NFJ * NFK);
          } else {
            fk = row % NFK;
            fl = row / NFK;
          }
          return ((!CART && LK >= 2) || fk == mk) && ((!CART && LL >= 2) || fl == ml) &&
                 (!DERIV || (!CART && LJ >= 2) || fj == mj);
        };
        mine = false;
        earlier = false;
#pragma unroll
        for (int row = 0; row < NROW; ++row) {
          if (!reaches(row)) continue;
          if (row / RPC == ch) mine = true;
          if (row / RPC < ch) earlier = true;
        }
        double v = 0;
        if (!mine) return v;
#pragma unroll
        for (int rl = 0; rl < RPC; ++rl) {
          const int row = ch * RPC + rl;
          if (row >= NROW) continue;
          int fj = 0, fk, fl;
          if constexpr (DERIV) {
            fj = row % NFJ;
            fk = (row / NFJ) % NFK;
            fl = row / (NFJ * NFK);
          } else {
            fk = row % NFK;
            fl = row / NFK;
          }
          double cf;
          if constexpr (CART) {
            cf = (fk == mk && fl == ml && (!DERIV || fj == mj)) ? 1.0 : 0.0;
          } else {
            cf = c2s_coef(a.c2s, LK, mk, fk) * c2s_coef(a.c2s, LL, ml, fl);
            if constexpr (DERIV) cf *= c2s_coef(a.c2s, LJ, mj, fj);
          }
          v += cf * d[rl * NSB + mb];
        }
        return v;
      };
      if constexpr (S::JK) {
        {
          // fused J/K: this chunk's outputs are staged in LDS (lane
          // interleaved, a range of (k, l) output pairs per pass) and each
          // lane digests its own rows (the reductions are linear, so the
          // chunks' shares and the passes digest separately); the block
          // never reaches global memory.  With s and p row axes a chunk
          // reaches the (k, l) pairs of its own rows only.
          constexpr int NKL = NSK * NSL;
          constexpr bool ONE_TO_ONE = LK < 2 && LL < 2;
          const int KL0 = ONE_TO_ONE ? ch * RPC : 0;
          const int KL1 = ONE_TO_ONE ? cmin(ch * RPC + RPC, NKL) : NKL;
          constexpr int RPP = S::jk_rows_per_pass();
          // one digest site: one chunk, one pass (the loops are unrolled),
          // over the rows [0, ONE_NR)
          constexpr bool ONE_SITE = NCH == 1 && (ONE_TO_ONE ? RPC : NKL) <= RPP;
          constexpr int ONE_NR = ONE_TO_ONE ? (RPC < NKL ? RPC : NKL) : NKL;
          extern __shared__ double jks[];
          const int sa = a.pair_sh[pq.x].x, sbb = a.pair_sh[pq.x].y;
          const int sc = a.pair_sh[pq.y].x, sd = a.pair_sh[pq.y].y;
          const int sh[4] = {sa, sbb, sc, sd};
          const int off[4] = {a.ao_off[sa], a.ao_off[sbb], a.ao_off[sc], a.ao_off[sd]};
          const int lane = threadIdx.x;
#pragma unroll
          for (int r0 = KL0; r0 < KL1; r0 += RPP) {
            const int nr = cmin(RPP, KL1 - r0);
#pragma unroll
            for (int r = 0; r < nr; ++r) {
              const int mk = (r0 + r) % NSK, ml = (r0 + r) / NSK;
#pragma unroll
              for (int mb = 0; mb < NSB; ++mb) {
                bool mine, earlier;
                const double v = out_value(ml, mk, 0, mb, mine, earlier);
                jks[(r * NSB + mb) * 256 + lane] = mine ? v : 0.0;
              }
            }
            __syncthreads();
            if constexpr (ONE_SITE)  // rows [0, ONE_NR), a compile-time range
              jk_digest_rows_fixed<NSI, NSJ, NSK, NSL, ONE_NR>(
                  lane, 256, active, sh[0], sh[1], sh[2], sh[3], off[0], off[1], off[2], off[3],
                  a.jk_D, a.jk_J, a.jk_K, a.nao);
            else
              jk_digest_rows_call<NSI, NSJ, NSK, NSL>(lane, 256, r0, nr, active, sh[0], sh[1],
                                                      sh[2], sh[3], off[0], off[1], off[2], off[3],
                                                      a.jk_D, a.jk_J, a.jk_K, a.nao);
            __syncthreads();
          }
          continue;
        }
      }
      if (!active) continue;
      double* o = a.out + a.out_base + q * a.block + comp * a.st_comp;
#pragma unroll
      for (int ml = 0; ml < NSL; ++ml)
#pragma unroll
        for (int mk = 0; mk < NSK; ++mk)
#pragma unroll
          for (int mj = 0; mj < (DERIV ? NSJ : 1); ++mj)
#pragma unroll
            for (int mb = 0; mb < NSB; ++mb) {
              bool mine, earlier;
              const double v = out_value(ml, mk, mj, mb, mine, earlier);
              if (!mine) continue;
              const int mi = DERIV ? mb : mb % NSI;
              const int mjj = DERIV ? mj : mb / NSI;
              double* dst = o + mi * a.st[0] + mjj * a.st[1] + mk * a.st[2] + ml * a.st[3];
              *dst = earlier ? *dst + v : v;
            }
    }
  }
  }
  } else {
    // Wide tuples: P primitive tuples per step.  Phase 1 (P*NR root slots):
    // screening, per-primitive scalars and one Rys root each; phase 2
    // (P*3*NR chain slots): one recurrence chain each; phase 3: every lane
    // accumulates its rows over its groups' primitive tuples.  Slots beyond
    // T lanes are strided (T < 3*NR happens for small tuples).
    constexpr int P = S::P;
    constexpr int NSC = S::NSC;
    constexpr int BLK = DK * DL;
    constexpr int SI = DJ * DK * DL;
    constexpr bool F64 = sizeof(R) == sizeof(double);
    double* const ubase = lds + slot * UD;
    R* const sg = reinterpret_cast<R*>(ubase);  // [P][GS]
    double* const suw = lds + Q * UD + slot * P * (3 * NR + NSC);  // [P][3NR]
    double* const sps = suw + P * 3 * NR;                          // [P][NSC]
    b0 = ubase;
    b1 = ubase + NF;

    // this lane's rows: packed base positions x | y << 10 | z << 20
    const int grp = T >= NROW ? t / NROW : 0;
    int rowid[RPL];
    int rb[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int row = T >= NROW ? t - grp * NROW : t + T * k;
      rowid[k] = (grp < G && row < NROW) ? row : -1;
      const int rr = rowid[k] < 0 ? 0 : row;
      int bx, by, bz;
      if constexpr (!IROWS) {
        int kx, ky, kz, lx, ly, lz;
        cart_decode(LK, rr % NFK, kx, ky, kz);
        cart_decode(LL, rr / NFK, lx, ly, lz);
        bx = kx * DL + lx;
        by = ky * DL + ly;
        bz = kz * DL + lz;
      } else {
        int jx, jy, jz, kx, ky, kz, lx, ly, lz;
        const int fj = rr % NFJ, r1 = rr / NFJ;
        cart_decode(LJ, fj, jx, jy, jz);
        cart_decode(LK, r1 % NFK, kx, ky, kz);
        cart_decode(LL, r1 / NFK, lx, ly, lz);
        bx = (jx * DK + kx) * DL + lx;
        by = (jy * DK + ky) * DL + ly;
        bz = (jz * DK + kz) * DL + lz;
      }
      rb[k] = bx | (by << 10) | (bz << 20);
    }
    // this step's partial sums (FP32 path) or the accumulators themselves
    R svf[F64 ? 1 : ACC];
    auto tgt = [&](int i) -> R& {
      if constexpr (F64)
        return acc[i];
      else
        return svf[i];
    };

    const int npq = mij * mkl;
    for (int base = 0; base < npq; base += P) {
      bool any = false;
      for (int ur = t; ur < P * NR; ur += T) {
        const int pp = ur / NR, r = ur - pp * NR;
        const int k = base + pp;
        const int kp = k / mij, ip = k - kp * mij;
        const bool kin = k < npq && kp < np_kl && ip < np_ij;
        const int kidx = koff + (kin ? kp : 0), iidx = boff + (kin ? ip : 0);
        const double* pk = a.prim + 8 * (size_t)kidx;
        const double* pi = a.prim + 8 * (size_t)iidx;
        const double ccekl = a.three_center ? 0.0 : pk[6] - a.lkl_ceil * logd_kl;
        const double cceij = pi[6] - a.lij_ceil * logd_ij;
        const bool ok = active && kin && ccekl <= a.expcutoff &&
                        cceij <= a.expcutoff - ccekl && cceij < a.expcutoff;
        double* uwp = suw + pp * 3 * NR;
        double* sc = sps + pp * NSC;
        bool flip = false;  // PB: the bra base of this primitive tuple is not libcint's
        if (ok) {
          const double aij = pi[0], akl = pk[0];
          // bra transfer base: libcint's rule, or (PB) the centre of the
          // larger exponent (a_i = pi[7], a_j = aij - a_i)
          const bool ib = S::PB ? (pi[7] >= aij - pi[7]) : IBASE;
          flip = ib != IBASE;
          double dx[3], pb[3], qb[3], rr = 0;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double pr = pi[2] * rji[d], qr = pk[2] * rlk[d];  // P - R_i, Q - R_k
            dx[d] = (rik[d] + pr) - qr;
            rr += dx[d] * dx[d];
            pb[d] = ib ? pr : -pi[1] * rji[d];
            qb[d] = KBASE ? qr : -pk[1] * rlk[d];
          }
          const double inv_sum = 1.0 / (aij + akl);
          const double hij = pi[5], hkl = pk[5];
          const double fac = a.common * pi[4] * pk[4];
          const double fac1 = 4.0 * hij * hkl * sqrt(inv_sum) * fac;
          const double x = aij * akl * inv_sum * rr;
          double sr, qr, w;
          rys_eval(a.rys, NR, r, x, sr, qr, w);
          uwp[r] = sr;
          uwp[NR + r] = w * fac1;
          uwp[2 * NR + r] = qr;
          if (r == 0) {
            sc[0] = inv_sum;
            sc[1] = hij;
            sc[2] = hkl;
            sc[3] = aij;
            sc[4] = akl;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              sc[5 + d] = dx[d];
              sc[8 + d] = pb[d];
              sc[11 + d] = qb[d];
            }
            sc[14] = pi[7];
            ++nsurv;
          }
        }
        // 0: screened out, 1: survives, 2: survives with the other bra base
        if (r == 0) sc[15] = ok ? (flip ? 2.0 : 1.0) : 0.0;
        any = any || ok;
      }
      if (!tuple_any<T>(any)) continue;
      // which of the P slots hold a surviving primitive tuple (read before
      // the next barrier; phase 1 of the next step rewrites the flags)
      unsigned okmask = 0;
#pragma unroll
      for (int pp = 0; pp < P; ++pp)
        if (sps[pp * NSC + 15] != 0.0) okmask |= 1u << pp;
      for (int uc = t; uc < P * 3 * NR; uc += T) {
        const int pp = uc / (3 * NR), c = uc - pp * 3 * NR;
        if (!((okmask >> pp) & 1u)) continue;
        const int axis = c / NR, r = c - axis * NR;
        const double* sc = sps + pp * NSC;
        const double* uwp = suw + pp * 3 * NR;
        // the whole chain (VRR and both transfers) runs in registers; LDS
        // only receives the G4 finished values, so no transfer step waits
        // on an LDS round trip
        R loc[G4];
        const double rij_a = axis == 0 ? rij[0] : (axis == 1 ? rij[1] : rij[2]);
        const double rkl_a = axis == 0 ? rkl[0] : (axis == 1 ? rkl[1] : rkl[2]);
        if (S::PB && sc[15] == 2.0)  // the other bra base: the transfer step changes sign
          build_chain<LIc, LJ, LK, LL, R, 1, S::PB>(
              loc, 0, 0, uwp[r], uwp[2 * NR + r], uwp[NR + r], sc[0], sc[1], sc[2], sc[3], sc[4],
              sc[5 + axis], sc[8 + axis], sc[11 + axis], axis == 2 ? 0.0 : 1.0, -rij_a, rkl_a);
        else
          build_chain<LIc, LJ, LK, LL, R, 1>(loc, 0, 0, uwp[r], uwp[2 * NR + r], uwp[NR + r],
                                             sc[0], sc[1], sc[2], sc[3], sc[4], sc[5 + axis],
                                             sc[8 + axis], sc[11 + axis], axis == 2 ? 0.0 : 1.0,
                                             rij_a, rkl_a);
        R* gp = sg + pp * GS + (size_t)axis * G4 * NR + r;
#pragma unroll
        for (int k = 0; k < G4; ++k) gp[k * NR] = loc[k];
      }
      tuple_sync<T>();
      if constexpr (!F64) {
#pragma unroll
        for (int kk = 0; kk < ACC; ++kk) svf[kk] = 0;
      }
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        if (rowid[k] < 0) continue;
        const int bx = rb[k] & 1023, by = (rb[k] >> 10) & 1023, bz = rb[k] >> 20;
        // not unrolled: unrolling lets the compiler hoist the LDS loads of
        // every slot and costs occupancy
#pragma unroll 1
        for (int pp = grp; pp < P; pp += G) {
          if (!((okmask >> pp) & 1u)) continue;
          const R* gpp = sg + pp * GS;
          if constexpr (S::IROW) {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              R xv[DI], yv[DI], zv[DI];
#pragma unroll
              for (int i = 0; i < DI; ++i) {
                xv[i] = gpp[(i * SI + bx) * NR + r];
                yv[i] = gpp[(G4 + i * SI + by) * NR + r];
                zv[i] = gpp[(2 * G4 + i * SI + bz) * NR + r];
              }
#pragma unroll
              for (int ib = 0; ib < NB; ++ib) {
                int ix, iy, iz;
                cart_decode(LI, ib, ix, iy, iz);
                tgt(k * NB + ib) += xv[ix] * yv[iy] * zv[iz];
              }
            }
          } else if constexpr (!DERIV) {
            constexpr int NIJ = DI * DJ;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              R xv[NIJ], yv[NIJ], zv[NIJ];
#pragma unroll
              for (int ij = 0; ij < NIJ; ++ij) {
                xv[ij] = gpp[(ij * BLK + bx) * NR + r];
                yv[ij] = gpp[(G4 + ij * BLK + by) * NR + r];
                zv[ij] = gpp[(2 * G4 + ij * BLK + bz) * NR + r];
              }
#pragma unroll
              for (int ib = 0; ib < NB; ++ib) {
                int ix, iy, iz, jx, jy, jz;
                cart_decode(LI, ib % NFI, ix, iy, iz);
                cart_decode(LJ, ib / NFI, jx, jy, jz);
                tgt(k * NB + ib) += xv[ix * DJ + jx] * yv[iy * DJ + jy] * zv[iz * DJ + jz];
              }
            }
          } else {
            const R ai2 = (R)(-2.0 * sps[pp * NSC + 14]);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              R xv[DI], yv[DI], zv[DI];
#pragma unroll
              for (int i = 0; i < DI; ++i) {
                xv[i] = gpp[(i * SI + bx) * NR + r];
                yv[i] = gpp[(G4 + i * SI + by) * NR + r];
                zv[i] = gpp[(2 * G4 + i * SI + bz) * NR + r];
              }
#pragma unroll
              for (int ib = 0; ib < NB; ++ib) {
                int ix, iy, iz;
                cart_decode(LI, ib, ix, iy, iz);
                R fx = ai2 * xv[ix + 1];
                R fy = ai2 * yv[iy + 1];
                R fz = ai2 * zv[iz + 1];
                if (ix > 0) fx += (R)ix * xv[ix - 1];
                if (iy > 0) fy += (R)iy * yv[iy - 1];
                if (iz > 0) fz += (R)iz * zv[iz - 1];
                tgt((k * 3 + 0) * NB + ib) += fx * yv[iy] * zv[iz];
                tgt((k * 3 + 1) * NB + ib) += xv[ix] * fy * zv[iz];
                tgt((k * 3 + 2) * NB + ib) += xv[ix] * yv[iy] * fz;
              }
            }
          }
        }
      }
      if constexpr (!F64) {
#pragma unroll
        for (int kk = 0; kk < ACC; ++kk) acc[kk] += (double)svf[kk];
      }
      // a tuple wider than a wavefront: phase 1 of the next step rewrites the
      // slot scalars that phase 3 reads (nabla_i)
      if constexpr (DERIV && T > 64) __syncthreads();
    }
    (void)rowid;
  }
  if (a.counters) {
    // work counters (surviving primitive tuples, tuples with any): the wide
    // path counts in the root-0 lane of every primitive slot, so add the
    // lanes of a tuple first
    unsigned int tot = nsurv;
    if constexpr (T > 1) {
      __shared__ unsigned int scnt[Q];
      if (t == 0) scnt[slot] = 0;
      tuple_sync<T>();
      if (nsurv) atomicAdd(&scnt[slot], nsurv);
      tuple_sync<T>();
      tot = scnt[slot];
    }
    if (t == 0 && tot) {
      atomicAdd(a.counters, (unsigned long long)tot);
      atomicAdd(a.counters + 1, 1ull);
    }
  }

  // Cartesian -> spherical (unless Cartesian output), then the strided write.
  constexpr int NSI = S::nout(LI), NSJ = S::nout(LJ), NSK = S::nout(LK), NSL = S::nout(LL);
  constexpr int NS = NSI * NSJ * NSK * NSL;
  if constexpr (REG) {
    // written chunk by chunk above
    } else {
  tuple_sync<T>();
  for (int comp = 0; comp < NCOMP; ++comp) {
    // lane groups add their partial rows in turn
    {
      const int grp = T >= NROW ? t / NROW : 0;
#pragma unroll 1
      for (int gg = 0; gg < G; ++gg) {
        if (grp == gg) {
#pragma unroll
          for (int k = 0; k < RPL; ++k) {
            const int row = T >= NROW ? t - grp * NROW : t + T * k;
            if (row < NROW) {
#pragma unroll
              for (int ib = 0; ib < NB; ++ib) {
                const double v = acc[(k * NCOMP + comp) * NB + ib];
                double* d = b0 + row * NB + ib;
                *d = gg == 0 ? v : *d + v;
              }
            }
          }
        }
        tuple_sync<T>();
      }
    }
    int dims[4] = {NFI, NFJ, NFK, NFL};
    const int ls[4] = {LI, LJ, LK, LL};
    double* src = b0;
    double* dst = b1;
#pragma unroll
    for (int ax = 0; ax < 4; ++ax) {
      if (ls[ax] < 2 || S::CART) continue;  // s and p: identity
      if constexpr (T % 64 == 0)
        c2s_pass_mfma(dst, src, dims, ax, ls[ax], a.c2s, t, T);
      else
        c2s_pass(dst, src, dims, ax, ls[ax], a.c2s, t, T);
      dims[ax] = nsph(ls[ax]);
      double* tmp = src;
      src = dst;
      dst = tmp;
      tuple_sync<T>();
    }
    bool fused = false;
    if constexpr (S::JK) {
      if (active) {  // fused J/K: digest the staged block
        const int2 sb = a.pair_sh[pq.x], sq = a.pair_sh[pq.y];
        const int sh[4] = {sb.x, sb.y, sq.x, sq.y};
        const int ns[4] = {NSI, NSJ, NSK, NSL};
        const int off[4] = {a.ao_off[sb.x], a.ao_off[sb.y], a.ao_off[sq.x], a.ao_off[sq.y]};
        jk_digest_block(src, ns, sh, off, t, T, a.jk_D, a.jk_J, a.jk_K, a.nao);
      }
      fused = true;
    }
    if (active && !fused) {
      double* o = a.out + a.out_base + q * a.block + comp * a.st_comp;
      for (int e = t; e < NS; e += T) {
        int mi = e % NSI, e1 = e / NSI;
        int mj = e1 % NSJ, e2 = e1 / NSJ;
        int mk = e2 % NSK, ml = e2 / NSK;
        o[mi * a.st[0] + mj * a.st[1] + mk * a.st[2] + ml * a.st[3]] = src[e];
      }
    }
    tuple_sync<T>();
  }
  }
}

}  // namespace cx

namespace cx {

template <int LI, int LJ, int LK, int LL, int DERIV, typename R, int TT, int PP>
hipError_t launch_class(const EriArgs& a, hipStream_t s) {
  using S = ClassShape<LI, LJ, LK, LL, DERIV, TT, PP>;
  constexpr int Q = S::template Q<R>();
  const int64_t nblk = (a.nq + Q - 1) / Q;
  // the one-lane kernels stage their blocks for a fused J/K in dynamic LDS
  size_t dyn = 0;
  if constexpr (S::T == 1 && S::JK) {
    static_assert(S::jk_lds_bytes() <= 128 * 1024, "fused J/K staging");
    dyn = S::jk_lds_bytes();
  }
  hipLaunchKernelGGL((eri_class_kernel<LI, LJ, LK, LL, DERIV, R, TT, PP>), dim3((unsigned)nblk),
                     dim3(S::T * Q), dyn, s, a);
  return hipGetLastError();
}

template <int LI, int LJ, int LK, int LL, int DERIV, typename R, int TT, int PP>
KernelInfo class_info() {
  using S = ClassShape<LI, LJ, LK, LL, DERIV, TT, PP>;
  return KernelInfo{&launch_class<LI, LJ, LK, LL, DERIV, R, TT, PP>, S::template Q<R>(), S::T,
                    S::T == 1 ? S::NCH : S::P};
}

}  // namespace cx
