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
// cint2e.c:961-1180, nabla_i g2e.c:4574-4620, screening cint2e.c:199-236.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
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
__device__ __forceinline__ void rys_eval(const RysTable& tab, int n, int r, double x, double& s,
                                         double& q, double& w) {
  if (!(x < rys_largex(n))) {
    const double* L = tab.large + n * (n - 1);
    s = L[r] / x;
    q = 1.0 - s;
    w = L[n + r] / sqrt(x);
    return;
  }
  int iv = (int)x;
  double t = 2.0 * (x - (double)iv) - 1.0;
  double t2 = t + t;
  const double* cu = tab.coef + tab.off[n] + ((size_t)iv * 2 * n + r) * kRysNc;
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
// g[((axis*DI + i)*DJ + j)*DK + k)*DL + l)*NR + r].
//
// Register budget: the VRR is streamed over the ket power m (columns m-1, m,
// m+1 live), each finished column goes straight through the bra transfer
// (in place, one extra column) into g at the (i, j) block's first DK*DL
// slots, and the ket transfer then runs block by block in place.  This keeps
// about 4*(NMAX+1) + (MMAX+1) values live instead of the whole 2D and 3D
// intermediates.
template <int LIc, int LJ, int LK, int LL, typename R>
__device__ __forceinline__ void build_chain(R* __restrict__ g, int axis, int r, double s,
                                            double q, double wz, double inv_sum, double hij, double hkl,
                                            double aij, double akl, double dx, double prb,
                                            double qrk, double unit, double rij_ax,
                                            double rkl_ax) {
  constexpr int NR = (LIc + LJ + LK + LL) / 2 + 1;
  constexpr int NMAX = LIc + LJ, MMAX = LK + LL;
  constexpr bool IBASE = LIc > LJ, KBASE = LK > LL;
  constexpr int DI = LIc + 1, DJ = LJ + 1, DK = LK + 1, DL = LL + 1;
  constexpr int A = IBASE ? LIc : LJ, B = IBASE ? LJ : LIc;
  constexpr int C = KBASE ? LK : LL, D = KBASE ? LL : LK;
  constexpr int BLK = DK * DL;  // positions per (i, j) block

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
  R* __restrict__ ga = g + (size_t)axis * DI * DJ * BLK * NR + r;

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
        ga[((i * DJ + j) * BLK + m) * NR] = t[a];
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
      R* blk = ga + (size_t)ij * BLK * NR;
      R t[MMAX + 1];
#pragma unroll
      for (int c = 0; c <= MMAX; ++c) t[c] = blk[c * NR];
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
          blk[(k * DL + l) * NR] = t[c];
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

// Lanes per tuple and tuples per workgroup for a class.
template <int LI, int LJ, int LK, int LL, int DERIV>
struct ClassShape {
  static constexpr int LIc = LI + DERIV;
  static constexpr int NR = (LIc + LJ + LK + LL) / 2 + 1;
  static constexpr int DI = LIc + 1, DJ = LJ + 1, DK = LK + 1, DL = LL + 1;
  static constexpr int G4 = DI * DJ * DK * DL;
  static constexpr int NF = ncart(LI) * ncart(LJ) * ncart(LK) * ncart(LL);
  static constexpr int NCOMP = DERIV ? 3 : 1;
  static constexpr int NE = NF * NCOMP;
  static constexpr int T = NE <= 8 ? 1 : (NE <= 48 ? 4 : (NE <= 192 ? 16 : (NE <= 768 ? 64 : 256)));
  static constexpr int NFT = (NF + T - 1) / T;
  // primitive tuples in flight per step (wide tuples): enough to keep the
  // lanes of the root and recurrence phases busy
  static constexpr int NSC = 16;  // per-primitive scalars shared in LDS
  // (LDS per primitive slot is sized for FP64 so both precisions share P)
  static constexpr int PLDS = (60 * 1024 - 2 * NF * 8) / (3 * G4 * NR * 8 + (3 * NR + NSC) * 8);
  static constexpr int P = T == 1 ? 1 : cmax(1, cmin(cmin(8, T / (3 * NR)), PLDS));
  template <typename R>
  static constexpr int lds_per_tuple() {
    return T == 1 ? (int)(2 * NF * sizeof(double))
                  : (int)(P * (3 * G4 * NR * sizeof(R) + (3 * NR + NSC) * sizeof(double)) +
                          2 * NF * sizeof(double));
  }
  template <typename R>
  static constexpr int Q() {
    int q = 256 / T;
    while (q > 1 && q * lds_per_tuple<R>() > 60 * 1024) q /= 2;
    return q;
  }
};

template <int LI, int LJ, int LK, int LL, int DERIV, typename R>
__global__ void __launch_bounds__((ClassShape<LI, LJ, LK, LL, DERIV>::T *
                                   ClassShape<LI, LJ, LK, LL, DERIV>::template Q<R>()))
    eri_class_kernel(EriArgs a) {
  using S = ClassShape<LI, LJ, LK, LL, DERIV>;
  constexpr int T = S::T;
  constexpr int Q = S::template Q<R>();
  constexpr int LIc = S::LIc;
  constexpr int NR = S::NR;
  constexpr int DI = S::DI, DJ = S::DJ, DK = S::DK, DL = S::DL;
  constexpr int G4 = S::G4;
  constexpr int NFI = ncart(LI), NFJ = ncart(LJ), NFK = ncart(LK), NFL = ncart(LL);
  constexpr int NF = S::NF;
  constexpr int NCOMP = S::NCOMP;
  constexpr int NE = S::NE;
  constexpr int NFT = S::NFT;
  constexpr bool IBASE = LIc > LJ, KBASE = LK > LL;
  constexpr int GS = 3 * G4 * NR;

  // One lane per tuple keeps g and the roots in registers (every index is a
  // compile-time constant once the loops are unrolled); wider tuples share
  // them through LDS.
  constexpr bool REG = (T == 1);
  __shared__ double sbuf[Q][2][NF];
  R greg[REG ? GS : 1];
  double uwreg[REG ? 3 * NR : 1];

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
  double rij[3], rkl[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    rij[d] = IBASE ? ri[d] - rj[d] : rj[d] - ri[d];
    rkl[d] = KBASE ? rk[d] - rl[d] : rl[d] - rk[d];
  }

  // packed LDS positions of this lane's elements: x | y << 10 | z << 20
  int pos[NFT];
#pragma unroll
  for (int kk = 0; kk < NFT; ++kk) {
    int e = t + kk * T;
    if (e >= NF) e = 0;
    int fi = e % NFI, e1 = e / NFI;
    int fj = e1 % NFJ, e2 = e1 / NFJ;
    int fk = e2 % NFK, fl = e2 / NFK;
    int ix, iy, iz, jx, jy, jz, kx, ky, kz, lx, ly, lz;
    cart_decode(LI, fi, ix, iy, iz);
    cart_decode(LJ, fj, jx, jy, jz);
    cart_decode(LK, fk, kx, ky, kz);
    cart_decode(LL, fl, lx, ly, lz);
    int px = ((ix * DJ + jx) * DK + kx) * DL + lx;
    int py = ((iy * DJ + jy) * DK + ky) * DL + ly;
    int pz = ((iz * DJ + jz) * DK + kz) * DL + lz;
    pos[kk] = px | (py << 10) | (pz << 20);
  }

  double acc[NFT * NCOMP];
#pragma unroll
  for (int kk = 0; kk < NFT * NCOMP; ++kk) acc[kk] = 0.0;
  unsigned int nsurv = 0;

  // Primitive counts are per tuple; a workgroup of several tuples loops to
  // its largest counts (tuples are sorted by (npij, npkl) so that is rarely
  // more) and masks the excess iterations.
  const int np_ij = a.pair_np[pq.x], np_kl = a.pair_np[pq.y];
  int mij = np_ij, mkl = np_kl;
  if constexpr (T > 1 || Q > 1) {
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
  }
  if constexpr (REG) {
  R* __restrict__ g = greg;
  double* __restrict__ uw = uwreg;

  for (int kp = 0; kp < mkl; ++kp) {
    const bool kin = kp < np_kl;
    const int kidx = koff + (kin ? kp : 0);
    const double akl = a.aij[kidx];
    const double Qv[3] = {a.px[kidx], a.py[kidx], a.pz[kidx]};
    const double kkl = a.kfac[kidx];
    const double hkl = a.hinv[kidx];
    const double ccekl = a.three_center ? 0.0 : a.cce[kidx] - a.lkl_ceil * logd_kl;
    const bool klok = active && kin && ccekl <= a.expcutoff;
    const double eijcut = a.expcutoff - ccekl;
    for (int ip = 0; ip < mij; ++ip) {
      const bool iin = ip < np_ij;
      const int iidx = boff + (iin ? ip : 0);
      const double cceij = a.cce[iidx] - a.lij_ceil * logd_ij;
      const bool ok = klok && iin && cceij <= eijcut;
      if constexpr (T > 1) {
        if (!__syncthreads_or(ok)) continue;
      } else {
        if (!ok) continue;
      }
      PrimTuple p;
      p.aij = a.aij[iidx];
      p.akl = akl;
      const double Pv[3] = {a.px[iidx], a.py[iidx], a.pz[iidx]};
      double rr = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        p.dx[d] = Pv[d] - Qv[d];
        rr += p.dx[d] * p.dx[d];
        p.prb[d] = Pv[d] - (IBASE ? ri[d] : rj[d]);
        p.qrk[d] = Qv[d] - (KBASE ? rk[d] : rl[d]);
      }
      p.inv_sum = 1.0 / (p.aij + akl);
      p.hij = a.hinv[iidx];
      p.hkl = hkl;
      p.a1 = p.aij * akl;
      p.a0 = p.a1 * p.inv_sum;
      // sqrt(a0/a1^3) = 1/(a1 sqrt(aij+akl)) = 4 hij hkl sqrt(inv_sum)
      const double fac = ok ? a.common * a.kfac[iidx] * kkl : 0.0;
      const double fac1 = 4.0 * p.hij * hkl * sqrt(p.inv_sum) * fac;
      const double x = p.a0 * rr;
      if (ok && t == 0) ++nsurv;

      // roots
#pragma unroll(T == 1 ? NR : 1)
      for (int r = t; r < NR; r += T) {
        double sr, qr, w;
        rys_eval(a.rys, NR, r, x, sr, qr, w);
        uw[r] = sr;
        uw[NR + r] = w * fac1;
        uw[2 * NR + r] = qr;
      }
      if constexpr (T > 1) __syncthreads();
      // recurrence chains
#pragma unroll(T == 1 ? 3 * NR : 1)
      for (int c = t; c < 3 * NR; c += T) {
        const int axis = c / NR;
        const int r = c - axis * NR;
        // the axis selects compile-time members, so no private array is
        // indexed at run time (that would put it in scratch memory)
        if (axis == 0)
          build_chain<LIc, LJ, LK, LL, R>(g, 0, r, uw[r], uw[2 * NR + r], uw[NR + r], p.inv_sum,
                                          p.hij, p.hkl,
                                          p.aij, p.akl, p.dx[0], p.prb[0], p.qrk[0], 1.0,
                                          rij[0], rkl[0]);
        else if (axis == 1)
          build_chain<LIc, LJ, LK, LL, R>(g, 1, r, uw[r], uw[2 * NR + r], uw[NR + r], p.inv_sum,
                                          p.hij, p.hkl,
                                          p.aij, p.akl, p.dx[1], p.prb[1], p.qrk[1], 1.0,
                                          rij[1], rkl[1]);
        else
          build_chain<LIc, LJ, LK, LL, R>(g, 2, r, uw[r], uw[2 * NR + r], uw[NR + r], p.inv_sum,
                                          p.hij, p.hkl,
                                          p.aij, p.akl, p.dx[2], p.prb[2], p.qrk[2], 0.0,
                                          rij[2], rkl[2]);
      }
      if constexpr (T > 1) __syncthreads();
      // Cartesian elements
      if constexpr (!DERIV) {
#pragma unroll
        for (int kk = 0; kk < NFT; ++kk) {
          if (kk * T + t < NF) {
            const R* gx = g + (pos[kk] & 1023) * NR;
            const R* gy = g + (G4 + ((pos[kk] >> 10) & 1023)) * NR;
            const R* gz = g + (2 * G4 + (pos[kk] >> 20)) * NR;
            R s = 0;
#pragma unroll
            for (int r = 0; r < NR; ++r) s += gx[r] * gy[r] * gz[r];
            acc[kk] += (double)s;
          }
        }
      } else {
        const R ai2 = (R)(-2.0 * a.ai[iidx]);
        constexpr int SI = DJ * DK * DL;  // stride of the i power
#pragma unroll
        for (int kk = 0; kk < NFT; ++kk) {
          const int e = kk * T + t;
          if (e < NF) {
            const int px = pos[kk] & 1023, py = (pos[kk] >> 10) & 1023, pz = pos[kk] >> 20;
            const int fi = e % NFI;
            int ix, iy, iz;
            cart_decode(LI, fi, ix, iy, iz);
            const R* gx = g + px * NR;
            const R* gy = g + (G4 + py) * NR;
            const R* gz = g + (2 * G4 + pz) * NR;
            R sx = 0, sy = 0, sz = 0;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              R fx = ai2 * gx[SI * NR + r];
              R fy = ai2 * gy[SI * NR + r];
              R fz = ai2 * gz[SI * NR + r];
              if (ix > 0) fx += (R)ix * gx[r - SI * NR];
              if (iy > 0) fy += (R)iy * gy[r - SI * NR];
              if (iz > 0) fz += (R)iz * gz[r - SI * NR];
              sx += fx * gy[r] * gz[r];
              sy += gx[r] * fy * gz[r];
              sz += gx[r] * gy[r] * fz;
            }
            acc[kk] += (double)sx;
            acc[NFT + kk] += (double)sy;
            acc[2 * NFT + kk] += (double)sz;
          }
        }
      }
    }
  }
  } else {
    // Wide tuples: P primitive tuples per step.  Phase 1 (lanes < P*NR):
    // screening, per-primitive scalars and one Rys root each; phase 2 (lanes
    // < P*3*NR): one recurrence chain each; phase 3: every lane accumulates
    // its elements over the P*NR quadrature points.
    constexpr int P = S::P;
    constexpr int NSC = S::NSC;
    __shared__ R sgb[Q][P][GS];
    __shared__ double suwb[Q][P][3 * NR];
    __shared__ double sps[Q][P][NSC];
    const int npq = mij * mkl;
    for (int base = 0; base < npq; base += P) {
      bool any = false;
      // strided: P*NR and P*3*NR can exceed T when T < 3*NR (P is then 1)
      for (int ur = t; ur < P * NR; ur += T) {
        const int pp = ur / NR, r = ur - pp * NR;
        const int k = base + pp;
        const int kp = k / mij, ip = k - kp * mij;
        const bool kin = k < npq && kp < np_kl && ip < np_ij;
        const int kidx = koff + (kin ? kp : 0), iidx = boff + (kin ? ip : 0);
        const double ccekl = a.three_center ? 0.0 : a.cce[kidx] - a.lkl_ceil * logd_kl;
        const double cceij = a.cce[iidx] - a.lij_ceil * logd_ij;
        const bool ok = active && kin && ccekl <= a.expcutoff && cceij <= a.expcutoff - ccekl;
        const double aij = a.aij[iidx], akl = a.aij[kidx];
        const double Pv[3] = {a.px[iidx], a.py[iidx], a.pz[iidx]};
        const double Qv[3] = {a.px[kidx], a.py[kidx], a.pz[kidx]};
        double dx[3], rr = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          dx[d] = Pv[d] - Qv[d];
          rr += dx[d] * dx[d];
        }
        const double inv_sum = 1.0 / (aij + akl);
        const double hij = a.hinv[iidx], hkl = a.hinv[kidx];
        const double fac = ok ? a.common * a.kfac[iidx] * a.kfac[kidx] : 0.0;
        const double fac1 = 4.0 * hij * hkl * sqrt(inv_sum) * fac;
        const double x = aij * akl * inv_sum * rr;
        double sr, qr, w;
        rys_eval(a.rys, NR, r, x, sr, qr, w);
        suwb[slot][pp][r] = sr;
        suwb[slot][pp][NR + r] = w * fac1;
        suwb[slot][pp][2 * NR + r] = qr;
        if (r == 0) {
          double* sc = sps[slot][pp];
          sc[0] = inv_sum;
          sc[1] = hij;
          sc[2] = hkl;
          sc[3] = aij;
          sc[4] = akl;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            sc[5 + d] = dx[d];
            sc[8 + d] = Pv[d] - (IBASE ? ri[d] : rj[d]);
            sc[11 + d] = Qv[d] - (KBASE ? rk[d] : rl[d]);
          }
          sc[14] = a.ai[iidx];
          if (ok) ++nsurv;
        }
        any = any || ok;
      }
      if (!__syncthreads_or(any)) continue;
      for (int uc = t; uc < P * 3 * NR; uc += T) {
        const int pp = uc / (3 * NR), c = uc - pp * 3 * NR;
        const int axis = c / NR, r = c - axis * NR;
        const double* sc = sps[slot][pp];
        const double* uwp = suwb[slot][pp];
        R* gp = sgb[slot][pp];
        if (axis == 0)
          build_chain<LIc, LJ, LK, LL, R>(gp, 0, r, uwp[r], uwp[2 * NR + r], uwp[NR + r], sc[0],
                                          sc[1], sc[2], sc[3], sc[4], sc[5], sc[8], sc[11], 1.0,
                                          rij[0], rkl[0]);
        else if (axis == 1)
          build_chain<LIc, LJ, LK, LL, R>(gp, 1, r, uwp[r], uwp[2 * NR + r], uwp[NR + r], sc[0],
                                          sc[1], sc[2], sc[3], sc[4], sc[6], sc[9], sc[12], 1.0,
                                          rij[1], rkl[1]);
        else
          build_chain<LIc, LJ, LK, LL, R>(gp, 2, r, uwp[r], uwp[2 * NR + r], uwp[NR + r], sc[0],
                                          sc[1], sc[2], sc[3], sc[4], sc[7], sc[10], sc[13], 0.0,
                                          rij[2], rkl[2]);
      }
      __syncthreads();
      if constexpr (!DERIV) {
#pragma unroll
        for (int kk = 0; kk < NFT; ++kk) {
          if (kk * T + t < NF) {
            R sum = 0;
#pragma unroll 1
            for (int pp = 0; pp < P; ++pp) {
              const R* gpp = sgb[slot][pp];
              const R* gx = gpp + (pos[kk] & 1023) * NR;
              const R* gy = gpp + (G4 + ((pos[kk] >> 10) & 1023)) * NR;
              const R* gz = gpp + (2 * G4 + (pos[kk] >> 20)) * NR;
#pragma unroll
              for (int r = 0; r < NR; ++r) sum += gx[r] * gy[r] * gz[r];
            }
            acc[kk] += (double)sum;
          }
        }
      } else {
        constexpr int SI = DJ * DK * DL;
#pragma unroll
        for (int kk = 0; kk < NFT; ++kk) {
          const int e = kk * T + t;
          if (e < NF) {
            const int px = pos[kk] & 1023, py = (pos[kk] >> 10) & 1023, pz = pos[kk] >> 20;
            int ix, iy, iz;
            cart_decode(LI, e % NFI, ix, iy, iz);
            R sx = 0, sy = 0, sz = 0;
#pragma unroll 1
            for (int pp = 0; pp < P; ++pp) {
              const R* gpp = sgb[slot][pp];
              const R ai2 = (R)(-2.0 * sps[slot][pp][14]);
              const R* gx = gpp + px * NR;
              const R* gy = gpp + (G4 + py) * NR;
              const R* gz = gpp + (2 * G4 + pz) * NR;
#pragma unroll
              for (int r = 0; r < NR; ++r) {
                R fx = ai2 * gx[SI * NR + r];
                R fy = ai2 * gy[SI * NR + r];
                R fz = ai2 * gz[SI * NR + r];
                if (ix > 0) fx += (R)ix * gx[r - SI * NR];
                if (iy > 0) fy += (R)iy * gy[r - SI * NR];
                if (iz > 0) fz += (R)iz * gz[r - SI * NR];
                sx += fx * gy[r] * gz[r];
                sy += gx[r] * fy * gz[r];
                sz += gx[r] * gy[r] * fz;
              }
            }
            acc[kk] += (double)sx;
            acc[NFT + kk] += (double)sy;
            acc[2 * NFT + kk] += (double)sz;
          }
        }
        __syncthreads();  // phase 1 of the next step rewrites a_i in sps
      }
    }
  }
  if (a.counters && t == 0 && nsurv) {
    atomicAdd(a.counters, (unsigned long long)nsurv);
    atomicAdd(a.counters + 1, 1ull);
  }

  // Cartesian -> spherical, then the strided write.
  if constexpr (T > 1) __syncthreads();
  double* b0 = sbuf[slot][0];
  double* b1 = sbuf[slot][1];
  constexpr int NSI = nsph(LI), NSJ = nsph(LJ), NSK = nsph(LK), NSL = nsph(LL);
  constexpr int NS = NSI * NSJ * NSK * NSL;
  for (int comp = 0; comp < NCOMP; ++comp) {
#pragma unroll
    for (int kk = 0; kk < NFT; ++kk) {
      const int e = kk * T + t;
      if (e < NF) b0[e] = acc[comp * NFT + kk];
    }
    if constexpr (T > 1) __syncthreads();
    int dims[4] = {NFI, NFJ, NFK, NFL};
    const int ls[4] = {LI, LJ, LK, LL};
    double* src = b0;
    double* dst = b1;
#pragma unroll
    for (int ax = 0; ax < 4; ++ax) {
      if (ls[ax] < 2) continue;  // s and p: identity
      c2s_pass(dst, src, dims, ax, ls[ax], a.c2s, t, T);
      dims[ax] = nsph(ls[ax]);
      double* tmp = src;
      src = dst;
      dst = tmp;
      if constexpr (T > 1) __syncthreads();
    }
    if (active) {
      double* o = a.out + a.out_base + q * a.block + comp * a.st_comp;
      for (int e = t; e < NS; e += T) {
        int mi = e % NSI, e1 = e / NSI;
        int mj = e1 % NSJ, e2 = e1 / NSJ;
        int mk = e2 % NSK, ml = e2 / NSK;
        o[mi * a.st[0] + mj * a.st[1] + mk * a.st[2] + ml * a.st[3]] = src[e];
      }
    }
    if constexpr (T > 1) __syncthreads();
  }
}

}  // namespace cx

namespace cx {

template <int LI, int LJ, int LK, int LL, int DERIV, typename R>
hipError_t launch_class(const EriArgs& a, hipStream_t s) {
  using S = ClassShape<LI, LJ, LK, LL, DERIV>;
  constexpr int Q = S::template Q<R>();
  const int64_t nblk = (a.nq + Q - 1) / Q;
  hipLaunchKernelGGL((eri_class_kernel<LI, LJ, LK, LL, DERIV, R>), dim3((unsigned)nblk),
                     dim3(S::T * Q), 0, s, a);
  return hipGetLastError();
}

template <int LI, int LJ, int LK, int LL, int DERIV, typename R>
KernelInfo class_info() {
  using S = ClassShape<LI, LJ, LK, LL, DERIV>;
  return KernelInfo{&launch_class<LI, LJ, LK, LL, DERIV, R>, S::template Q<R>(), S::T};
}

}  // namespace cx
