// Runtime-class fallback kernel: any (li, lj, lk, ll) with l <= 4 (+1 for
// nabla_i), used for the classes that have no compile-time instantiation
// (high angular momentum: (ff|..), (gg|gg), the s..g basis workload).  One
// shell tuple per 256-lane workgroup, grid-stride over tuples.  The g tensor
// lives in LDS when it fits in 150 KB, else in a per-workgroup global scratch
// slice; the Cartesian accumulators live in registers (8 elements per lane,
// larger blocks in chunks); the chunk results and the
// cart->sph buffers are in the slice.  Arithmetic is the same as
// eri_class_kernel (libcint g2e.c:272-690, 4425-4620).
#include "eri_device.h"

namespace cx {

struct GenShape {
  int l[4];
  int lic;  // l[0] + deriv
  int deriv, nr;
  int di, dj, dk, dl, g4;
  int nf, ncomp;
  int ibase, kbase;
};

// The recurrences of one primitive tuple in three lane-parallel phases.
// generic_vrr, one (axis, root) per lane: the VRR columns m <= MMAX, column m
// staged in slot m of the first NMAX + 1 (i, j) blocks (entry n in block n;
// (A+1)(B+1) >= A+B+1 blocks exist).  generic_bra, one (axis, root, m) per
// lane: reads its staged column, then the bra transfer into g(i, j, slot m).
// generic_ket, one (axis, root, (i, j) block) per lane: the ket transfer of
// that block from the slots m to (k, l).  The work arrays are indexed with
// compile-time positions only (loops unrolled to the largest size, guarded by
// the runtime bounds) so they stay in registers.
constexpr int kChainMax = 10;  // NMAX, MMAX <= 9 (nabla_i on (gg|..))
__device__ void generic_vrr(double* __restrict__ g, const GenShape& S, int axis, int r, double sr,
                            double qr, double wz, double inv_sum, double hij, double hkl,
                            double aij, double akl, double dx, double prb, double qrk) {
  const int NR = S.nr;
  const int NMAX = S.lic + S.l[1], MMAX = S.l[2] + S.l[3];
  const int BLK = S.dk * S.dl;
  const double sa = sr * inv_sum, oms = qr;
  const double b00 = 0.5 * sa, b10 = 0.5 * sa + oms * hij, b01 = 0.5 * sa + oms * hkl;
  const double c00 = prb - sa * akl * dx;
  const double c0p = qrk + sa * aij * dx;
  double* ga = g + (size_t)axis * S.g4 * NR + r;
  double cm1[kChainMax], c0[kChainMax];
#pragma unroll
  for (int n = 0; n < kChainMax; ++n) {
    cm1[n] = 0.0;
    c0[n] = 0.0;
  }
  c0[0] = (axis == 2) ? wz : 1.0;
#pragma unroll
  for (int n = 0; n + 1 < kChainMax; ++n)
    if (n < NMAX) c0[n + 1] = c00 * c0[n] + (n > 0 ? n * b10 * c0[n - 1] : 0.0);
  for (int m = 0; m <= MMAX; ++m) {
    if (m > 0) {
      // descending n: c0[n - 1] is still the previous column when read
#pragma unroll
      for (int n = kChainMax - 1; n >= 0; --n) {
        double v = c0p * c0[n] + (m - 1) * b01 * cm1[n];
        if (n > 0) v += n * b00 * c0[n - 1];
        cm1[n] = c0[n];
        c0[n] = v;
      }
    }
#pragma unroll
    for (int n = 0; n < kChainMax; ++n)
      if (n <= NMAX) ga[((size_t)n * BLK + m) * NR] = c0[n];
  }
}

__device__ void generic_bra(double* __restrict__ g, const GenShape& S, int axis, int r, int m,
                            double drb, bool ibase) {
  const int NR = S.nr;
  const int NMAX = S.lic + S.l[1];
  const int A = ibase ? S.lic : S.l[1], B = ibase ? S.l[1] : S.lic;
  const int DJ = S.dj, BLK = S.dk * S.dl;
  double* ga = g + (size_t)axis * S.g4 * NR + r + (size_t)m * NR;
  double t[kChainMax];
#pragma unroll
  for (int a = 0; a < kChainMax; ++a) t[a] = a <= NMAX ? ga[(size_t)a * BLK * NR] : 0.0;
#pragma unroll
  for (int b = 0; b < kChainMax; ++b) {
    if (b > B) break;
    if (b > 0) {
#pragma unroll
      for (int a = 0; a + 1 < kChainMax; ++a)
        if (a <= A + B - b) t[a] = t[a + 1] + drb * t[a];
    }
#pragma unroll
    for (int a = 0; a < kChainMax; ++a) {
      if (a > A) break;
      const int i = ibase ? a : b, j = ibase ? b : a;
      ga[(size_t)(i * DJ + j) * BLK * NR] = t[a];
    }
  }
}

__device__ void generic_ket(double* __restrict__ g, const GenShape& S, int axis, int r, int ij,
                            double drk, bool kbase) {
  const int NR = S.nr;
  const int MMAX = S.l[2] + S.l[3];
  const int C = kbase ? S.l[2] : S.l[3], D = kbase ? S.l[3] : S.l[2];
  const int DL = S.dl, BLK = S.dk * S.dl;
  double* blk = g + (size_t)axis * S.g4 * NR + r + (size_t)ij * BLK * NR;
  double t[kChainMax];
#pragma unroll
  for (int c = 0; c < kChainMax; ++c) t[c] = c <= MMAX ? blk[c * NR] : 0.0;
#pragma unroll
  for (int d = 0; d < kChainMax; ++d) {
    if (d > D) break;
    if (d > 0) {
#pragma unroll
      for (int c = 0; c + 1 < kChainMax; ++c)
        if (c <= C + D - d) t[c] = t[c + 1] + drk * t[c];
    }
#pragma unroll
    for (int c = 0; c < kChainMax; ++c) {
      if (c > C) break;
      const int k = kbase ? c : d, l = kbase ? d : c;
      blk[(k * DL + l) * NR] = t[c];
    }
  }
}

// GLDS: the g tensor in (dynamic) LDS instead of the global slice.
// Row-blocked gout: a lane owns RPL rows of the block, a row being the
// ncart(LI) elements that share their j, k, l components (LI, the canonical
// i shell's l, is a template parameter so that the row's elements index
// registers).  Per root the lane loads the LI + 1 (nabla_i: LI + 2) g values
// of each axis once and forms every element of the row from them: about one
// LDS read per element instead of three.  The accumulators stay in
// registers (two workgroups per CU; one for nabla_i, whose three components
// per element need the larger register budget); a block with more rows than
// one chunk (256 RPL rows) is done chunk by chunk, each chunk a full pass
// over the primitive tuples with the roots and chains recomputed.  WIDE:
// such a block runs with about 60 accumulators per lane (nabla_i: 90) at
// one workgroup per CU (the register budget of one wave per SIMD), so a
// (g f|g p) block takes 3 chunks instead of 9.
constexpr size_t kGenLdsBytes = 150 * 1024;  // dynamic LDS for g (gfx950: 160 KB per CU)

// rows per lane: about 16 accumulators (nabla_i: 45, WIDE: 60 (48 for an s
// shell i, whose 60 rows spill), WIDE nabla_i: 90, each with one workgroup
// per CU)
__host__ __device__ constexpr int gen_rows_per_lane(int li, bool deriv, bool wide = false) {
  return deriv ? cmax(1, (wide ? 30 : 15) / ncart(li))
               : cmax(1, (wide ? (li > 0 ? 60 : 48) : 16) / ncart(li));
}

template <bool GLDS, bool DERIV, int LI, bool WIDE = false>
__global__ void __launch_bounds__(256, (DERIV || WIDE ? 1 : 2))
    eri_generic_kernel(EriArgs a, GenShape S, int64_t per_block) {
  constexpr int NC = DERIV ? 3 : 1;
  constexpr int NCI = ncart(LI);                           // elements per row
  constexpr int RPL = gen_rows_per_lane(LI, DERIV, WIDE);  // rows per lane
  constexpr int NLD = LI + 1 + (DERIV ? 1 : 0);       // i powers loaded per axis
  constexpr int ACC = RPL * NCI;                      // elements per lane
  constexpr int CHUNK = 256 * RPL;                    // rows per chunk
  extern __shared__ double gl[];
  const int t = threadIdx.x;
  const int T = blockDim.x;
  const int NR = S.nr;
  double* slice = a.scratch + (size_t)blockIdx.x * per_block;
  double* g = GLDS ? gl : slice;
  double* acc = slice + (size_t)3 * S.g4 * NR;  // [ncomp][nf]: chunk results
  double* b0 = acc + (size_t)S.nf * S.ncomp;
  double* b1 = b0 + S.nf;
  __shared__ double uw[3 * kMaxRoots];
  __shared__ unsigned long long cnt[2];
  const int nfi = ncart(S.l[0]), nfj = ncart(S.l[1]), nfk = ncart(S.l[2]), nfl = ncart(S.l[3]);
  const int SI = S.dj * S.dk * S.dl;
  const int IJ = S.di * S.dj;
  const int ket_tasks = (S.l[2] > 0 && S.l[3] > 0) ? 3 * IJ * NR : 0;
  const int NM = S.l[2] + S.l[3] + 1;  // VRR columns
  const int bra_tasks = 3 * NM * NR;
  // this lane's rows: offsets (in doubles) of their x, y, z g entries with
  // i power 0
  const int nrow = nfj * nfk * nfl;
  int ox[RPL], oy[RPL], oz[RPL];
  double racc[NC * ACC];
  for (int64_t q = blockIdx.x; q < a.nq; q += gridDim.x) {
    const int2 pq = a.quartets[q];
    const int boff = a.pair_off[pq.x], koff = a.pair_off[pq.y];
    const double* gb = a.pair_geo + 8 * (size_t)pq.x;
    const double* gk = a.pair_geo + 8 * (size_t)pq.y;
    // coordinate differences (positions are formed from these and the
    // primitive weights, never from an absolute product centre)
    double rji[3], rlk[3], rik[3];
    for (int d = 0; d < 3; ++d) {
      rji[d] = gb[3 + d] - gb[d];
      rlk[d] = gk[3 + d] - gk[d];
      rik[d] = gb[d] - gk[d];
    }
    if (t == 0) cnt[0] = 0;
    const double logd_ij = gb[6], logd_kl = gk[6];
    const int np_ij = a.pair_np[pq.x], np_kl = a.pair_np[pq.y];
    for (int r0 = 0; r0 < nrow; r0 += CHUNK) {
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int row = r0 + t + k * 256;
        int px = 0, py = 0, pz = 0;
        if (row < nrow) {  // rows past the block read entry 0 and are not stored
          const int fj = row % nfj, e2 = row / nfj;
          const int fk = e2 % nfk, fl = e2 / nfk;
          int jx, jy, jz, kx, ky, kz, lx_, ly_, lz_;
          cart_decode(S.l[1], fj, jx, jy, jz);
          cart_decode(S.l[2], fk, kx, ky, kz);
          cart_decode(S.l[3], fl, lx_, ly_, lz_);
          px = (jx * S.dk + kx) * S.dl + lx_;
          py = (jy * S.dk + ky) * S.dl + ly_;
          pz = (jz * S.dk + kz) * S.dl + lz_;
        }
        ox[k] = px * NR;
        oy[k] = (S.g4 + py) * NR;
        oz[k] = (2 * S.g4 + pz) * NR;
      }
#pragma unroll
      for (int k = 0; k < NC * ACC; ++k) racc[k] = 0.0;
      __syncthreads();
      for (int kp = 0; kp < np_kl; ++kp) {
        const int kidx = koff + kp;
        const double* pk = a.prim + 8 * (size_t)kidx;
        const double akl = pk[0];
        const double ccekl = a.three_center ? 0.0 : pk[6] - a.lkl_ceil * logd_kl;
        if (ccekl > a.expcutoff) continue;  // uniform across the workgroup
        const double eijcut = a.expcutoff - ccekl;
        // transfer base of the ket: libcint's (lk > ll) unless both shells
        // carry angular momentum, then the centre of the larger exponent
        // (Q lies near it, so the VRR terms stay small and the transfer
        // adds instead of cancelling; FP64 results of (f|g)-type quartets
        // with distant centres otherwise lose up to 7 digits)
        const bool kb = (S.l[2] > 0 && S.l[3] > 0) ? (pk[7] >= akl - pk[7]) : (S.kbase != 0);
        double rkl[3], qk[3], qb[3];  // HRR step, Q - R_k, Q - R(ket base)
        for (int d = 0; d < 3; ++d) {
          rkl[d] = kb ? -rlk[d] : rlk[d];
          qk[d] = pk[2] * rlk[d];
          qb[d] = kb ? qk[d] : -pk[1] * rlk[d];
        }
        for (int ip = 0; ip < np_ij; ++ip) {
          const int iidx = boff + ip;
          const double* pi = a.prim + 8 * (size_t)iidx;
          const double cceij = pi[6] - a.lij_ceil * logd_ij;
          // (libcint zeroes the prefactor of a bra pair with cce >= expcutoff)
          if (cceij > eijcut || cceij >= a.expcutoff) continue;
          const double aij = pi[0];
          // the bra's transfer base chosen like the ket's (uniform across
          // the workgroup: one tuple, one primitive tuple at a time)
          const bool ib = (S.lic > 0 && S.l[1] > 0) ? (pi[7] >= aij - pi[7]) : (S.ibase != 0);
          double dx[3], prb[3], qrk[3], rij[3], rr = 0;
          for (int d = 0; d < 3; ++d) {
            const double pr = pi[2] * rji[d];  // P - R_i
            dx[d] = (rik[d] + pr) - qk[d];
            rr += dx[d] * dx[d];
            prb[d] = ib ? pr : -pi[1] * rji[d];
            qrk[d] = qb[d];
            rij[d] = ib ? -rji[d] : rji[d];
          }
          const double inv_sum = 1.0 / (aij + akl);
          const double hij = pi[5], hkl = pk[5];
          const double a0 = aij * akl * inv_sum;
          const double fac = a.common * pi[4] * pk[4];
          const double fac1 = 4.0 * hij * hkl * sqrt(inv_sum) * fac;
          const double x = a0 * rr;
          if (t < NR) {
            double sr, qr, w;
            rys_eval(a.rys, NR, t, x, sr, qr, w);
            uw[t] = sr;
            uw[NR + t] = w * fac1;
            uw[2 * NR + t] = qr;
          }
          if (t == 0 && r0 == 0) ++cnt[0];
          __syncthreads();
          for (int c = t; c < 3 * NR; c += T) {
            const int axis = c / NR, r = c - axis * NR;
            generic_vrr(g, S, axis, r, uw[r], uw[2 * NR + r], uw[NR + r], inv_sum, hij, hkl, aij,
                        akl, dx[axis], prb[axis], qrk[axis]);
          }
          __syncthreads();
          for (int c = t; c < bra_tasks; c += T) {
            const int r = c % NR, rest = c / NR;
            const int m = rest % NM, axis = rest / NM;
            generic_bra(g, S, axis, r, m, rij[axis], ib);
          }
          __syncthreads();
          if (ket_tasks) {  // (with min(lk, ll) = 0 the slots are already (k, l))
            for (int c = t; c < ket_tasks; c += T) {
              const int r = c % NR, rest = c / NR;
              const int ij = rest % IJ, axis = rest / IJ;
              generic_ket(g, S, axis, r, ij, rkl[axis], kb);
            }
            __syncthreads();
          }
          // roots outer, rows inner: a row's g values of one root are loaded
          // once (the loads of the lane's rows are independent, so they
          // overlap) and every element of the row is formed from them
          const double ai2 = -2.0 * pi[7];
          const int up = SI * NR;  // stride of the i power
          for (int r = 0; r < NR; ++r) {
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
              double xv[NLD], yv[NLD], zv[NLD];
#pragma unroll
              for (int i = 0; i < NLD; ++i) {
                xv[i] = g[ox[k] + i * up + r];
                yv[i] = g[oy[k] + i * up + r];
                zv[i] = g[oz[k] + i * up + r];
              }
#pragma unroll
              for (int fi = 0; fi < NCI; ++fi) {
                int ix, iy, iz;
                cart_decode(LI, fi, ix, iy, iz);
                if constexpr (!DERIV) {
                  racc[k * NCI + fi] += xv[ix] * yv[iy] * zv[iz];
                } else {
                  double fx = ai2 * xv[ix + 1];
                  double fy = ai2 * yv[iy + 1];
                  double fz = ai2 * zv[iz + 1];
                  if (ix > 0) fx += (double)ix * xv[ix - 1];
                  if (iy > 0) fy += (double)iy * yv[iy - 1];
                  if (iz > 0) fz += (double)iz * zv[iz - 1];
                  racc[k * NCI + fi] += fx * yv[iy] * zv[iz];
                  racc[ACC + k * NCI + fi] += xv[ix] * fy * zv[iz];
                  racc[2 * ACC + k * NCI + fi] += xv[ix] * yv[iy] * fz;
                }
              }
            }
          }
          __syncthreads();
        }
      }
      // the chunk's results to the slice (element fi of row `row` is
      // fi + NCI row: i fastest)
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int row = r0 + t + k * 256;
        if (row < nrow) {
#pragma unroll
          for (int comp = 0; comp < NC; ++comp)
#pragma unroll
            for (int fi = 0; fi < NCI; ++fi)
              acc[(size_t)comp * S.nf + (size_t)row * NCI + fi] = racc[comp * ACC + k * NCI + fi];
        }
      }
    }
    __syncthreads();
    if (a.counters && t == 0 && cnt[0]) {
      atomicAdd(a.counters, cnt[0]);
      atomicAdd(a.counters + 1, 1ull);
    }
    const int ns0 = a.cart ? nfi : nsph(S.l[0]), ns1 = a.cart ? nfj : nsph(S.l[1]);
    const int ns2 = a.cart ? nfk : nsph(S.l[2]), ns3 = a.cart ? nfl : nsph(S.l[3]);
    const int NS = ns0 * ns1 * ns2 * ns3;
    for (int comp = 0; comp < S.ncomp; ++comp) {
      for (int e = t; e < S.nf; e += T) b0[e] = acc[(size_t)comp * S.nf + e];
      __syncthreads();
      int dims[4] = {nfi, nfj, nfk, nfl};
      double* src = b0;
      double* dst = b1;
      for (int ax = 0; ax < 4; ++ax) {
        if (S.l[ax] < 2 || a.cart) continue;
        c2s_pass_mfma(dst, src, dims, ax, S.l[ax], a.c2s, t, T);
        dims[ax] = nsph(S.l[ax]);
        double* tmp = src;
        src = dst;
        dst = tmp;
        __syncthreads();
      }
      double* o = a.out + a.out_base + q * a.block + comp * a.st_comp;
      for (int e = t; e < NS; e += T) {
        int mi = e % ns0, e1 = e / ns0;
        int mj = e1 % ns1, e2 = e1 / ns1;
        int mk = e2 % ns2, ml = e2 / ns2;
        o[mi * a.st[0] + mj * a.st[1] + mk * a.st[2] + ml * a.st[3]] = src[e];
      }
      __syncthreads();
    }
  }
}

int64_t generic_scratch_per_block(const int* l, int deriv) {
  int lic = l[0] + deriv;
  int nr = (lic + l[1] + l[2] + l[3]) / 2 + 1;
  int64_t g4 = (int64_t)(lic + 1) * (l[1] + 1) * (l[2] + 1) * (l[3] + 1);
  int64_t nf = (int64_t)ncart(l[0]) * ncart(l[1]) * ncart(l[2]) * ncart(l[3]);
  int64_t n = 3 * g4 * nr + nf * (deriv ? 3 : 1) + 2 * nf;
  return (n + 31) / 32 * 32;
}

hipError_t launch_generic(const EriArgs& a, const int* l, int deriv, int grid, hipStream_t s) {
  GenShape S;
  for (int c = 0; c < 4; ++c) S.l[c] = l[c];
  S.deriv = deriv;
  S.lic = l[0] + deriv;
  S.nr = (S.lic + l[1] + l[2] + l[3]) / 2 + 1;
  S.di = S.lic + 1;
  S.dj = l[1] + 1;
  S.dk = l[2] + 1;
  S.dl = l[3] + 1;
  S.g4 = S.di * S.dj * S.dk * S.dl;
  S.nf = ncart(l[0]) * ncart(l[1]) * ncart(l[2]) * ncart(l[3]);
  S.ncomp = deriv ? 3 : 1;
  S.ibase = S.lic > l[1];
  S.kbase = l[2] > l[3];
  int64_t per = generic_scratch_per_block(l, deriv);
  const size_t gbytes = (size_t)3 * S.g4 * S.nr * sizeof(double);
  const bool glds = gbytes <= kGenLdsBytes;
  const size_t dyn = glds ? gbytes : 0;
  // more rows than one chunk of the narrow rows: the wide kernel (nabla_i
  // with an s shell i keeps the narrow one: 30 rows per lane spill)
  const int nrow = ncart(l[1]) * ncart(l[2]) * ncart(l[3]);
  const bool wide = nrow > 256 * gen_rows_per_lane(l[0], deriv != 0) && !(deriv && l[0] == 0);
  switch (l[0]) {
#define CX_GEN_LAUNCH(LI)                                                                     \
  case LI:                                                                                    \
    if (glds && deriv && wide)                                                                \
      hipLaunchKernelGGL((eri_generic_kernel<true, true, LI, true>), dim3(grid), dim3(256), dyn, \
                         s, a, S, per);                                                       \
    else if (glds && deriv)                                                                   \
      hipLaunchKernelGGL((eri_generic_kernel<true, true, LI>), dim3(grid), dim3(256), dyn, s, a, \
                         S, per);                                                             \
    else if (glds && wide)                                                                    \
      hipLaunchKernelGGL((eri_generic_kernel<true, false, LI, true>), dim3(grid), dim3(256),   \
                         dyn, s, a, S, per);                                                  \
    else if (glds)                                                                            \
      hipLaunchKernelGGL((eri_generic_kernel<true, false, LI>), dim3(grid), dim3(256), dyn, s,  \
                         a, S, per);                                                          \
    else if (deriv)                                                                           \
      hipLaunchKernelGGL((eri_generic_kernel<false, true, LI>), dim3(grid), dim3(256), 0, s, a, \
                         S, per);                                                             \
    else                                                                                      \
      hipLaunchKernelGGL((eri_generic_kernel<false, false, LI>), dim3(grid), dim3(256), 0, s,   \
                         a, S, per);                                                          \
    break;
    CX_GEN_LAUNCH(0)
    CX_GEN_LAUNCH(1)
    CX_GEN_LAUNCH(2)
    CX_GEN_LAUNCH(3)
    CX_GEN_LAUNCH(4)
#undef CX_GEN_LAUNCH
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace cx
