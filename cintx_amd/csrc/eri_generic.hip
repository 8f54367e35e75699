// Runtime-class fallback kernel: any (li, lj, lk, ll) with l <= 4 (+1 for
// nabla_i), used for the classes that have no compile-time instantiation
// (high angular momentum: (ff|..), (gg|gg), the s..g basis workload).  One
// shell tuple per 256-lane workgroup, grid-stride over tuples.  The g tensor
// lives in LDS when it fits in 56 KB, else in a per-workgroup global scratch
// slice; the Cartesian accumulators live in registers (16 elements per lane)
// up to 4096 elements, else in the slice; the cart->sph buffers are in the
// slice.  Arithmetic is the same as eri_class_kernel (libcint g2e.c:272-690,
// 4425-4620).
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

__device__ void generic_chain(double* __restrict__ g, const GenShape& S, int axis, int r,
                              double sr, double qr, double wz, double inv_sum, double hij, double hkl,
                              double aij, double akl, double dx, double prb, double qrk,
                              double drb, double drk) {
  const int NR = S.nr;
  const int NMAX = S.lic + S.l[1], MMAX = S.l[2] + S.l[3];
  const int A = S.ibase ? S.lic : S.l[1], B = S.ibase ? S.l[1] : S.lic;
  const int C = S.kbase ? S.l[2] : S.l[3], D = S.kbase ? S.l[3] : S.l[2];
  const int DJ = S.dj, DL = S.dl, BLK = S.dk * S.dl;
  const double sa = sr * inv_sum, oms = qr;
  const double b00 = 0.5 * sa, b10 = 0.5 * sa + oms * hij, b01 = 0.5 * sa + oms * hkl;
  const double c00 = prb - sa * akl * dx;
  const double c0p = qrk + sa * aij * dx;
  double* ga = g + (size_t)axis * S.g4 * NR + r;
  double cm1[12], c0[12], c1[12], t[12];
  c0[0] = (axis == 2) ? wz : 1.0;
  for (int n = 0; n < NMAX; ++n) c0[n + 1] = c00 * c0[n] + (n > 0 ? n * b10 * c0[n - 1] : 0.0);
  for (int m = 0; m <= MMAX; ++m) {
    if (m > 0) {
      for (int n = 0; n <= NMAX; ++n) {
        double v = c0p * c0[n];
        if (m > 1) v += (m - 1) * b01 * cm1[n];
        if (n > 0) v += n * b00 * c0[n - 1];
        c1[n] = v;
      }
      for (int n = 0; n <= NMAX; ++n) {
        cm1[n] = c0[n];
        c0[n] = c1[n];
      }
    }
    // bra transfer of column m into block(i, j) slot m
    for (int a = 0; a <= NMAX; ++a) t[a] = c0[a];
    for (int b = 0; b <= B; ++b) {
      if (b > 0)
        for (int a = 0; a <= A + B - b; ++a) t[a] = t[a + 1] + drb * t[a];
      for (int a = 0; a <= A; ++a) {
        const int i = S.ibase ? a : b, j = S.ibase ? b : a;
        ga[((i * DJ + j) * BLK + m) * NR] = t[a];
      }
    }
  }
  for (int ij = 0; ij < S.di * S.dj; ++ij) {
    double* blk = ga + (size_t)ij * BLK * NR;
    for (int c = 0; c <= MMAX; ++c) t[c] = blk[c * NR];
    for (int d = 0; d <= D; ++d) {
      if (d > 0)
        for (int c = 0; c <= C + D - d; ++c) t[c] = t[c + 1] + drk * t[c];
      for (int c = 0; c <= C; ++c) {
        const int k = S.kbase ? c : d, l = S.kbase ? d : c;
        blk[(k * DL + l) * NR] = t[c];
      }
    }
  }
}

// GLDS: the g tensor in (dynamic) LDS instead of the global slice.
// AREG: every lane keeps its at most kGenAcc elements (t + 256 k) and their
// packed g positions in registers instead of global accumulators.
constexpr int kGenAcc = 16;
constexpr size_t kGenLdsBytes = 56 * 1024;  // dynamic LDS for g (plus ~0.5 KB static)

template <bool GLDS, bool AREG>
__global__ void __launch_bounds__(256) eri_generic_kernel(EriArgs a, GenShape S,
                                                          int64_t per_block) {
  extern __shared__ double gl[];
  const int t = threadIdx.x;
  const int T = blockDim.x;
  const int NR = S.nr;
  double* slice = a.scratch + (size_t)blockIdx.x * per_block;
  double* g = GLDS ? gl : slice;
  double* acc = slice + (size_t)3 * S.g4 * NR;
  double* b0 = acc + (size_t)S.nf * S.ncomp;
  double* b1 = b0 + S.nf;
  __shared__ double uw[3 * kMaxRoots];
  __shared__ unsigned long long cnt[2];
  const int nfi = ncart(S.l[0]), nfj = ncart(S.l[1]), nfk = ncart(S.l[2]), nfl = ncart(S.l[3]);
  const int SI = S.dj * S.dk * S.dl;
  // this lane's elements: packed g positions x | y << 10 | z << 20 and the
  // i powers (nabla_i) x | y << 4 | z << 8
  int pos[AREG ? kGenAcc : 1], ipw[AREG ? kGenAcc : 1];
  double racc[AREG ? 3 * kGenAcc : 1];
  if constexpr (AREG) {
#pragma unroll
    for (int k = 0; k < kGenAcc; ++k) {
      const int e = t + k * 256;
      pos[k] = 0;
      ipw[k] = 0;
      if (e < S.nf) {
        int fi = e % nfi, e1 = e / nfi;
        int fj = e1 % nfj, e2 = e1 / nfj;
        int fk = e2 % nfk, fl = e2 / nfk;
        int ix, iy, iz, jx, jy, jz, kx, ky, kz, lx, ly, lz;
        cart_decode(S.l[0], fi, ix, iy, iz);
        cart_decode(S.l[1], fj, jx, jy, jz);
        cart_decode(S.l[2], fk, kx, ky, kz);
        cart_decode(S.l[3], fl, lx, ly, lz);
        const int px = ((ix * S.dj + jx) * S.dk + kx) * S.dl + lx;
        const int py = ((iy * S.dj + jy) * S.dk + ky) * S.dl + ly;
        const int pz = ((iz * S.dj + jz) * S.dk + kz) * S.dl + lz;
        pos[k] = px | (py << 10) | (pz << 20);
        ipw[k] = ix | (iy << 4) | (iz << 8);
      }
    }
  }
  for (int64_t q = blockIdx.x; q < a.nq; q += gridDim.x) {
    const int2 pq = a.quartets[q];
    const int boff = a.pair_off[pq.x], koff = a.pair_off[pq.y];
    const double* gb = a.pair_geo + 8 * (size_t)pq.x;
    const double* gk = a.pair_geo + 8 * (size_t)pq.y;
    double ri[3], rj[3], rk[3], rl[3], rij[3], rkl[3];
    for (int d = 0; d < 3; ++d) {
      ri[d] = gb[d];
      rj[d] = gb[3 + d];
      rk[d] = gk[d];
      rl[d] = gk[3 + d];
      rij[d] = S.ibase ? ri[d] - rj[d] : rj[d] - ri[d];
      rkl[d] = S.kbase ? rk[d] - rl[d] : rl[d] - rk[d];
    }
    if constexpr (AREG) {
#pragma unroll
      for (int k = 0; k < 3 * kGenAcc; ++k) racc[k] = 0.0;
    } else {
      for (int e = t; e < S.nf * S.ncomp; e += T) acc[e] = 0.0;
    }
    if (t == 0) cnt[0] = 0;
    __syncthreads();
    const double logd_ij = gb[6], logd_kl = gk[6];
    const int np_ij = a.pair_np[pq.x], np_kl = a.pair_np[pq.y];
    for (int kp = 0; kp < np_kl; ++kp) {
      const int kidx = koff + kp;
      const double* pk = a.prim + 8 * (size_t)kidx;
      const double akl = pk[0];
      const double Qv[3] = {pk[1], pk[2], pk[3]};
      const double ccekl = a.three_center ? 0.0 : pk[6] - a.lkl_ceil * logd_kl;
      if (ccekl > a.expcutoff) continue;  // uniform across the workgroup
      const double eijcut = a.expcutoff - ccekl;
      for (int ip = 0; ip < np_ij; ++ip) {
        const int iidx = boff + ip;
        const double* pi = a.prim + 8 * (size_t)iidx;
        const double cceij = pi[6] - a.lij_ceil * logd_ij;
        // (libcint zeroes the prefactor of a bra pair with cce >= expcutoff)
        if (cceij > eijcut || cceij >= a.expcutoff) continue;
        const double aij = pi[0];
        const double Pv[3] = {pi[1], pi[2], pi[3]};
        double dx[3], prb[3], qrk[3], rr = 0;
        for (int d = 0; d < 3; ++d) {
          dx[d] = Pv[d] - Qv[d];
          rr += dx[d] * dx[d];
          prb[d] = Pv[d] - (S.ibase ? ri[d] : rj[d]);
          qrk[d] = Qv[d] - (S.kbase ? rk[d] : rl[d]);
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
        if (t == 0) ++cnt[0];
        __syncthreads();
        for (int c = t; c < 3 * NR; c += T) {
          const int axis = c / NR, r = c - axis * NR;
          generic_chain(g, S, axis, r, uw[r], uw[2 * NR + r], uw[NR + r], inv_sum, hij, hkl, aij,
                        akl, dx[axis], prb[axis], qrk[axis], rij[axis], rkl[axis]);
        }
        __syncthreads();
        const double ai2 = -2.0 * pi[7];
        if constexpr (AREG) {
#pragma unroll
          for (int k = 0; k < kGenAcc; ++k) {
            if (t + k * 256 >= S.nf) break;
            const double* gx = g + (size_t)(pos[k] & 1023) * NR;
            const double* gy = g + (size_t)(S.g4 + ((pos[k] >> 10) & 1023)) * NR;
            const double* gz = g + (size_t)(2 * S.g4 + (pos[k] >> 20)) * NR;
            if (!S.deriv) {
              double sum = 0;
              for (int r = 0; r < NR; ++r) sum += gx[r] * gy[r] * gz[r];
              racc[k] += sum;
            } else {
              const int ix = ipw[k] & 15, iy = (ipw[k] >> 4) & 15, iz = ipw[k] >> 8;
              double sx = 0, sy = 0, sz = 0;
              for (int r = 0; r < NR; ++r) {
                double fx = ai2 * gx[SI * NR + r], fy = ai2 * gy[SI * NR + r],
                       fz = ai2 * gz[SI * NR + r];
                if (ix > 0) fx += ix * gx[r - SI * NR];
                if (iy > 0) fy += iy * gy[r - SI * NR];
                if (iz > 0) fz += iz * gz[r - SI * NR];
                sx += fx * gy[r] * gz[r];
                sy += gx[r] * fy * gz[r];
                sz += gx[r] * gy[r] * fz;
              }
              racc[k] += sx;
              racc[kGenAcc + k] += sy;
              racc[2 * kGenAcc + k] += sz;
            }
          }
        } else {
          for (int e = t; e < S.nf; e += T) {
            int fi = e % nfi, e1 = e / nfi;
            int fj = e1 % nfj, e2 = e1 / nfj;
            int fk = e2 % nfk, fl = e2 / nfk;
            int ix, iy, iz, jx, jy, jz, kx, ky, kz, lx, ly, lz;
            cart_decode(S.l[0], fi, ix, iy, iz);
            cart_decode(S.l[1], fj, jx, jy, jz);
            cart_decode(S.l[2], fk, kx, ky, kz);
            cart_decode(S.l[3], fl, lx, ly, lz);
            const double* gx = g + (size_t)(((ix * S.dj + jx) * S.dk + kx) * S.dl + lx) * NR;
            const double* gy =
                g + (size_t)(S.g4 + ((iy * S.dj + jy) * S.dk + ky) * S.dl + ly) * NR;
            const double* gz =
                g + (size_t)(2 * S.g4 + ((iz * S.dj + jz) * S.dk + kz) * S.dl + lz) * NR;
            if (!S.deriv) {
              double sum = 0;
              for (int r = 0; r < NR; ++r) sum += gx[r] * gy[r] * gz[r];
              acc[e] += sum;
            } else {
              double sx = 0, sy = 0, sz = 0;
              for (int r = 0; r < NR; ++r) {
                double fx = ai2 * gx[SI * NR + r], fy = ai2 * gy[SI * NR + r],
                       fz = ai2 * gz[SI * NR + r];
                if (ix > 0) fx += ix * gx[r - SI * NR];
                if (iy > 0) fy += iy * gy[r - SI * NR];
                if (iz > 0) fz += iz * gz[r - SI * NR];
                sx += fx * gy[r] * gz[r];
                sy += gx[r] * fy * gz[r];
                sz += gx[r] * gy[r] * fz;
              }
              acc[e] += sx;
              acc[S.nf + e] += sy;
              acc[2 * S.nf + e] += sz;
            }
          }
        }
        __syncthreads();
      }
    }
    if (a.counters && t == 0 && cnt[0]) {
      atomicAdd(a.counters, cnt[0]);
      atomicAdd(a.counters + 1, 1ull);
    }
    const int ns0 = a.cart ? nfi : nsph(S.l[0]), ns1 = a.cart ? nfj : nsph(S.l[1]);
    const int ns2 = a.cart ? nfk : nsph(S.l[2]), ns3 = a.cart ? nfl : nsph(S.l[3]);
    const int NS = ns0 * ns1 * ns2 * ns3;
    for (int comp = 0; comp < S.ncomp; ++comp) {
      if constexpr (AREG) {
#pragma unroll
        for (int k = 0; k < kGenAcc; ++k)
          if (t + k * 256 < S.nf) b0[t + k * 256] = racc[comp * kGenAcc + k];
      } else {
        for (int e = t; e < S.nf; e += T) b0[e] = acc[comp * S.nf + e];
      }
      __syncthreads();
      int dims[4] = {nfi, nfj, nfk, nfl};
      double* src = b0;
      double* dst = b1;
      for (int ax = 0; ax < 4; ++ax) {
        if (S.l[ax] < 2 || a.cart) continue;
        c2s_pass(dst, src, dims, ax, S.l[ax], a.c2s, t, T);
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
  const bool areg = S.nf <= 256 * kGenAcc;
  const size_t dyn = glds ? gbytes : 0;
  if (glds && areg)
    hipLaunchKernelGGL((eri_generic_kernel<true, true>), dim3(grid), dim3(256), dyn, s, a, S, per);
  else if (glds)
    hipLaunchKernelGGL((eri_generic_kernel<true, false>), dim3(grid), dim3(256), dyn, s, a, S,
                       per);
  else if (areg)
    hipLaunchKernelGGL((eri_generic_kernel<false, true>), dim3(grid), dim3(256), 0, s, a, S, per);
  else
    hipLaunchKernelGGL((eri_generic_kernel<false, false>), dim3(grid), dim3(256), 0, s, a, S,
                       per);
  return hipGetLastError();
}

}  // namespace cx
