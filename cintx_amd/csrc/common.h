// Shared host/device definitions for libcintx_hip.
//
// Data layout in HBM (all device-resident, built once per basis):
//   shells      : virtual (segmented, nctr == 1) shells; general contractions of
//                 the caller's basis are split into one virtual shell per
//                 contraction column, which is exact by linearity.
//   pairs       : one record per shell pair (i, j) that a plan needs.
//   prim pairs  : one 64-byte record per primitive pair of all pairs,
//                 contiguous per pair, in ascending screening key cce
//                 (CINTset_pairdata's jp-major order, libcint-master/src/
//                 optimizer.c:317-341, only breaks ties).
//   quartets    : (bra pair, ket pair) int32 tuples, binned by class.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace cx {

constexpr int kMaxL = 4;          // largest shell l on the hot path (g)
constexpr int kMaxRoots = 14;     // table covers n = 1 .. kMaxRoots
constexpr int kRysDeg = 11;       // Chebyshev degree per unit x-interval
constexpr int kRysNc = kRysDeg + 1;
constexpr int kC2sL = 6;          // c2s table covers l = 0 .. kC2sL
constexpr int kC2sM = 2 * kC2sL + 1;
constexpr int kC2sC = (kC2sL + 1) * (kC2sL + 2) / 2;

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nsph(int l) { return 2 * l + 1; }
__host__ __device__ constexpr double rys_largex(int n) { return 35.0 + 5.0 * n; }

// Rys table: for n in 1..kMaxRoots, x in [0, largex(n)) split into unit
// intervals; per interval, per root r: kRysNc Chebyshev coefficients of
// s_r = t_r^2 = u_r/(1+u_r) followed (after all n roots) by those of w_r.
// For x >= largex(n) the asymptotic rule s = R/x, w = W/sqrt(x) is used.
struct RysTable {
  const double* coef;   // device
  const double* large;  // device: per n at n(n-1): R[0..n-1], W[0..n-1]
  int off[kMaxRoots + 1];  // start of n's block in coef
};

// Per-launch arguments of a class kernel.
struct EriArgs {
  // per primitive pair one 64-byte record (a lane reads one cache line per
  // primitive pair): a_ij, w_i = a_i/a_ij, w_j = a_j/a_ij, 0,
  // kfac = exp(-a_i a_j/a_ij |R_ij|^2) c_i c_j, 0.5/a_ij, the screening key
  // cce (without the l-dependent term) and a_i (nabla_i).  The product
  // centre is P = R_i + w_j (R_j - R_i); kernels use only the differences
  // P - R_i = w_j (R_j - R_i) and P - R_j = -w_i (R_j - R_i).
  const double* prim;
  // per pair
  const int* pair_off;  // first prim pair of the pair
  const int* pair_np;   // primitive pairs of the pair
  const double* pair_geo;  // 8 per pair: ri[3], rj[3], log(|Rij|+1), unused
  // work list
  const int2* quartets;    // (bra pair, ket pair)
  int64_t out_base;        // element offset of this launch's first block
  int64_t block;           // elements per block (components included)
  int64_t nq;
  double common;           // pi^3 2/sqrt(pi) prod common_fac_sp
  double expcutoff;
  int lij_ceil, lkl_ceil;  // li_ceil+lj_ceil, lk_ceil+ll_ceil for the screen
  int three_center;        // ket pairs are pseudo pairs (no kl screen)
  int cart;                // generic kernel: Cartesian output, no cart -> sph
  // output strides for the four canonical axes and the component
  int64_t st[4];
  int64_t st_comp;
  double* out;
  RysTable rys;
  const double* c2s;       // [kC2sL+1][kC2sM][kC2sC]
  double* scratch;         // generic kernel only
  unsigned long long* counters;  // optional: [0] surviving primitive tuples
  // fused J/K (int2e_sph blocks of unique quartets): when jk_D is set the
  // class kernel digests each block into the accumulators jk_J / jk_K
  // (halves of J and K, symmetrised afterwards) instead of writing it
  const double* jk_D;      // symmetric density, nao x nao
  double* jk_J;
  double* jk_K;
  const int2* pair_sh;     // original shells (a, b) of every pair
  const int* ao_off;       // first spherical function of every shell
  int nao;
};

// ------------------------------------------------------------------
// Cartesian -> spinor coupling table: per l = 0..kMaxL, the 4l+2 spinors
// of the shell as rows (first the 2l of j = l-1/2, then the 2l+2 of
// j = l+1/2, each m_j = -j..j), each row 2 ncart(l) complex coefficients:
// the alpha then the beta component over libcint's Cartesian functions.
// kappa < 0 selects the j = l+1/2 rows, kappa > 0 the j = l-1/2 rows,
// kappa = 0 all rows in table order (libcint's KAPPA_OF convention,
// cart2sph.c:3537-3547).
// ------------------------------------------------------------------
__host__ __device__ constexpr int spinor_tab_off(int l) {
  return l <= 0 ? 0 : spinor_tab_off(l - 1) + (4 * (l - 1) + 2) * 2 * ncart(l - 1);
}
constexpr int kSpinorTab = spinor_tab_off(kMaxL + 1);  // complex entries
__host__ __device__ constexpr int spinor_len(int l, int kappa) {
  return kappa == 0 ? 4 * l + 2 : (kappa < 0 ? 2 * l + 2 : 2 * l);
}
__host__ __device__ constexpr int spinor_row0(int l, int kappa) { return kappa < 0 ? 2 * l : 0; }

// One contracted block of one tuple for the spinor transform.
struct SpinorJob {
  int64_t src;    // first double of the Cartesian block (canonical slot order)
  int64_t dst;    // first complex element of the output block
  int so[3];      // output strides (complex elements) of argument slots 1..3
  int cs[4];      // strides of argument slots 0..3 inside the Cartesian block
  int8_t l[4];    // angular momentum per argument slot
  int8_t kappa[4];  // sign of KAPPA_OF per argument slot (see above)
};

// complex elements of the per-workgroup workspace a job needs
__host__ __device__ inline int64_t spinor_job_slice(const SpinorJob& J, bool four) {
  const int fi = ncart(J.l[0]), fj = ncart(J.l[1]), fk = ncart(J.l[2]);
  const int fl = four ? ncart(J.l[3]) : 1;
  const int di = spinor_len(J.l[0], J.kappa[0]), dj = spinor_len(J.l[1], J.kappa[1]);
  const int dk = four ? spinor_len(J.l[2], J.kappa[2]) : nsph(J.l[2]);
  const int64_t n1 = 2LL * di * fj * fk * fl, n2 = (int64_t)di * dj * fk * fl;
  const int64_t n3 = four ? 2LL * di * dj * dk * fl : 0;
  return (n1 > n3 ? n1 : n3) + n2;
}

}  // namespace cx
