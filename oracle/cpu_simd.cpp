// cpu_simd.cpp -- CPU baseline for the int2e_sph throughput comparison.
//
// TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg and by
// tests/, never by the product library.  A C++ restatement of cintx's
// CPU/SIMD two-electron kernel (crates/cintx-simd/src/kernels/two_electron.rs:
// 102-410) for segmented shells:
//   * libcint's per-root coefficients (g2e.c:4523-4543), 2D VRR
//     (g2e.c:272-411) and HRR transfers (g2e.c:428-690), gout
//     (cint2e.c:961-1180), evaluated on f64x4 lanes over the primitive
//     index of the l shell (the `wide` f64x4 scheme of two_electron.rs:
//     208-230; unused lanes carry weight 0), primitives unscreened as in
//     cintx-simd;
//   * the Rys rule per lane from a degree-11 Chebyshev table per unit
//     x-interval (the asymptotic rule beyond 35+5n, rys_roots.c:67-77),
//     fitted at load to the oracle's long-double rule -- the same kind of
//     table-driven rule as the device kernels;
//   * cart -> sph with the oracle's coefficient table;
//   * OpenMP over rows of quartets that share a bra pair, schedule(dynamic, 2)
//     (libcint-master/examples/time_c60.c:196-227).
#include <math.h>
#include <omp.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "cint_oracle.h"

namespace {

typedef double v4 __attribute__((vector_size(32)));
constexpr int kLanes = 4;
constexpr int kMaxN = 9;   // (gg|gg)
constexpr int kMaxL = 4;
constexpr int kDeg = 11;
constexpr int kNc = kDeg + 1;

inline v4 splat(double x) { return (v4){x, x, x, x}; }

double largex(int n) { return 35.0 + 5.0 * n; }

struct RysTable {
  std::vector<double> coef[kMaxN + 1];  // [interval][root][s|w][kNc]
  double R[kMaxN + 1][kMaxN], W[kMaxN + 1][kMaxN];
};
RysTable g_tab;
double g_c2s[kMaxL + 1][2 * kMaxL + 1][(kMaxL + 1) * (kMaxL + 2) / 2];
std::once_flag g_once;

void build_tables() {
  const double pi = 3.14159265358979323846;
  double tj[kNc];
  for (int j = 0; j < kNc; ++j) tj[j] = cos(pi * (j + 0.5) / kNc);
  for (int n = 1; n <= kMaxN; ++n) {
    const int nint = (int)largex(n);
    std::vector<double>& c = g_tab.coef[n];
    c.assign((size_t)nint * n * 2 * kNc, 0.0);
    long double s[ORC_MAX_ROOTS], w[ORC_MAX_ROOTS];
    for (int iv = 0; iv < nint; ++iv) {
      // per root the smaller of s and 1 - s over the interval (decided at its
      // midpoint, so the fitted function is smooth), 1 - s stored negated
      bool use_q[kMaxN];
      orc_rys_rule_ld(n, iv + 0.5, s, w);
      for (int r = 0; r < n; ++r) use_q[r] = s[r] > 0.5L;
      double fs[kMaxN][kNc], fw[kMaxN][kNc];
      for (int j = 0; j < kNc; ++j) {
        orc_rys_rule_ld(n, iv + 0.5 * (tj[j] + 1.0), s, w);
        for (int r = 0; r < n; ++r) {
          fs[r][j] = use_q[r] ? -(double)(1.0L - s[r]) : (double)s[r];
          fw[r][j] = (double)w[r];
        }
      }
      for (int r = 0; r < n; ++r)
        for (int k = 0; k < kNc; ++k) {
          double as = 0, aw = 0;
          for (int j = 0; j < kNc; ++j) {
            const double t = cos(pi * k * (j + 0.5) / kNc);
            as += fs[r][j] * t;
            aw += fw[r][j] * t;
          }
          const double sc = (k == 0 ? 1.0 : 2.0) / kNc;
          c[(((size_t)iv * n + r) * 2 + 0) * kNc + k] = as * sc;
          c[(((size_t)iv * n + r) * 2 + 1) * kNc + k] = aw * sc;
        }
    }
    const double x0 = 1e8;  // s = R/x, w = W/sqrt(x) beyond largex(n)
    orc_rys_rule_ld(n, x0, s, w);
    for (int r = 0; r < n; ++r) {
      g_tab.R[n][r] = (double)(s[r] * x0);
      g_tab.W[n][r] = (double)(w[r] * sqrtl(x0));
    }
  }
  for (int l = 0; l <= kMaxL; ++l)
    for (int m = 0; m < 2 * l + 1; ++m)
      for (int c = 0; c < (l + 1) * (l + 2) / 2; ++c) g_c2s[l][m][c] = orc_c2s_coef(l, m, c);
}

// libcint's u = s/(1-s) and the weight, for n roots at x.
inline void rys_rule(int n, double x, double* u, double* w) {
  if (!(x < largex(n))) {
    for (int r = 0; r < n; ++r) {
      const double s = g_tab.R[n][r] / x;
      u[r] = s / (1.0 - s);
      w[r] = g_tab.W[n][r] / sqrt(x);
    }
    return;
  }
  const int iv = (int)x;
  const double t = 2.0 * (x - iv) - 1.0, t2 = 2.0 * t;
  const double* base = g_tab.coef[n].data() + (size_t)iv * n * 2 * kNc;
  for (int r = 0; r < n; ++r) {
    double v[2];
    for (int f = 0; f < 2; ++f) {
      const double* cf = base + ((size_t)r * 2 + f) * kNc;
      double b1 = 0, b2 = 0;
      for (int k = kDeg; k >= 1; --k) {
        const double b0 = t2 * b1 - b2 + cf[k];
        b2 = b1;
        b1 = b0;
      }
      v[f] = t * b1 - b2 + cf[0];
    }
    const double s = v[0] < 0 ? 1.0 + v[0] : v[0], q = v[0] < 0 ? -v[0] : 1.0 - v[0];
    u[r] = s / q;
    w[r] = v[1];
  }
}

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

struct Shell {
  int l, np;
  const double *r, *a, *c;
};

double common_fac_sp(int l) {
  if (l == 0) return 0.282094791773878143;
  if (l == 1) return 0.488602511902919921;
  return 1.0;
}

// cart (i fastest, l slowest) -> sph along every axis; out in the same order
void cart_to_sph(const double* cart, const int* ls, double* out, double* work) {
  int dim[4] = {ncart(ls[0]), ncart(ls[1]), ncart(ls[2]), ncart(ls[3])};
  const double* cur = cart;
  double* bufs[2] = {work, work + 50625};
  for (int a = 0; a < 4; ++a) {
    const int nc = dim[a], ns = 2 * ls[a] + 1;
    double* nxt = a == 3 ? out : bufs[a & 1];
    size_t inner = 1, outer = 1;
    for (int b = 0; b < a; ++b) inner *= dim[b];
    for (int b = a + 1; b < 4; ++b) outer *= dim[b];
    if (ls[a] < 2) {
      if (nxt != cur) memcpy(nxt, cur, outer * nc * inner * sizeof(double));
    } else {
      for (size_t o = 0; o < outer; ++o)
        for (int s = 0; s < ns; ++s)
          for (size_t in = 0; in < inner; ++in) {
            double acc = 0;
            for (int c = 0; c < nc; ++c) acc += g_c2s[ls[a]][s][c] * cur[(o * nc + c) * inner + in];
            nxt[(o * ns + s) * inner + in] = acc;
          }
    }
    dim[a] = ns;
    cur = nxt;
  }
}

// One shell quartet (segmented shells, l <= 4) into `out` (spherical,
// i fastest).  `work` holds the lane state.
struct Work {
  std::vector<v4> g;       // [axis][root][i][j][k][l]
  std::vector<v4> g2;      // VRR [n][m]
  std::vector<double> cart, tmp;
  Work() : g(3 * kMaxN * 5 * 5 * 5 * 5), g2(9 * 9 * 9), cart(50625), tmp(2 * 50625) {}
};

void quartet(const Shell* S, double* out, Work& W) {
  const int li = S[0].l, lj = S[1].l, lk = S[2].l, ll = S[3].l;
  const int nr = (li + lj + lk + ll) / 2 + 1;
  const bool ibase = li > lj, kbase = lk > ll;
  const int nmax = li + lj, mmax = lk + ll;
  const int NI = li + 1, NJ = lj + 1, NK = lk + 1, NL = ll + 1;
  const int nfi = ncart(li), nfj = ncart(lj), nfk = ncart(lk), nfl = ncart(ll);
  const int nf = nfi * nfj * nfk * nfl;
  double* cart = W.cart.data();
  for (int e = 0; e < nf; ++e) cart[e] = 0.0;
  const double common = M_PI * M_PI * M_PI * 2.0 / sqrt(M_PI) * common_fac_sp(li) *
                        common_fac_sp(lj) * common_fac_sp(lk) * common_fac_sp(ll);
  const double *ri = S[0].r, *rj = S[1].r, *rk = S[2].r, *rl = S[3].r;
  double rr_ij = 0, rr_kl = 0, rirj[3], rkrl[3];
  for (int d = 0; d < 3; ++d) {
    rr_ij += (ri[d] - rj[d]) * (ri[d] - rj[d]);
    rr_kl += (rk[d] - rl[d]) * (rk[d] - rl[d]);
    rirj[d] = ibase ? ri[d] - rj[d] : rj[d] - ri[d];
    rkrl[d] = kbase ? rk[d] - rl[d] : rl[d] - rk[d];
  }
  const double* Rb = ibase ? ri : rj;
  const double* Rq = kbase ? rk : rl;
  const int W2 = mmax + 1;
  auto G = [&](int ax, int r, int i, int j, int k, int l) -> v4& {
    return W.g[((((size_t)(ax * nr + r) * NI + i) * NJ + j) * NK + k) * NL + l];
  };
  // Cartesian components (libcint order)
  int cx[4][15], cy[4][15], cz[4][15];
  const int lv[4] = {li, lj, lk, ll};
  for (int s = 0; s < 4; ++s) {
    int n = 0;
    for (int x = lv[s]; x >= 0; --x)
      for (int y = lv[s] - x; y >= 0; --y, ++n) {
        cx[s][n] = x;
        cy[s][n] = y;
        cz[s][n] = lv[s] - x - y;
      }
  }
  for (int pi = 0; pi < S[0].np; ++pi)
    for (int pj = 0; pj < S[1].np; ++pj) {
      const double ai = S[0].a[pi], aj = S[1].a[pj], aij = ai + aj;
      const double fac_ij = exp(-ai * aj / aij * rr_ij);
      double P[3];
      for (int d = 0; d < 3; ++d) P[d] = (ai * ri[d] + aj * rj[d]) / aij;
      for (int pk = 0; pk < S[2].np; ++pk) {
        const double ak = S[2].a[pk];
        const double cijk = S[0].c[pi] * S[1].c[pj] * S[2].c[pk];
        for (int pl = 0; pl < S[3].np; pl += kLanes) {
          const int nl = S[3].np - pl < kLanes ? S[3].np - pl : kLanes;
          double alv[kLanes], wt[kLanes];
          for (int q = 0; q < kLanes; ++q) {
            alv[q] = q < nl ? S[3].a[pl + q] : 1.0;
            wt[q] = q < nl ? cijk * S[3].c[pl + q] : 0.0;
          }
          const v4 al = (v4){alv[0], alv[1], alv[2], alv[3]};
          const v4 weight = (v4){wt[0], wt[1], wt[2], wt[3]};
          const v4 akl = splat(ak) + al;
          const v4 a1 = splat(aij) * akl;
          const v4 a0 = a1 / (splat(aij) + akl);
          v4 Q[3], dx[3];
          v4 rr = splat(0.0);
          for (int d = 0; d < 3; ++d) {
            Q[d] = (splat(ak * rk[d]) + al * splat(rl[d])) / akl;
            dx[d] = splat(P[d]) - Q[d];
            rr += dx[d] * dx[d];
          }
          double fac1v[kLanes], xv[kLanes];
          for (int q = 0; q < kLanes; ++q) {
            const double fkl = exp(-ak * alv[q] / akl[q] * rr_kl);
            fac1v[q] = sqrt(a0[q] / (a1[q] * a1[q] * a1[q])) * common * fac_ij * fkl;
            xv[q] = a0[q] * rr[q];
          }
          double uu[kLanes][kMaxN], ww[kLanes][kMaxN];
          for (int q = 0; q < kLanes; ++q) rys_rule(nr, xv[q], uu[q], ww[q]);
          for (int r = 0; r < nr; ++r) {
            const v4 u = (v4){uu[0][r], uu[1][r], uu[2][r], uu[3][r]};
            const v4 wz = (v4){ww[0][r] * fac1v[0], ww[1][r] * fac1v[1], ww[2][r] * fac1v[2],
                               ww[3][r] * fac1v[3]};
            const v4 u2 = a0 * u;
            const v4 tmp4 = splat(0.5) / (u2 * (splat(aij) + akl) + a1);
            const v4 tmp5 = u2 * tmp4;
            const v4 tmp2 = splat(2.0) * tmp5 * akl;
            const v4 tmp3 = splat(2.0) * tmp5 * splat(aij);
            const v4 b00 = tmp5, b10 = tmp5 + tmp4 * akl, b01 = tmp5 + tmp4 * splat(aij);
            for (int ax = 0; ax < 3; ++ax) {
              const v4 c00 = splat(P[ax] - Rb[ax]) - tmp2 * dx[ax];
              const v4 c0p = (Q[ax] - splat(Rq[ax])) + tmp3 * dx[ax];
              v4* g2 = W.g2.data();
              g2[0] = ax == 2 ? wz : splat(1.0);
              if (nmax > 0) g2[W2] = c00 * g2[0];
              for (int n = 1; n < nmax; ++n)
                g2[(n + 1) * W2] = c00 * g2[n * W2] + splat((double)n) * b10 * g2[(n - 1) * W2];
              for (int m = 0; m < mmax; ++m)
                for (int n = 0; n <= nmax; ++n) {
                  v4 v = c0p * g2[n * W2 + m];
                  if (m > 0) v += splat((double)m) * b01 * g2[n * W2 + m - 1];
                  if (n > 0) v += splat((double)n) * b00 * g2[(n - 1) * W2 + m];
                  g2[n * W2 + m + 1] = v;
                }
              // bra transfer for every ket power, then the ket transfer
              const int A = ibase ? li : lj, B = ibase ? lj : li;
              const int C = kbase ? lk : ll, D = kbase ? ll : lk;
              const v4 drb = splat(rirj[ax]), drk = splat(rkrl[ax]);
              v4 h1[5][5][9];
              for (int m = 0; m <= mmax; ++m) {
                v4 t[9];
                for (int a = 0; a <= A + B; ++a) t[a] = g2[a * W2 + m];
                for (int b = 0; b <= B; ++b) {
                  if (b > 0)
                    for (int a = 0; a <= A + B - b; ++a) t[a] = t[a + 1] + drb * t[a];
                  for (int a = 0; a <= A; ++a) h1[ibase ? a : b][ibase ? b : a][m] = t[a];
                }
              }
              for (int i = 0; i < NI; ++i)
                for (int j = 0; j < NJ; ++j) {
                  v4 t[9];
                  for (int c = 0; c <= C + D; ++c) t[c] = h1[i][j][c];
                  for (int d = 0; d <= D; ++d) {
                    if (d > 0)
                      for (int c = 0; c <= C + D - d; ++c) t[c] = t[c + 1] + drk * t[c];
                    for (int c = 0; c <= C; ++c)
                      G(ax, r, i, j, kbase ? c : d, kbase ? d : c) = t[c];
                  }
                }
            }
          }
          // gout, weighted, reduced over the lanes
          int e = 0;
          for (int fl = 0; fl < nfl; ++fl)
            for (int fk = 0; fk < nfk; ++fk)
              for (int fj = 0; fj < nfj; ++fj)
                for (int fi = 0; fi < nfi; ++fi, ++e) {
                  v4 sum = splat(0.0);
                  for (int r = 0; r < nr; ++r)
                    sum += G(0, r, cx[0][fi], cx[1][fj], cx[2][fk], cx[3][fl]) *
                           G(1, r, cy[0][fi], cy[1][fj], cy[2][fk], cy[3][fl]) *
                           G(2, r, cz[0][fi], cz[1][fj], cz[2][fk], cz[3][fl]);
                  sum *= weight;
                  cart[e] += (sum[0] + sum[1]) + (sum[2] + sum[3]);
                }
        }
      }
    }
  cart_to_sph(cart, lv, out, W.tmp.data());
}

Shell get_shell(int sh, const int* atm, const int* bas, const double* env) {
  const int* b = bas + 8 * sh;
  return Shell{b[1], b[2], env + atm[6 * b[0] + 1], env + b[5], env + b[6]};
}

}  // namespace

extern "C" {

// Evaluate nq quartets (4 shells each, segmented, l <= 4).  out (nullable):
// block q at out + offsets[q], libcint layout.  Returns the sum of every
// element (so that the work cannot be elided) or NaN on unsupported input.
double cpus_int2e_sph(int nq, const int* shls, const long* offsets, double* out, const int* atm,
                      int natm, const int* bas, int nbas, const double* env, int nthreads) {
  (void)natm;
  std::call_once(g_once, build_tables);
  for (int s = 0; s < nbas; ++s)
    if (bas[8 * s + 3] != 1 || bas[8 * s + 1] > kMaxL) return NAN;
  // rows: runs of quartets sharing the bra pair
  std::vector<int> row;
  for (int q = 0; q < nq; ++q)
    if (q == 0 || shls[4 * q] != shls[4 * q - 4] || shls[4 * q + 1] != shls[4 * q - 3])
      row.push_back(q);
  row.push_back(nq);
  const int nrow = (int)row.size() - 1;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  double total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
  {
    Work W;
    std::vector<double> blk(6561);
#pragma omp for schedule(dynamic, 2)
    for (int k = 0; k < nrow; ++k)
      for (int q = row[k]; q < row[k + 1]; ++q) {
        Shell S[4];
        for (int a = 0; a < 4; ++a) S[a] = get_shell(shls[4 * q + a], atm, bas, env);
        double* o = out ? out + offsets[q] : blk.data();
        quartet(S, o, W);
        const int n = (2 * S[0].l + 1) * (2 * S[1].l + 1) * (2 * S[2].l + 1) * (2 * S[3].l + 1);
        for (int e = 0; e < n; ++e) total += o[e];
      }
  }
  return total;
}

// J_pq = sum_rs (pq|rs) D_rs and K_pr = sum_qs (pq|rs) D_qs (D: nao x nao
// row-major, taken as given) from nq unique quartets (each unordered shell
// quartet once): every distinct image of a quartet under the 8-fold
// permutational symmetry is added in full.  J, K: nao x nao, overwritten.
// Returns 0, or -1 on unsupported input.
int cpus_int2e_jk(int nq, const int* shls, const int* atm, int natm, const int* bas, int nbas,
                  const double* env, const double* D, double* J, double* K, int nthreads) {
  (void)natm;
  std::call_once(g_once, build_tables);
  for (int s = 0; s < nbas; ++s)
    if (bas[8 * s + 3] != 1 || bas[8 * s + 1] > kMaxL) return -1;
  std::vector<int> ao(nbas + 1, 0);
  for (int s = 0; s < nbas; ++s) ao[s + 1] = ao[s] + 2 * bas[8 * s + 1] + 1;
  const int nao = ao[nbas];
  const size_t nn = (size_t)nao * nao;
  for (size_t e = 0; e < nn; ++e) J[e] = K[e] = 0.0;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
  {
    Work W;
    std::vector<double> blk(6561), jl(nn, 0.0), kl(nn, 0.0);
#pragma omp for schedule(dynamic, 16)
    for (int q = 0; q < nq; ++q) {
      const int* sh = shls + 4 * q;
      Shell S[4];
      for (int a = 0; a < 4; ++a) S[a] = get_shell(sh[a], atm, bas, env);
      quartet(S, blk.data(), W);
      int n[4], st[4];
      for (int a = 0; a < 4; ++a) n[a] = 2 * S[a].l + 1;
      st[0] = 1;
      st[1] = n[0];
      st[2] = n[0] * n[1];
      st[3] = n[0] * n[1] * n[2];
      // the images (p q|r s) as permutations of the slots (0 1|2 3)
      static const int perm8[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                      {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};
      int seen[8][4], nseen = 0;
      for (int g = 0; g < 8; ++g) {
        int img[4];
        for (int a = 0; a < 4; ++a) img[a] = sh[perm8[g][a]];
        bool dup = false;
        for (int u = 0; u < nseen && !dup; ++u)
          dup = seen[u][0] == img[0] && seen[u][1] == img[1] && seen[u][2] == img[2] &&
                seen[u][3] == img[3];
        if (dup) continue;
        for (int a = 0; a < 4; ++a) seen[nseen][a] = img[a];
        ++nseen;
        const int* pm = perm8[g];
        const int np = n[pm[0]], nq_ = n[pm[1]], nr = n[pm[2]], ns = n[pm[3]];
        const int op = ao[img[0]], oq = ao[img[1]], orr = ao[img[2]], os = ao[img[3]];
        for (int ms = 0; ms < ns; ++ms)
          for (int mr = 0; mr < nr; ++mr)
            for (int mq = 0; mq < nq_; ++mq)
              for (int mp = 0; mp < np; ++mp) {
                const double v = blk[mp * st[pm[0]] + mq * st[pm[1]] + mr * st[pm[2]] +
                                     ms * st[pm[3]]];
                jl[(size_t)(op + mp) * nao + oq + mq] += v * D[(size_t)(orr + mr) * nao + os + ms];
                kl[(size_t)(op + mp) * nao + orr + mr] += v * D[(size_t)(oq + mq) * nao + os + ms];
              }
      }
    }
#pragma omp critical(cpus_jk)
    for (size_t e = 0; e < nn; ++e) {
      J[e] += jl[e];
      K[e] += kl[e];
    }
  }
  return 0;
}

}  // extern "C"
