// Host-side generator of the device Rys tables.
//
// The n-point Rys rule for argument x is the Gauss rule of the measure
// exp(-x t^2) dt on t in [0,1] expressed in s = t^2 (u = s/(1-s)); libcint
// evaluates it with fitted polynomials and moment methods
// (libcint-master/src/rys_roots.c:57-123).  Here it is computed once per
// table node in 80-bit long double by discretising the measure with
// Gauss-Legendre nodes and running the Stieltjes procedure (W. Gautschi,
// SIAM J. Sci. Stat. Comput. 3 (1982) 289) followed by the Golub-Welsch
// eigen-solve (implicit QL).  The device evaluates degree-11 Chebyshev fits
// of s_r(x) = t_r^2 = u_r/(1+u_r) (or of q_r = 1 - s_r where s_r > 1/2, stored
// negated) and w_r(x) on unit x-intervals below 35+5n (the LARGEX bound of
// rys_roots.c:67) and the asymptotic Hermite rule s = R/x, w = W/sqrt(x)
// above it.  Keeping s and q rather than u makes every per-root coefficient
// of the recurrences division-free without losing accuracy near s = 1.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "rys_table.h"

namespace cx {

namespace {

constexpr int kGlMax = 256;

struct GaussLegendre {
  long double t[kGlMax], w[kGlMax];
  int m = 0;
};

void gauss_legendre_01(int m, GaussLegendre& gl) {
  const long double pi = 3.14159265358979323846264338327950288L;
  gl.m = m;
  for (int k = 0; k < (m + 1) / 2; ++k) {
    long double z = cosl(pi * (k + 0.75L) / (m + 0.5L));
    long double p0 = 1, p1 = z, dp = 1;
    for (int it = 0; it < 100; ++it) {
      p0 = 1;
      p1 = z;
      for (int j = 2; j <= m; ++j) {
        long double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      dp = m * (z * p1 - p0) / (z * z - 1);
      long double dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-21L) break;
    }
    p0 = 1;
    p1 = z;
    for (int j = 2; j <= m; ++j) {
      long double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
      p0 = p1;
      p1 = p2;
    }
    dp = m * (z * p1 - p0) / (z * z - 1);
    long double wk = 2.0L / ((1 - z * z) * dp * dp);
    gl.t[k] = 0.5L * (1 - z);
    gl.w[k] = 0.5L * wk;
    gl.t[m - 1 - k] = 0.5L * (1 + z);
    gl.w[m - 1 - k] = 0.5L * wk;
  }
}

// Eigenvalues of the symmetric tridiagonal matrix (d, e) and the first
// components of the normalised eigenvectors, by implicit QL.
void ql_first_row(int n, long double* d, long double* e, long double* z0) {
  for (int i = 0; i < n; ++i) z0[i] = (i == 0) ? 1.0L : 0.0L;
  if (n == 1) return;
  e[n - 1] = 0;
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 300; ++iter) {
      int m = l;
      for (; m < n - 1; ++m) {
        long double dd = fabsl(d[m]) + fabsl(d[m + 1]);
        if (fabsl(e[m]) <= LDBL_EPSILON * dd) break;
      }
      if (m == l) break;
      long double g = (d[l + 1] - d[l]) / (2.0L * e[l]);
      long double r = hypotl(g, 1.0L);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      long double s = 1, c = 1, p = 0;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        long double f = s * e[i], b = c * e[i];
        r = hypotl(f, g);
        e[i + 1] = r;
        if (r == 0) {
          d[i + 1] -= p;
          e[m] = 0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0L * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        long double zf = z0[i + 1];
        z0[i + 1] = s * z0[i] + c * zf;
        z0[i] = c * z0[i] - s * zf;
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0;
    }
  }
}

// Gauss-Legendre rules of every order, each built once on first use (shared
// by the table builder's threads; no per-thread copies)
const GaussLegendre& gauss_legendre_rule(int m) {
  static GaussLegendre rules[kGlMax + 1];
  static std::once_flag built[kGlMax + 1];
  std::call_once(built[m], [m] { gauss_legendre_01(m, rules[m]); });
  return rules[m];
}

}  // namespace

void rys_rule_ld(int n, long double x, long double* s_out, long double* w_out) {
  const long double T = 100.0L + 4.0L * n;
  long double tmax = 1.0L;
  if (x * tmax * tmax > T) tmax = sqrtl(T / x);
  long double xeff = x * tmax * tmax;
  int m = 24 + 2 * n + (int)(1.5L * xeff);
  if (m > kGlMax) m = kGlMax;
  const GaussLegendre& gl = gauss_legendre_rule(m);
  long double s[kGlMax], W[kGlMax], p0[kGlMax], p1[kGlMax];
  long double mu0 = 0;
  for (int k = 0; k < m; ++k) {
    long double t = gl.t[k] * tmax;
    s[k] = t * t;
    W[k] = gl.w[k] * tmax * expl(-x * s[k]);
    p0[k] = 0;
    p1[k] = 1;
    mu0 += W[k];
  }
  long double alpha[kMaxRoots + 2], beta[kMaxRoots + 2];
  long double norm = mu0;
  for (int j = 0; j < n; ++j) {
    long double num = 0;
    for (int k = 0; k < m; ++k) num += W[k] * s[k] * p1[k] * p1[k];
    alpha[j] = num / norm;
    if (j == n - 1) break;
    long double bj = (j == 0) ? 0.0L : beta[j];
    long double nn = 0;
    for (int k = 0; k < m; ++k) {
      long double pn = (s[k] - alpha[j]) * p1[k] - bj * p0[k];
      p0[k] = p1[k];
      p1[k] = pn;
      nn += W[k] * pn * pn;
    }
    beta[j + 1] = nn / norm;
    norm = nn;
  }
  long double d[kMaxRoots + 2], e[kMaxRoots + 2], z0[kMaxRoots + 2];
  for (int j = 0; j < n; ++j) {
    d[j] = alpha[j];
    e[j] = (j + 1 < n) ? sqrtl(beta[j + 1]) : 0.0L;
  }
  ql_first_row(n, d, e, z0);
  int idx[kMaxRoots + 2];
  for (int i = 0; i < n; ++i) idx[i] = i;
  for (int i = 1; i < n; ++i) {
    int v = idx[i], j = i - 1;
    while (j >= 0 && d[idx[j]] > d[v]) {
      idx[j + 1] = idx[j];
      --j;
    }
    idx[j + 1] = v;
  }
  for (int i = 0; i < n; ++i) {
    s_out[i] = d[idx[i]];
    w_out[i] = mu0 * z0[idx[i]] * z0[idx[i]];
  }
}

void rys_table_build(std::vector<double>& coef, std::vector<double>& large,
                     int off[kMaxRoots + 1]) {
  // layout and offsets
  size_t total = 0;
  off[0] = 0;
  for (int n = 1; n <= kMaxRoots; ++n) {
    off[n] = (int)total;
    int nint = (int)rys_largex(n);
    total += (size_t)nint * 2 * n * kRysNc;
  }
  coef.assign(total, 0.0);
  large.assign((size_t)kMaxRoots * (kMaxRoots + 1), 0.0);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int n = 1; n <= kMaxRoots; ++n) {
    int nint = (int)rys_largex(n);
#pragma omp parallel for schedule(dynamic, 1)
    for (int iv = 0; iv < nint; ++iv) {
      long double su[kRysNc][kMaxRoots], sw[kRysNc][kMaxRoots];
      for (int k = 0; k < kRysNc; ++k) {
        long double tk = cosl(pi * (k + 0.5L) / kRysNc);
        long double x = iv + 0.5L * (tk + 1.0L);
        long double s[kMaxRoots + 2], w[kMaxRoots + 2];
        rys_rule_ld(n, x, s, w);
        for (int r = 0; r < n; ++r) {
          su[k][r] = s[r];  // t^2; u = s/(1-s) is formed only where needed
          sw[k][r] = w[r];
        }
      }
      double* blk = coef.data() + off[n] + (size_t)iv * 2 * n * kRysNc;
      for (int r = 0; r < n; ++r) {
        // Fit whichever of s and q = 1-s is the smaller on this interval, so
        // both come out with full relative accuracy (q matters near s = 1);
        // a fitted q is stored negated, which the evaluator detects by sign.
        long double smid = 0;
        for (int k = 0; k < kRysNc; ++k) smid += su[k][r];
        const bool fit_q = smid / kRysNc > 0.5L;
        for (int j = 0; j < kRysNc; ++j) {
          long double cu = 0, cw = 0;
          for (int k = 0; k < kRysNc; ++k) {
            long double tj = cosl(pi * j * (k + 0.5L) / kRysNc);
            cu += (fit_q ? -(1.0L - su[k][r]) : su[k][r]) * tj;
            cw += sw[k][r] * tj;
          }
          long double sc = (j == 0 ? 1.0L : 2.0L) / kRysNc;
          blk[(size_t)r * kRysNc + j] = (double)(cu * sc);
          blk[(size_t)(n + r) * kRysNc + j] = (double)(cw * sc);
        }
      }
    }
    // asymptotic rule: R = x s, W = w sqrt(x) at a very large x
    long double x0 = 1.0e8L;
    long double s[kMaxRoots + 2], w[kMaxRoots + 2];
    rys_rule_ld(n, x0, s, w);
    double* L = large.data() + (size_t)n * (n - 1);
    for (int r = 0; r < n; ++r) {
      L[r] = (double)(x0 * s[r]);
      L[n + r] = (double)(w[r] * sqrtl(x0));
    }
  }
}

}  // namespace cx
