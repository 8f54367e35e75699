// Host runtime of libcintx_hip: input validation, device-resident basis and
// primitive-pair data, class binning of shell tuples, kernel launches,
// hipGraph replay, and the three C-ABI layers declared in cintx_hip.h.
//
// Reference behaviour mirrored here:
//   * libcint raw entry points and their out == NULL cache query
//     (libcint-master/src/cint2e.c:794-869, cint3c2e.c:556-625);
//   * cintx-compat eval_raw validation and dims contract
//     (crates/cintx-compat/src/raw.rs:874-1125, layout.rs:25-134);
//   * cintx-capi status codes / thread-local last error
//     (crates/cintx-capi/src/shim.rs:313-450, errors.rs:9-245).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cintx_hip.h"
#include "common.h"
#include "launch.h"
#include "rys_table.h"

namespace cx {

hipError_t launch_rys_eval(const RysTable& tab, int n, const double* x, int64_t nx, double* u,
                           double* w, hipStream_t s);
hipError_t launch_generic(const EriArgs& a, const int* l, int deriv, int grid, hipStream_t s);
int64_t generic_scratch_per_block(const int* l, int deriv);
hipError_t mlp_train_device(const RysTable& tab, float* prm, int64_t nsamples, int epochs,
                            int batch, float lr, uint64_t seed, double* final_loss,
                            hipStream_t st);
hipError_t mlp_grad_device(const float* prm, const float* z, const int* nr, const float* tgt,
                           int batch, float* grad, float* loss, hipStream_t st);
hipError_t mlp_eval_device(const float* prm, const RysTable& tab, int n, const double* x,
                           int64_t nx, int steps, int mode, double* u, double* w, int* status,
                           double* perr, hipStream_t st);
int mlp_nparam();
hipError_t launch_fill(double* dst, int64_t n, double v, hipStream_t s);
hipError_t launch_jk_symmetrize(double* A, int n, double scale, hipStream_t s);
hipError_t launch_jk_sym_copy(const double* src, double* dst, int n, hipStream_t s);
hipError_t jk_scratch_address(double** ptr);
size_t jk_scratch_bytes();
hipError_t launch_jk_digest(const double* eri, int64_t out_base, int64_t block, const int2* q,
                            int64_t nq, const int2* pair_sh, const int* ao_off, const int* shl_l,
                            const double* D, double* J, double* K, int nao, hipStream_t s);
hipError_t launch_spinor(const SpinorJob* jobs, int64_t njobs, const double* cart, double* out,
                         const double* tab, const double* c2s, double* scratch, int64_t slice,
                         int grid, bool four, hipStream_t s);
static constexpr int kGenericGrid = 1024;

// ------------------------------------------------------------------
// Errors: fail-closed, typed, thread-local last error.
// ------------------------------------------------------------------
struct Failure {
  int status;
  std::string msg;
};

static thread_local int t_status = 0;
static thread_local std::string t_msg;

static void set_error(int st, const std::string& m) {
  t_status = st;
  t_msg = m;
}
static void clear_error() {
  t_status = 0;
  t_msg.clear();
}

[[noreturn]] static void fail(int st, const std::string& m) { throw Failure{st, m}; }

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess)                                                                  \
      fail(CINTX_STATUS_EXECUTION_FAILED, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    clear_error();
    return CINTX_STATUS_SUCCESS;
  } catch (const Failure& e) {
    set_error(e.status, e.msg);
    return e.status;
  } catch (const std::bad_alloc&) {
    set_error(CINTX_STATUS_ALLOCATION_FAILED, "host allocation failed");
    return CINTX_STATUS_ALLOCATION_FAILED;
  } catch (...) {
    set_error(CINTX_STATUS_PANIC, "unexpected exception");
    return CINTX_STATUS_PANIC;
  }
}

// ------------------------------------------------------------------
// Families
// ------------------------------------------------------------------
// Canonical slots (0..3) are the kernel's (i j|k l); canon_perm says which
// argument shell sits in each.  `three`/`two` describe the canonical ket/bra,
// `deriv` puts nabla on canonical slot 0, `dcenter` names the argument that
// carries it in the libcint call (-1: none).
struct FamilyInfo {
  bool three;  // the canonical ket is one Gaussian
  int deriv;
  int ncomp;
  int arity;
  const char* name;
  bool two;     // the canonical bra is one Gaussian
  int dcenter;  // argument shell under nabla
  bool cart = false;  // Cartesian output (no cart -> sph)
  // complex two-component output: the Cartesian blocks (cart) go through
  // the spinor transform (slot k of a three-centre family stays spherical)
  bool spinor = false;
};
static constexpr int kNumFamilies = 11;
static const FamilyInfo kFam[kNumFamilies] = {
    {false, 0, 1, 4, "int2e_sph", false, -1},
    {false, 1, 3, 4, "int2e_ip1_sph", false, 0},
    {true, 0, 1, 3, "int3c2e_sph", false, -1},
    {true, 1, 3, 3, "int3c2e_ip1_sph", false, 0},
    {true, 0, 1, 2, "int2c2e_sph", true, -1},
    // (ij|nabla k l) = (nabla k l|ij): the ip1 kernels with (k l) as the bra
    {false, 1, 3, 4, "int2e_ip2_sph", false, 2},
    // (ij|nabla k): bra = k alone, ket = the (i j) pair
    {false, 1, 3, 3, "int3c2e_ip2_sph", true, 2},
    {false, 0, 1, 4, "int2e_cart", false, -1, true},
    {true, 0, 1, 3, "int3c2e_cart", false, -1, true},
    {false, 0, 1, 4, "int2e_spinor", false, -1, true, true},
    {true, 0, 1, 3, "int3c2e_spinor", false, -1, true, true}};

// functions per shell in a family's canonical (kernel) blocks
static int nfun(const FamilyInfo& f, int l) { return f.cart ? ncart(l) : nsph(l); }
// functions of argument slot a in a family's output (libcint's counts,
// CINTcgto_spinor for the spinor slots)
static int out_fun(const FamilyInfo& f, int a, int l, int kappa) {
  if (!f.spinor) return nfun(f, l);
  return (f.three && a == 2) ? nsph(l) : spinor_len(l, kappa);
}
// doubles per output element
static int out_width(const FamilyInfo& f) { return f.spinor ? 2 : 1; }

static constexpr double kExpCutoff = 60.0;     // cint_config.h.in EXPCUTOFF
static constexpr double kMinExpCutoff = 40.0;  // cint_config.h.in MIN_EXPCUTOFF

static double common_fac_sp(int l) {  // g1e.c:566-573
  if (l == 0) return 0.282094791773878143;
  if (l == 1) return 0.488602511902919921;
  return 1.0;
}

// ------------------------------------------------------------------
// Real solid harmonics table (same convention as cart2sph.c:21):
// coefficient of x^a y^b z^c (libcint Cartesian order) in r^l Y_lm with
// unit-sphere-normalised real Y_lm, rows m = -l..l; s and p are identity.
// Closed form: Helgaker, Jorgensen, Olsen, Molecular Electronic-Structure
// Theory, eqs. 6.4.47-6.4.50.
// ------------------------------------------------------------------
static void build_c2s_table(std::vector<double>& tab) {
  tab.assign((size_t)(kC2sL + 1) * kC2sM * kC2sC, 0.0);
  auto at = [&](int l, int m, int c) -> double& {
    return tab[((size_t)l * kC2sM + m) * kC2sC + c];
  };
  at(0, 0, 0) = 1.0;
  for (int i = 0; i < 3; ++i) at(1, i, i) = 1.0;
  auto binom = [](int n, int k) -> long double {
    if (k < 0 || k > n) return 0.0L;
    long double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
  };
  auto fact = [](int n) -> long double {
    long double r = 1;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
  };
  auto cidx = [](int l, int lx, int ly) {
    int pos = 0;
    for (int ax = l; ax > lx; --ax) pos += l - ax + 1;
    return pos + (l - lx - ly);
  };
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int l = 2; l <= kC2sL; ++l) {
    long double ynorm = sqrtl((2 * l + 1) / (4 * pi));
    for (int m = -l; m <= l; ++m) {
      int am = std::abs(m);
      int vm2 = m < 0 ? 1 : 0;
      long double nlm = 1.0L / (powl(2.0L, am) * fact(l)) *
                        sqrtl(2.0L * fact(l + am) * fact(l - am) / (m == 0 ? 2.0L : 1.0L));
      for (int t = 0; t <= (l - am) / 2; ++t)
        for (int uu = 0; uu <= t; ++uu) {
          int vvmax = am - vm2 >= 0 ? (am - vm2) / 2 : -1;
          for (int vv = 0; vv <= vvmax; ++vv) {
            int v2 = 2 * vv + vm2;
            long double c = (((t + vv) & 1) ? -1.0L : 1.0L) * powl(0.25L, t) * binom(l, t) *
                            binom(l - t, am + t) * binom(t, uu) * binom(am, v2);
            int px = 2 * t + am - 2 * uu - v2, py = 2 * uu + v2;
            if (px < 0) continue;
            at(l, m + l, cidx(l, px, py)) += (double)(ynorm * nlm * c);
          }
        }
    }
  }
}

// ------------------------------------------------------------------
// Cartesian -> spinor coupling table (layout in common.h): the two-component
// spinors |l, j, m_j> built from the complex spherical harmonics (Condon-
// Shortley phase) and the Clebsch-Gordan coefficients <l m 1/2 s|j m_j>:
//   j = l+1/2: sqrt((2l+1+2m_j)/(2(2l+1))) Y_l^{m_j-1/2} alpha
//            + sqrt((2l+1-2m_j)/(2(2l+1))) Y_l^{m_j+1/2} beta
//   j = l-1/2: -sqrt((2l+1-2m_j)/(2(2l+1))) Y_l^{m_j-1/2} alpha
//            + sqrt((2l+1+2m_j)/(2(2l+1))) Y_l^{m_j+1/2} beta
// with Y_l^{+m} = (-1)^m (S_lm + i S_l,-m)/sqrt(2), Y_l^{-m} =
// (S_lm - i S_l,-m)/sqrt(2), Y_l^0 = S_l0 over the real solid harmonics S
// of the c2s table (p: x, y, z are S_11, S_1-1, S_10), so the coefficients
// carry libcint's Cartesian normalisation.  (Same convention as libcint's
// g_trans_cart2jR/I, cart2sph.c:3561, scripts/cart2sph.py:187-265.)
// `tab` receives 2 kSpinorTab doubles (re, im interleaved).
// ------------------------------------------------------------------
static void build_spinor_table(const std::vector<double>& c2s, std::vector<double>& tab) {
  tab.assign((size_t)2 * kSpinorTab, 0.0);
  for (int l = 0; l <= kMaxL; ++l) {
    const int nf = ncart(l);
    // complex harmonics Y[m + l][cart] (re, im)
    std::vector<long double> yr((size_t)(2 * l + 1) * nf, 0.0L), yi(yr.size(), 0.0L);
    auto real_row = [&](int m) -> const double* {
      int row = m + l;
      if (l == 1) row = m == 1 ? 0 : (m == -1 ? 1 : 2);  // p rows are x, y, z
      return c2s.data() + ((size_t)l * kC2sM + row) * kC2sC;
    };
    if (l == 0) {
      yr[0] = 1.0L;
    } else {
      for (int c = 0; c < nf; ++c) yr[(size_t)l * nf + c] = real_row(0)[c];
      const long double h = sqrtl(0.5L);
      for (int m = 1; m <= l; ++m) {
        const long double sg = (m & 1) ? -1.0L : 1.0L;
        for (int c = 0; c < nf; ++c) {
          const long double cr = real_row(m)[c], ci = real_row(-m)[c];
          yr[(size_t)(l + m) * nf + c] = sg * h * cr;
          yi[(size_t)(l + m) * nf + c] = sg * h * ci;
          yr[(size_t)(l - m) * nf + c] = h * cr;
          yi[(size_t)(l - m) * nf + c] = -h * ci;
        }
      }
    }
    double* out = tab.data() + 2 * (size_t)spinor_tab_off(l);
    int row = 0;
    auto put = [&](int twomj, long double ca, long double cb) {
      const int ma = (twomj - 1) / 2, mb = (twomj + 1) / 2;
      double* r = out + (size_t)row * 2 * nf * 2;
      for (int c = 0; c < nf; ++c) {
        if (ma >= -l && ma <= l) {
          r[2 * c] = (double)(ca * yr[(size_t)(ma + l) * nf + c]);
          r[2 * c + 1] = (double)(ca * yi[(size_t)(ma + l) * nf + c]);
        }
        if (mb >= -l && mb <= l) {
          r[2 * (nf + c)] = (double)(cb * yr[(size_t)(mb + l) * nf + c]);
          r[2 * (nf + c) + 1] = (double)(cb * yi[(size_t)(mb + l) * nf + c]);
        }
      }
      ++row;
    };
    const long double n2 = 2.0L * (2 * l + 1);
    for (int t = -(2 * l - 1); t <= 2 * l - 1; t += 2)  // j = l - 1/2
      put(t, -sqrtl((2 * l + 1 - t) / n2), sqrtl((2 * l + 1 + t) / n2));
    for (int t = -(2 * l + 1); t <= 2 * l + 1; t += 2)  // j = l + 1/2
      put(t, sqrtl((2 * l + 1 + t) / n2), sqrtl((2 * l + 1 - t) / n2));
  }
}

// ------------------------------------------------------------------
// Device context (one per process; current device at first use).
// ------------------------------------------------------------------
constexpr int kChains = 8;  // parallel chains of a captured plan graph

struct DevCtx {
  int device = -1;
  hipStream_t stream = nullptr;
  hipStream_t branch[kChains - 1] = {};  // further chains of a captured plan graph
  RysTable rys{};
  double* d_rys = nullptr;
  double* d_large = nullptr;
  double* d_c2s = nullptr;
  double* d_spinor = nullptr;  // Cartesian -> spinor table (complex)
  Registry* reg = nullptr;
  double* d_jk_scratch = nullptr;  // the static scratch of cintx_plan_jk_noalloc
  hipEvent_t jk_ev[kChains] = {};  // fork/join events of cintx_plan_jk_noalloc's chains
  // autotuned class-kernel variant per (deriv, form, class); -1: untuned
  signed char tuned[2][kForms][kMaxL + 1][kMaxL + 1][kMaxL + 1][kMaxL + 1];
};

static std::atomic<DevCtx*> g_ctx{nullptr};

// The process's device context, created on first use on the device current
// then; later calls read it without a lock.
static DevCtx& ctx() {
  if (DevCtx* c = g_ctx.load(std::memory_order_acquire)) return *c;
  static std::mutex init_mu;
  std::lock_guard<std::mutex> lk(init_mu);
  if (DevCtx* c = g_ctx.load(std::memory_order_relaxed)) return *c;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    fail(CINTX_STATUS_EXECUTION_FAILED, "no HIP device visible");
  DevCtx* c = new DevCtx();
  HIPCHK(hipGetDevice(&c->device));
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int b = 0; b < kChains - 1; ++b)
    HIPCHK(hipStreamCreateWithFlags(&c->branch[b], hipStreamNonBlocking));
  std::vector<double> coef, large;
  rys_table_build(coef, large, c->rys.off);
  HIPCHK(hipMalloc(&c->d_rys, coef.size() * sizeof(double)));
  HIPCHK(hipMalloc(&c->d_large, large.size() * sizeof(double)));
  HIPCHK(hipMemcpy(c->d_rys, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_large, large.data(), large.size() * sizeof(double),
                   hipMemcpyHostToDevice));
  c->rys.coef = c->d_rys;
  c->rys.large = c->d_large;
  std::vector<double> c2s;
  build_c2s_table(c2s);
  HIPCHK(hipMalloc(&c->d_c2s, c2s.size() * sizeof(double)));
  HIPCHK(hipMemcpy(c->d_c2s, c2s.data(), c2s.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> sp;
  build_spinor_table(c2s, sp);
  HIPCHK(hipMalloc(&c->d_spinor, sp.size() * sizeof(double)));
  HIPCHK(hipMemcpy(c->d_spinor, sp.data(), sp.size() * sizeof(double), hipMemcpyHostToDevice));
  // address of the allocation-free J/K's static scratch on this device, and
  // the events its chains fork and join on (made here, so that the call
  // itself creates nothing)
  HIPCHK(jk_scratch_address(&c->d_jk_scratch));
  for (int k = 0; k < kChains; ++k)
    HIPCHK(hipEventCreateWithFlags(&c->jk_ev[k], hipEventDisableTiming));
  c->reg = new Registry();
  std::memset(c->reg, 0, sizeof(Registry));
  register_class_kernels(*c->reg);
  std::memset(c->tuned, -1, sizeof c->tuned);
  g_ctx.store(c, std::memory_order_release);
  return *c;
}

// ------------------------------------------------------------------
// Basis ingestion (libcint slots, cint.h.in:49-67).  General contractions
// are split into virtual shells, one per contraction column.
// ------------------------------------------------------------------
struct VShell {
  int l, nprim, orig, ctr;
  double r[3];
  int off;  // into exps / coefs
};

struct HostBasis {
  std::vector<VShell> v;
  std::vector<double> exps, coefs;
  std::vector<int> first_v;  // original shell -> first virtual shell
  std::vector<int> l, nctr, nprim, kappa;
  int nbas = 0;
};

static HostBasis ingest(const int* atm, int natm, const int* bas, int nbas, const double* env,
                        size_t env_len /* 0: unknown */) {
  if (!atm) fail(CINTX_STATUS_NULL_POINTER, "atm is null");
  if (!bas) fail(CINTX_STATUS_NULL_POINTER, "bas is null");
  if (!env) fail(CINTX_STATUS_NULL_POINTER, "env is null");
  if (natm <= 0) fail(CINTX_STATUS_INVALID_INPUT, "natm must be positive");
  if (nbas <= 0) fail(CINTX_STATUS_INVALID_INPUT, "nbas must be positive");
  HostBasis B;
  B.nbas = nbas;
  B.first_v.resize(nbas);
  B.l.resize(nbas);
  B.nctr.resize(nbas);
  B.nprim.resize(nbas);
  B.kappa.resize(nbas);
  auto env_ok = [&](long p, long n) {
    return p >= 0 && (env_len == 0 || (size_t)(p + n) <= env_len);
  };
  for (int s = 0; s < nbas; ++s) {
    const int* b = bas + 8 * (size_t)s;
    int ia = b[0], l = b[1], np = b[2], nc = b[3], pe = b[5], pc = b[6];
    if (ia < 0 || ia >= natm)
      fail(CINTX_STATUS_INVALID_INPUT, "bas[" + std::to_string(s) + "] ATOM_OF out of range");
    if (l < 0) fail(CINTX_STATUS_INVALID_INPUT, "negative angular momentum");
    if (l > kMaxL)
      fail(CINTX_STATUS_UNSUPPORTED_API,
           "angular momentum " + std::to_string(l) + " above the supported maximum 4");
    if (np <= 0 || nc <= 0) fail(CINTX_STATUS_INVALID_INPUT, "NPRIM_OF/NCTR_OF must be positive");
    if (np > 64) fail(CINTX_STATUS_UNSUPPORTED_API, "more than 64 primitives in a shell");
    const int* a = atm + 6 * (size_t)ia;
    if (!env_ok(a[1], 3) || !env_ok(pe, np) || !env_ok(pc, (long)np * nc))
      fail(CINTX_STATUS_INVALID_INPUT, "env offset out of range for shell " + std::to_string(s));
    B.first_v[s] = (int)B.v.size();
    B.l[s] = l;
    B.nctr[s] = nc;
    B.nprim[s] = np;
    B.kappa[s] = b[4];
    for (int c = 0; c < nc; ++c) {
      VShell vs;
      vs.l = l;
      vs.nprim = np;
      vs.orig = s;
      vs.ctr = c;
      for (int d = 0; d < 3; ++d) vs.r[d] = env[a[1] + d];
      vs.off = (int)B.exps.size();
      for (int p = 0; p < np; ++p) {
        B.exps.push_back(env[pe + p]);
        B.coefs.push_back(env[pc + (size_t)c * np + p]);  // column-major [ctr][prim]
      }
      B.v.push_back(vs);
    }
  }
  return B;
}

// ------------------------------------------------------------------
// Primitive-pair data (CINTset_pairdata, optimizer.c:288-342; screening
// keys cint2e.c:93-123, 199-236).  cce stored without the l-dependent
// l_ij*log(|R_ij|+1) term, which the kernel subtracts per class.  The
// product centre is kept as the weights w_i = a_i/a_ij, w_j = a_j/a_ij
// instead of P itself: the kernels form P - R_i = w_j (R_j - R_i) and
// P - Q from coordinate differences, so no absolute coordinate enters a
// difference (one-centre terms vanish exactly, as they do analytically).
// ------------------------------------------------------------------
struct PairSet {
  std::vector<double> aij, wi, wj, kfac, ai, cce, hinv;
  std::vector<int> off, np;
  std::vector<int2> shells;  // original shells (a, b) of each pair (b = -1: pseudo)
  std::vector<double> geo;
  std::unordered_map<uint64_t, int> index;
  std::vector<int> dense;  // (va, vb) -> id when nv^2 is small enough
  int nv = 0;

  void init(const HostBasis& B) {
    nv = (int)B.v.size();
    if ((int64_t)nv * (nv + 1) <= ((int64_t)1 << 26)) dense.assign((size_t)nv * (nv + 1), -1);
  }
  int* slot(int va, int vb) {  // vb == nv: pseudo pair of va
    if (dense.empty()) return nullptr;
    return &dense[(size_t)va * (nv + 1) + vb];
  }
  // id of a pair added before (vb == nv: the pseudo pair of va), -1 if none
  int find(int va, int vb) const {
    if (!dense.empty()) return dense[(size_t)va * (nv + 1) + vb];
    const uint64_t key = vb == nv ? (((uint64_t)0xffffffffu << 32) | (uint32_t)va)
                                  : (((uint64_t)(uint32_t)va << 32) | (uint32_t)vb);
    auto it = index.find(key);
    return it == index.end() ? -1 : it->second;
  }

  int add(const HostBasis& B, int va, int vb) {
    int* sl = slot(va, vb);
    if (sl && *sl >= 0) return *sl;
    uint64_t key = ((uint64_t)(uint32_t)va << 32) | (uint32_t)vb;
    if (!sl) {
      auto it = index.find(key);
      if (it != index.end()) return it->second;
    }
    const VShell& A = B.v[va];
    const VShell& Bs = B.v[vb];
    int id = (int)off.size();
    off.push_back((int)aij.size());
    np.push_back(A.nprim * Bs.nprim);
    shells.push_back(make_int2(A.orig, Bs.orig));
    double rr = 0;
    for (int d = 0; d < 3; ++d) rr += (A.r[d] - Bs.r[d]) * (A.r[d] - Bs.r[d]);
    double aij_last = B.exps[A.off + A.nprim - 1] + B.exps[Bs.off + Bs.nprim - 1];
    double log_rr0 = 1.7 - 1.5 * std::log(aij_last);
    // primitive pairs in ascending cce (libcint's jp-major order otherwise,
    // which only fixes the summation order): the survivors of the screen
    // cce_ij + cce_kl <= cutoff are then a prefix of each list, which the
    // one-lane kernels walk without visiting screened primitive tuples
    std::vector<std::pair<double, int>> ord;
    for (int jp = 0; jp < Bs.nprim; ++jp)
      for (int ip = 0; ip < A.nprim; ++ip) {
        double a_i = B.exps[A.off + ip], a_j = B.exps[Bs.off + jp];
        double ci = B.coefs[A.off + ip], cj = B.coefs[Bs.off + jp];
        double inv = 1.0 / (a_i + a_j);
        double eij = rr * a_i * a_j * inv;
        ord.emplace_back(eij - log_rr0 - std::log(std::fabs(ci)) - std::log(std::fabs(cj)),
                         jp * A.nprim + ip);
      }
    std::stable_sort(ord.begin(), ord.end(),
                     [](const std::pair<double, int>& x, const std::pair<double, int>& y) {
                       return x.first < y.first;
                     });
    for (const auto& o : ord) {
      const int jp = o.second / A.nprim, ip = o.second % A.nprim;
      double a_i = B.exps[A.off + ip], a_j = B.exps[Bs.off + jp];
      double ci = B.coefs[A.off + ip], cj = B.coefs[Bs.off + jp];
      double inv = 1.0 / (a_i + a_j);
      double eij = rr * a_i * a_j * inv;
      aij.push_back(a_i + a_j);
      hinv.push_back(0.5 * inv);
      wi.push_back(a_i * inv);
      wj.push_back(a_j * inv);
      kfac.push_back(std::exp(-eij) * ci * cj);
      ai.push_back(a_i);
      cce.push_back(o.first);
    }
    for (int d = 0; d < 3; ++d) geo.push_back(A.r[d]);
    for (int d = 0; d < 3; ++d) geo.push_back(Bs.r[d]);
    geo.push_back(std::log(std::sqrt(rr) + 1.0));
    geo.push_back(0.0);
    if (sl)
      *sl = id;
    else
      index.emplace(key, id);
    return id;
  }

  // int3c2e ket: shell k alone (the l slot of g3c2e.c:68-69 carries k, al=0).
  int add_pseudo(const HostBasis& B, int vk) {
    int* sl = slot(vk, nv);
    if (sl && *sl >= 0) return *sl;
    uint64_t key = ((uint64_t)0xffffffffu << 32) | (uint32_t)vk;
    if (!sl) {
      auto it = index.find(key);
      if (it != index.end()) return it->second;
    }
    const VShell& K = B.v[vk];
    int id = (int)off.size();
    off.push_back((int)aij.size());
    np.push_back(K.nprim);
    shells.push_back(make_int2(K.orig, -1));
    for (int kp = 0; kp < K.nprim; ++kp) {
      aij.push_back(B.exps[K.off + kp]);
      hinv.push_back(0.5 / B.exps[K.off + kp]);
      wi.push_back(1.0);
      wj.push_back(0.0);
      kfac.push_back(B.coefs[K.off + kp]);
      ai.push_back(B.exps[K.off + kp]);
      cce.push_back(0.0);
    }
    for (int d = 0; d < 3; ++d) geo.push_back(K.r[d]);
    for (int d = 0; d < 3; ++d) geo.push_back(K.r[d]);
    geo.push_back(0.0);
    geo.push_back(0.0);
    if (sl)
      *sl = id;
    else
      index.emplace(key, id);
    return id;
  }
};

// ------------------------------------------------------------------
// Plans
// ------------------------------------------------------------------
struct Bin {
  int l[4];
  int prec;
  double last_ms = 0;           // last timed execute
  double prim_tuples = 0;       // from the last work count
  double flops = 0;
  bool generic = false;  // no compile-time instantiation: runtime-class kernel
  int variant = 0;       // class-kernel variant (launch shape) in use
  std::vector<int2> q;
  int64_t out_base = 0;  // element offset of the bin's first block
  int64_t block = 0;     // elements per block (NS * ncomp)
  // device pointer (inside the plan blob)
  int2* d_q = nullptr;
};

struct VQ {  // one virtual (segmented) tuple of a caller tuple
  int bin;
  int64_t idx;
  uint8_t cc[4];  // contraction index per canonical axis
};

struct TupleRec {
  int shl[4];
  int perm[4];  // canonical axis c -> caller axis perm[c]
  int64_t first_vq, nvq;
  int64_t nat_off;  // natural-layout offset in the host output
  int64_t nat_size;
};

}  // namespace cx

struct cintx_plan {
  int family = 0;
  int64_t ntuples_fp32 = 0;
  double expcutoff = 60.0;
  cx::HostBasis basis;
  cx::PairSet pairs;
  std::vector<cx::Bin> bins;
  std::vector<cx::VQ> vqs;
  std::vector<cx::TupleRec> tuples;
  int64_t total = 0;  // output elements
  char* blob = nullptr;
  size_t blob_bytes = 0;
  double* d_out = nullptr;
  double* d_scratch = nullptr;  // generic-kernel workspace: one slice per chain
  int64_t scratch_stride = 0;   // doubles per chain's slice
  int scratch_chains = 1;       // chains that may run runtime-class bins
  unsigned long long* d_counters = nullptr;  // 2 per bin
  // J/K digestion tables: AO offset and l of every shell (spherical,
  // segmented shells), original shells of every pair
  int* d_ao_off = nullptr;
  int* d_shl_l = nullptr;
  int2* d_pair_sh = nullptr;
  int nao = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;  // ev0/ev1 bracket an execute whose time sync may read
  double last_ms = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double* graph_out = nullptr;
  // spinor families: the bins' Cartesian blocks and the transform jobs
  // (one per contracted block) that turn them into d_out's layout
  int64_t cart_total = 0;  // doubles of the Cartesian blocks
  double* d_cart = nullptr;
  cx::SpinorJob* d_jobs = nullptr;
  int64_t njobs = 0;
  double* d_spin_scratch = nullptr;
  int64_t spin_slice = 0;  // complex elements per workgroup
  int spin_grid = 0;
  // device pointers of the pair SoA
  double *d_geo = nullptr, *d_rec = nullptr;
  int* d_poff = nullptr;
  int* d_pnp = nullptr;
};

namespace cx {

static void canon_perm(int fam, const int* l, int p[4]) {
  for (int c = 0; c < 4; ++c) p[c] = c;
  const FamilyInfo& f = kFam[fam];
  if (f.arity == 2) {
    // (i|k): i in the i slot, k in the k slot, the j and l slots empty
    // (g2c2e.c:76-84 sets aj = al = 0 and uses the 2e recurrences)
    p[1] = 2;
    p[2] = 1;
    return;
  }
  if (f.dcenter == 2) {
    // nabla on k (grad2.c:99-101, autocode/int3c2e.c:98-100): the (k l)
    // side becomes the canonical bra, (i j) the ket, swapped to li >= lj
    p[0] = 2;
    p[1] = 3;
    p[2] = 0;
    p[3] = 1;
    if (l[0] < l[1]) std::swap(p[2], p[3]);
    return;
  }
  if (!f.deriv && l[0] < l[1]) std::swap(p[0], p[1]);
  if (!f.three && l[2] < l[3]) std::swap(p[2], p[3]);
  if (!f.three && !f.deriv) {
    int b0 = l[p[0]], b1 = l[p[1]], k0 = l[p[2]], k1 = l[p[3]];
    if (b0 < k0 || (b0 == k0 && b1 < k1)) {
      std::swap(p[0], p[2]);
      std::swap(p[1], p[3]);
    }
  }
}

// output extent of argument slot a of a tuple (contractions included)
static int out_extent(const FamilyInfo& f, const HostBasis& B, int a, int sh) {
  return out_fun(f, a, B.l[sh], B.kappa[sh]) * B.nctr[sh];
}

// doubles of a tuple's output block
static int64_t nat_size(int fam, const HostBasis& B, const int* shl) {
  const FamilyInfo& f = kFam[fam];
  int64_t n = (int64_t)f.ncomp * out_width(f);
  for (int a = 0; a < f.arity; ++a) n *= out_extent(f, B, a, shl[a]);
  return n;
}

static double fam_expcutoff(int fam, const double* env) {
  double e0 = env[0];
  if (e0 == 0) return kExpCutoff;
  double c = std::max(kMinExpCutoff, e0);
  if (kFam[fam].arity == 4) c += 1.0;  // g2e.c:57-62 vs g3c2e.c:56-60, g2c2e.c:46-50
  return c;
}

// Add nt caller tuples (4 ints each) in order; precision `prec` for all, or
// per tuple from prec_of when that is given.
static void plan_add_tuples(cintx_plan* P, const int* tuples, int64_t nt, int prec,
                            const uint8_t* prec_of = nullptr) {
  const FamilyInfo& f = kFam[P->family];
  if ((prec != 0 || prec_of) && f.cart)
    fail(CINTX_STATUS_INVALID_INPUT, "Cartesian families run in FP64 only");
  const HostBasis& B = P->basis;
  std::unordered_map<uint64_t, int> bin_of;
  for (int64_t t = 0; t < nt; ++t) {
    TupleRec tr;
    int l[4] = {0, 0, 0, 0};
    for (int a = 0; a < 4; ++a) tr.shl[a] = 0;
    for (int a = 0; a < f.arity; ++a) {
      int s = tuples[4 * t + a];
      if (s < 0 || s >= B.nbas) fail(CINTX_STATUS_INVALID_INPUT, "shell index out of range");
      tr.shl[a] = s;
      l[a] = B.l[s];
    }
    canon_perm(P->family, l, tr.perm);
    int cl[4], cn[4], cs[4];
    for (int c = 0; c < 4; ++c) {
      int a = tr.perm[c];
      bool dummy = a >= f.arity;
      cl[c] = dummy ? 0 : l[a];
      cn[c] = dummy ? 1 : B.nctr[tr.shl[a]];
      cs[c] = dummy ? -1 : tr.shl[a];
    }
    tr.first_vq = (int64_t)P->vqs.size();
    tr.nvq = (int64_t)cn[0] * cn[1] * cn[2] * cn[3];
    tr.nat_off = P->total;
    tr.nat_size = nat_size(P->family, B, tr.shl);
    P->total += tr.nat_size;
    const int tprec = prec_of ? prec_of[t] : prec;
    uint64_t key = 0;
    for (int c = 0; c < 4; ++c) key = key * 8 + cl[c];
    key = key * 2 + tprec;
    int bi;
    auto it = bin_of.find(key);
    if (it == bin_of.end()) {
      Bin b;
      for (int c = 0; c < 4; ++c) b.l[c] = cl[c];
      b.prec = tprec;
      b.block = (int64_t)f.ncomp * nfun(f, cl[0]) * nfun(f, cl[1]) * nfun(f, cl[2]) * nfun(f, cl[3]);
      bi = (int)P->bins.size();
      P->bins.push_back(b);
      bin_of.emplace(key, bi);
    } else {
      bi = it->second;
    }
    for (int c3 = 0; c3 < cn[3]; ++c3)
      for (int c2 = 0; c2 < cn[2]; ++c2)
        for (int c1 = 0; c1 < cn[1]; ++c1)
          for (int c0 = 0; c0 < cn[0]; ++c0) {
            int v0 = B.first_v[cs[0]] + c0;
            int bra = f.two ? P->pairs.add_pseudo(B, v0)
                            : P->pairs.add(B, v0, B.first_v[cs[1]] + c1);
            int ket;
            if (f.three)
              ket = P->pairs.add_pseudo(B, B.first_v[cs[2]] + c2);
            else
              ket = P->pairs.add(B, B.first_v[cs[2]] + c2, B.first_v[cs[3]] + c3);
            VQ vq;
            vq.bin = bi;
            vq.idx = (int64_t)P->bins[bi].q.size();
            vq.cc[0] = (uint8_t)c0;
            vq.cc[1] = (uint8_t)c1;
            vq.cc[2] = (uint8_t)c2;
            vq.cc[3] = (uint8_t)c3;
            P->bins[bi].q.push_back(make_int2(bra, ket));
            P->vqs.push_back(vq);
          }
    P->tuples.push_back(tr);
  }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Primitive tuples of (bra, ket) that pass the screen of a class with
// l-sums lij, lkl (cce_ij + cce_kl <= cutoff, cce_ij < cutoff, each with
// its l log(|R|+1) term; no ket term for one-Gaussian kets).  The
// primitive pairs are stored in ascending cce, so the survivors against
// each ket primitive are a prefix of the bra list (two pointers).
static int64_t surviving_prims(const PairSet& S, int bra, int ket, int lij, int lkl, bool three,
                               double cutoff) {
  const double* cb = S.cce.data() + S.off[bra];
  const double* ck = S.cce.data() + S.off[ket];
  const double tb = lij * S.geo[8 * (size_t)bra + 6], tk = lkl * S.geo[8 * (size_t)ket + 6];
  int n = S.np[bra];
  int64_t total = 0;
  for (int kp = 0; kp < S.np[ket]; ++kp) {
    const double ccekl = three ? 0.0 : ck[kp] - tk;
    if (!(ccekl <= cutoff)) break;
    const double cut = cutoff - ccekl;
    while (n > 0 && !(cb[n - 1] - tb <= cut && cb[n - 1] - tb < cutoff)) --n;
    if (n == 0) break;
    total += n;
  }
  return total;
}

static const KernelInfo* bin_kernels(const cintx_plan* P, const Bin& b);

// Lay out and upload pairs, bins and offsets in one device blob.
// Order every bin's tuples by (bra, ket) primitive-pair counts (counting
// sort), so that the tuples sharing a workgroup of the wide kernels have
// equal loop trip counts; for a class with a one-lane kernel instead by
// surviving primitive tuples (stable), so that the lanes of a wavefront,
// which each walk their own survivors, run equal numbers of iterations.
static void plan_sort_bins(cintx_plan* P) {
  const PairSet& S = P->pairs;
  const FamilyInfo& f = kFam[P->family];
  std::vector<std::vector<int64_t>> remap(P->bins.size());
  for (size_t bi = 0; bi < P->bins.size(); ++bi) {
    Bin& b = P->bins[bi];
    const int64_t n = (int64_t)b.q.size();
    auto counting_sort = [&](int K, const std::vector<int>& key, std::vector<int64_t>& pos) {
      std::vector<int64_t> count((size_t)K + 1, 0);
      for (int64_t i = 0; i < n; ++i) ++count[(size_t)key[(size_t)i] + 1];
      for (int k = 0; k < K; ++k) count[k + 1] += count[k];
      pos.resize((size_t)n);
      for (int64_t i = 0; i < n; ++i) pos[(size_t)i] = count[(size_t)key[(size_t)i]]++;
    };
    // first by primitive counts, then (stable) by survivors
    std::vector<int> key((size_t)n);
    for (int64_t i = 0; i < n; ++i)
      key[(size_t)i] = std::min(S.np[b.q[i].x], 64) * 65 + std::min(S.np[b.q[i].y], 64);
    std::vector<int64_t> pos1, pos2;
    counting_sort(65 * 65, key, pos1);
    std::vector<int2> by_np((size_t)n);
    for (int64_t i = 0; i < n; ++i) by_np[(size_t)pos1[(size_t)i]] = b.q[i];
    bool one_lane = false;
    if (const KernelInfo* ks = bin_kernels(P, b))
      for (int v = 0; v < kVariants; ++v)
        one_lane = one_lane || (ks[v].fn && ks[v].lanes_per_tuple == 1);
    if (!one_lane) {
      remap[bi] = pos1;
      b.q.swap(by_np);
      continue;
    }
    // one-lane classes: by survivors alone (stable), so that within equal
    // survivor counts the enumeration's bra-major order stays and
    // neighbouring lanes often share their bra pair (the fused J/K sums
    // such lanes before its atomics); a lane walks its survivors only, so
    // the primitive counts do not set its trip count
    const int lij = b.l[0] + b.l[1] + f.deriv, lkl = b.l[2] + b.l[3];
    constexpr int kSurvCap = 1 << 16;
    for (int64_t i = 0; i < n; ++i)
      key[(size_t)i] = (int)std::min<int64_t>(
          surviving_prims(S, b.q[i].x, b.q[i].y, lij, lkl, f.three, P->expcutoff),
          kSurvCap - 1);
    counting_sort(kSurvCap, key, pos2);
    std::vector<int2> sorted((size_t)n);
    for (int64_t i = 0; i < n; ++i) sorted[(size_t)pos2[(size_t)i]] = b.q[(size_t)i];
    remap[bi] = pos2;
    b.q.swap(sorted);
  }
  for (VQ& v : P->vqs) v.idx = remap[v.bin][(size_t)v.idx];
}

// The spinor-transform job of one contracted block: argument-slot shells
// shl, canonical permutation perm (canonical slot c holds argument
// perm[c]) and classes cl, contraction columns cc (canonical order), the
// block at `src` in the Cartesian buffer, the output block of the tuple at
// complex element dst0 with argument-slot strides so (complex elements).
static SpinorJob spinor_job(const FamilyInfo& f, const HostBasis& B, const int* shl,
                            const int* perm, const int* cl, const int* cc, int64_t src,
                            int64_t dst0, const int64_t* so) {
  SpinorJob J;
  std::memset(&J, 0, sizeof J);
  J.src = src;
  int cst = 1, cca[4] = {0, 0, 0, 0};
  for (int c = 0; c < 4; ++c) {
    const int a = perm[c];
    if (a < f.arity) {
      J.cs[a] = cst;
      cca[a] = cc[c];
    }
    cst *= ncart(cl[c]);
  }
  int64_t dst = dst0;
  for (int a = 0; a < f.arity; ++a) {
    const int sh = shl[a], l = B.l[sh], kap = B.kappa[sh];
    J.l[a] = (int8_t)l;
    J.kappa[a] = (int8_t)(kap < 0 ? -1 : (kap > 0 ? 1 : 0));
    dst += (int64_t)cca[a] * out_fun(f, a, l, kap) * so[a];
    if (a > 0) {
      if (so[a] > INT32_MAX) fail(CINTX_STATUS_UNSUPPORTED_API, "spinor block too large");
      J.so[a - 1] = (int)so[a];
    }
  }
  J.dst = dst;
  return J;
}

// Workgroups of a spinor launch: enough to fill the device, with their
// workspace held under `budget` bytes.
static int spinor_grid(int64_t njobs, int64_t slice, size_t budget) {
  const int64_t fit = std::max<int64_t>(1, (int64_t)(budget / ((size_t)slice * 16)));
  return (int)std::max<int64_t>(1, std::min<int64_t>({njobs, fit, (int64_t)2048}));
}

static void plan_upload(cintx_plan* P) {
  DevCtx& C = ctx();
  const FamilyInfo& fam = kFam[P->family];
  plan_sort_bins(P);
  int64_t base = 0;
  for (auto& b : P->bins) {
    b.out_base = base;
    base += b.block * (int64_t)b.q.size();
  }
  if (fam.spinor)
    P->cart_total = base;
  else if (base != P->total)
    fail(CINTX_STATUS_PANIC, "output size mismatch");
  std::vector<SpinorJob> jobs;
  if (fam.spinor) {
    const HostBasis& B = P->basis;
    P->spin_slice = 1;
    for (const TupleRec& tr : P->tuples) {
      int64_t so[4] = {1, 1, 1, 1};
      for (int a = 1; a < fam.arity; ++a) so[a] = so[a - 1] * out_extent(fam, B, a - 1, tr.shl[a - 1]);
      for (int64_t k = 0; k < tr.nvq; ++k) {
        const VQ& vq = P->vqs[tr.first_vq + k];
        const Bin& b = P->bins[vq.bin];
        const int cc[4] = {vq.cc[0], vq.cc[1], vq.cc[2], vq.cc[3]};
        jobs.push_back(spinor_job(fam, B, tr.shl, tr.perm, b.l, cc, b.out_base + vq.idx * b.block,
                                  tr.nat_off / 2, so));
        P->spin_slice = std::max(P->spin_slice, spinor_job_slice(jobs.back(), fam.arity == 4));
      }
    }
    P->njobs = (int64_t)jobs.size();
    P->spin_grid = spinor_grid(P->njobs, P->spin_slice, (size_t)1 << 30);
  }
  const PairSet& S = P->pairs;
  size_t npp = S.aij.size(), npair = S.off.size();
  std::vector<std::pair<size_t, size_t>> segs;  // (offset, bytes)
  size_t cur = 0;
  auto take = [&](size_t bytes) {
    size_t o = cur;
    cur = align256(cur + bytes);
    return o;
  };
  size_t o_rec = take(npp * 64);
  size_t o_poff = take(npair * 4), o_pnp = take(npair * 4), o_geo = take(npair * 64);
  std::vector<size_t> o_q(P->bins.size());
  for (size_t i = 0; i < P->bins.size(); ++i) o_q[i] = take(P->bins[i].q.size() * sizeof(int2));
  size_t o_cnt = take(P->bins.size() * 2 * 8 + 8);
  const int nbas = P->basis.nbas;
  size_t o_ao = take((size_t)nbas * 4), o_l = take((size_t)nbas * 4),
         o_psh = take(npair * sizeof(int2));
  int64_t scratch = 0;
  for (auto& b : P->bins)
    if (b.generic) {
      int64_t grid = std::min<int64_t>((int64_t)b.q.size(), kGenericGrid);
      scratch = std::max(scratch, grid * generic_scratch_per_block(b.l, kFam[P->family].deriv));
    }
  size_t o_out = take((size_t)P->total * 8);
  // one workspace per graph chain that runs runtime-class bins (so those of
  // different chains run concurrently): as many slices as such bins, up to
  // the chain count, fewer when that would pass 8 GiB
  int ngeneric = 0;
  for (auto& b : P->bins) ngeneric += b.generic ? 1 : 0;
  P->scratch_chains = 1;
  while (P->scratch_chains < kChains && 2 * P->scratch_chains <= ngeneric &&
         (size_t)scratch * 8 * (size_t)(2 * P->scratch_chains) <= ((size_t)8 << 30))
    P->scratch_chains *= 2;
  P->scratch_stride = scratch;
  size_t o_scr = take((size_t)scratch * 8 * (size_t)P->scratch_chains);
  size_t o_jobs = take(jobs.size() * sizeof(SpinorJob));
  size_t o_cart = take((size_t)P->cart_total * 8);
  size_t o_sscr = take(fam.spinor ? (size_t)P->spin_grid * (size_t)P->spin_slice * 16 : 0);
  size_t host_bytes = o_out;  // everything before the output is uploaded
  P->blob_bytes = cur;
  HIPCHK(hipMalloc(&P->blob, P->blob_bytes));
  std::vector<char> h(host_bytes, 0);
  auto put = [&](size_t o, const void* src, size_t bytes) {
    if (bytes) std::memcpy(h.data() + o, src, bytes);
  };
  {
    std::vector<double> rec(npp * 8);
    for (size_t k = 0; k < npp; ++k) {
      double* r = rec.data() + 8 * k;
      r[0] = S.aij[k];
      r[1] = S.wi[k];
      r[2] = S.wj[k];
      r[3] = 0.0;
      r[4] = S.kfac[k];
      r[5] = S.hinv[k];
      r[6] = S.cce[k];
      r[7] = S.ai[k];
    }
    put(o_rec, rec.data(), npp * 64);
  }
  put(o_poff, S.off.data(), npair * 4);
  put(o_pnp, S.np.data(), npair * 4);
  put(o_geo, S.geo.data(), npair * 64);
  {
    std::vector<int> ao(nbas);
    P->nao = 0;
    for (int sh = 0; sh < nbas; ++sh) {
      ao[sh] = P->nao;
      P->nao += nsph(P->basis.l[sh]) * P->basis.nctr[sh];
    }
    put(o_ao, ao.data(), (size_t)nbas * 4);
    put(o_l, P->basis.l.data(), (size_t)nbas * 4);
    put(o_psh, S.shells.data(), npair * sizeof(int2));
  }
  for (size_t i = 0; i < P->bins.size(); ++i) {
    Bin& b = P->bins[i];
    put(o_q[i], b.q.data(), b.q.size() * sizeof(int2));
    b.d_q = (int2*)(P->blob + o_q[i]);
  }
  HIPCHK(hipMemcpyAsync(P->blob, h.data(), host_bytes, hipMemcpyHostToDevice, C.stream));
  if (!jobs.empty())
    HIPCHK(hipMemcpyAsync(P->blob + o_jobs, jobs.data(), jobs.size() * sizeof(SpinorJob),
                          hipMemcpyHostToDevice, C.stream));
  HIPCHK(hipStreamSynchronize(C.stream));
  if (fam.spinor) {
    P->d_jobs = (SpinorJob*)(P->blob + o_jobs);
    P->d_cart = (double*)(P->blob + o_cart);
    P->d_spin_scratch = (double*)(P->blob + o_sscr);
  }
  P->d_rec = (double*)(P->blob + o_rec);
  P->d_poff = (int*)(P->blob + o_poff);
  P->d_pnp = (int*)(P->blob + o_pnp);
  P->d_geo = (double*)(P->blob + o_geo);
  P->d_counters = (unsigned long long*)(P->blob + o_cnt);
  P->d_ao_off = (int*)(P->blob + o_ao);
  P->d_shl_l = (int*)(P->blob + o_l);
  P->d_pair_sh = (int2*)(P->blob + o_psh);
  P->d_out = (double*)(P->blob + o_out);
  P->d_scratch = scratch ? (double*)(P->blob + o_scr) : nullptr;
  HIPCHK(hipEventCreate(&P->ev0));
  HIPCHK(hipEventCreate(&P->ev1));
}

// registry form of a bin: 0 FP64, 1 FP32, 2 FP64 with Cartesian output
static int bin_form(const cintx_plan* P, const Bin& b) {
  return kFam[P->family].cart ? 2 : b.prec;
}

static const KernelInfo* bin_kernels(const cintx_plan* P, const Bin& b) {
  const FamilyInfo& f = kFam[P->family];
  return ctx().reg->k[f.deriv][bin_form(P, b)][b.l[0]][b.l[1]][b.l[2]][b.l[3]];
}

// Variant v of a bin: v < kVariants picks the bin's own precision; an FP32
// bin of a mixed plan may also run the FP64 instantiation (v - kVariants),
// which the autotuner picks where FP32 does not pay (FP32 without packed math
// issues at the FP64 rate on CDNA4, so it only helps through registers and
// LDS bytes).
static const KernelInfo& variant_info(const cintx_plan* P, const Bin& b, int v) {
  static const KernelInfo none{};
  if (v < 0 || v >= 2 * kVariants) return none;
  if (v < kVariants) return bin_kernels(P, b)[v];
  if (b.prec != 1) return none;
  const FamilyInfo& f = kFam[P->family];
  return ctx().reg->k[f.deriv][0][b.l[0]][b.l[1]][b.l[2]][b.l[3]][v - kVariants];
}

// The fused J/K build of variant v (same launch shape; none for other
// families, runtime-class bins or shapes without an instantiation).
static const KernelInfo& variant_info_jk(const cintx_plan* P, const Bin& b, int v) {
  static const KernelInfo none{};
  if (P->family != 0 || b.generic || v < 0 || v >= 2 * kVariants) return none;
  if (v < kVariants) return ctx().reg->jk[b.prec][b.l[0]][b.l[1]][b.l[2]][b.l[3]][v];
  if (b.prec != 1) return none;
  return ctx().reg->jk[0][b.l[0]][b.l[1]][b.l[2]][b.l[3]][v - kVariants];
}

static signed char& tuned_slot(const cintx_plan* P, const Bin& b) {
  const FamilyInfo& f = kFam[P->family];
  return ctx().tuned[f.deriv][bin_form(P, b)][b.l[0]][b.l[1]][b.l[2]][b.l[3]];
}

static void plan_drop_graph(cintx_plan* P) {
  if (P->graph_exec) (void)hipGraphExecDestroy(P->graph_exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  P->graph_exec = nullptr;
  P->graph = nullptr;
}

static void plan_check_kernels(cintx_plan* P) {
  for (auto& b : P->bins) {
    const KernelInfo* ks = bin_kernels(P, b);
    b.generic = (ks[0].fn == nullptr);
    const int v = tuned_slot(P, b);
    b.variant = (v > 0 && variant_info(P, b, v).fn) ? v : 0;
  }
}

// Device pair data a launch reads (a plan's blob or an optimizer's).
struct PairDev {
  const double* prim;
  const int* off;
  const int* np;
  const double* geo;
};

// One class bin's launches: nq tuples d_q[0..nq) of class l (canonical
// slots) and form prec, blocks of `block` elements written from out_base.
// Variant `k` of the compiled class kernels, or the runtime-class kernel
// (generic) with `scratch` of kGenericGrid slices.
// Fused J/K of a launch: the class kernel digests its blocks into J/K.
struct JkFuse {
  const double* D;
  double* J;
  double* K;
  const int2* pair_sh;
  const int* ao_off;
  int nao;
};

static void launch_bin_kernels(int fam, const int* l, int variant, bool generic,
                               const KernelInfo& k, const PairDev& pd, double expcutoff,
                               const int2* d_q, int64_t nq, double* out, int64_t out_base,
                               int64_t block, double* scratch, unsigned long long* counters,
                               hipStream_t st, const JkFuse* jk = nullptr) {
  DevCtx& C = ctx();
  const FamilyInfo& f = kFam[fam];
  const double common_base = (M_PI * M_PI * M_PI) * 2.0 / std::sqrt(M_PI);
  EriArgs a;
  std::memset(&a, 0, sizeof a);
  a.prim = pd.prim;
  a.pair_off = pd.off;
  a.pair_np = pd.np;
  a.pair_geo = pd.geo;
  // one CINTcommon_fac_sp per real shell (g2e.c:54-56, g3c2e.c, g2c2e.c:44-45)
  a.common = common_base * common_fac_sp(l[0]) * (f.two ? 1.0 : common_fac_sp(l[1])) *
             common_fac_sp(l[2]) * (f.three ? 1.0 : common_fac_sp(l[3]));
  a.expcutoff = expcutoff;
  a.lij_ceil = l[0] + l[1] + f.deriv;
  a.lkl_ceil = l[2] + l[3];
  a.three_center = f.three ? 1 : 0;
  a.cart = f.cart ? 1 : 0;
  int ns0 = nfun(f, l[0]), ns1 = nfun(f, l[1]), ns2 = nfun(f, l[2]);
  a.st[0] = 1;
  a.st[1] = ns0;
  a.st[2] = (int64_t)ns0 * ns1;
  a.st[3] = (int64_t)ns0 * ns1 * ns2;
  a.st_comp = block / f.ncomp;
  a.out = out;
  a.rys = C.rys;
  a.c2s = C.d_c2s;
  a.counters = counters;
  a.scratch = scratch;
  if (jk) {
    if (generic || f.deriv || f.cart || f.three || f.two)
      fail(CINTX_STATUS_PANIC, "fused J/K needs an int2e_sph class kernel");
    a.jk_D = jk->D;
    a.jk_J = jk->J;
    a.jk_K = jk->K;
    a.pair_sh = jk->pair_sh;
    a.ao_off = jk->ao_off;
    a.nao = jk->nao;
  }
  // a dispatch holds at most 2^32 - 1 work-items per grid dimension: launch
  // in chunks of at most 2^31 lanes (2^28 tuples for the runtime-class
  // kernel, whose grid is capped anyway)
  const int64_t chunk = generic ? ((int64_t)1 << 28)
                                : (((int64_t)1 << 31) / std::max(1, k.lanes_per_tuple));
  for (int64_t s = 0; s < nq; s += chunk) {
    a.quartets = d_q + s;
    a.out_base = out_base + s * block;
    a.block = block;
    a.nq = std::min(chunk, nq - s);
    hipError_t e;
    // the launch wrappers read the thread's last HIP error after launching;
    // drop any error an earlier, deliberately unchecked call left behind
    (void)hipGetLastError();
    if (generic) {
      int grid = (int)std::min<int64_t>(a.nq, kGenericGrid);
      e = launch_generic(a, l, f.deriv, grid, st);
    } else {
      e = k.fn(a, st);
    }
    if (e != hipSuccess)
      fail(CINTX_STATUS_EXECUTION_FAILED,
           "kernel launch of class (" + std::to_string(l[0]) + std::to_string(l[1]) + "|" +
               std::to_string(l[2]) + std::to_string(l[3]) + ") " +
               (generic ? std::string("runtime-class kernel")
                        : "variant " + std::to_string(variant) + " (" +
                              std::to_string(k.lanes_per_tuple) + " lanes, " +
                              std::to_string(k.tuples_per_block) + " tuples per workgroup)") +
               ", " + std::to_string(a.nq) + " tuples: " + hipGetErrorString(e));
  }
}

static void plan_launch_bin(cintx_plan* P, size_t i, double* out, bool count,
                            hipStream_t st = nullptr, int chain = 0,
                            const JkFuse* jk = nullptr) {
  DevCtx& C = ctx();
  if (!st) st = C.stream;
  Bin& b = P->bins[i];
  if (kFam[P->family].spinor) out = P->d_cart;  // the transform reads them from there
  const KernelInfo& k = jk ? variant_info_jk(P, b, b.variant) : variant_info(P, b, b.variant);
  const PairDev pd{P->d_rec, P->d_poff, P->d_pnp, P->d_geo};
  double* scratch = b.generic ? P->d_scratch + (size_t)chain * (size_t)P->scratch_stride : nullptr;
  launch_bin_kernels(P->family, b.l, b.variant, b.generic, k, pd, P->expcutoff, b.d_q,
                     (int64_t)b.q.size(), out, b.out_base, b.block, scratch,
                     count ? P->d_counters + 2 * i : nullptr, st, jk);
}

// Spinor families: the Cartesian blocks (all bins done) -> `out`.
static void plan_launch_spinor(cintx_plan* P, double* out, hipStream_t st) {
  if (!kFam[P->family].spinor || P->njobs == 0) return;
  DevCtx& C = ctx();
  (void)hipGetLastError();
  hipError_t e = launch_spinor(P->d_jobs, P->njobs, P->d_cart, out, C.d_spinor, C.d_c2s,
                               P->d_spin_scratch, P->spin_slice, P->spin_grid,
                               kFam[P->family].arity == 4, st);
  if (e != hipSuccess)
    fail(CINTX_STATUS_EXECUTION_FAILED, std::string("spinor transform launch: ") + hipGetErrorString(e));
}

static void plan_launch_all(cintx_plan* P, double* out, bool count,
                            std::vector<hipEvent_t>* evs = nullptr) {
  DevCtx& C = ctx();
  for (size_t i = 0; i < P->bins.size(); ++i) {
    if (evs) HIPCHK(hipEventRecord((*evs)[2 * i], C.stream));
    plan_launch_bin(P, i, out, count);
    if (evs) HIPCHK(hipEventRecord((*evs)[2 * i + 1], C.stream));
  }
  plan_launch_spinor(P, out, C.stream);
}

static cintx_plan* plan_new(int family, const int* atm, int natm, const int* bas, int nbas,
                            const double* env, size_t env_len) {
  if (family < 0 || family >= kNumFamilies) fail(CINTX_STATUS_UNSUPPORTED_API, "unknown family");
  cintx_plan* P = new cintx_plan();
  P->family = family;
  try {
    P->basis = ingest(atm, natm, bas, nbas, env, env_len);
    P->pairs.init(P->basis);
    P->expcutoff = fam_expcutoff(family, env);
  } catch (...) {
    delete P;
    throw;
  }
  return P;
}

static void plan_free(cintx_plan* P) {
  if (!P) return;
  plan_drop_graph(P);
  if (P->ev0) (void)hipEventDestroy(P->ev0);
  if (P->ev1) (void)hipEventDestroy(P->ev1);
  if (P->blob) (void)hipFree(P->blob);
  delete P;
}

// Convert the canonical device blocks (already on the host in `canon`) into
// the libcint layout of every tuple (cart2sph.c:5359-5374 placement).
static void plan_to_natural(const cintx_plan* P, const double* canon, double* nat) {
  const FamilyInfo& f = kFam[P->family];
  const HostBasis& B = P->basis;
  for (const TupleRec& tr : P->tuples) {
    int n[4] = {1, 1, 1, 1};
    for (int a = 0; a < f.arity; ++a) n[a] = nfun(f, B.l[tr.shl[a]]) * B.nctr[tr.shl[a]];
    int64_t so[4] = {1, n[0], (int64_t)n[0] * n[1], (int64_t)n[0] * n[1] * n[2]};
    int64_t nout = (int64_t)n[0] * n[1] * n[2] * n[3];
    double* dst = nat + tr.nat_off;
    for (int64_t k = 0; k < tr.nvq; ++k) {
      const VQ& vq = P->vqs[tr.first_vq + k];
      const Bin& b = P->bins[vq.bin];
      const double* src = canon + b.out_base + vq.idx * b.block;
      int ns[4], base[4];
      int64_t st[4];
      for (int c = 0; c < 4; ++c) {
        ns[c] = nfun(f, b.l[c]);
        int a = tr.perm[c];
        base[c] = vq.cc[c] * ns[c];
        st[c] = a >= f.arity ? 0 : so[a];
      }
      int64_t nsall = (int64_t)ns[0] * ns[1] * ns[2] * ns[3];
      for (int comp = 0; comp < f.ncomp; ++comp) {
        const double* s = src + comp * nsall;
        double* d = dst + comp * nout;
        int64_t e = 0;
        for (int m3 = 0; m3 < ns[3]; ++m3)
          for (int m2 = 0; m2 < ns[2]; ++m2)
            for (int m1 = 0; m1 < ns[1]; ++m1)
              for (int m0 = 0; m0 < ns[0]; ++m0, ++e)
                d[(base[0] + m0) * st[0] + (base[1] + m1) * st[1] + (base[2] + m2) * st[2] +
                  (base[3] + m3) * st[3]] = s[e];
      }
    }
  }
}

// ------------------------------------------------------------------
// libcint cache-size formulas (cint2e.c:780-816, cint3c2e.c:543-567).
// ------------------------------------------------------------------
static int cache_size(int fam, const int* shls, const int* bas) {
  const FamilyInfo& f = kFam[fam];
  auto B = [&](int s, int slot) { return bas[8 * (size_t)s + slot]; };
  int l[4] = {0, 0, 0, 0}, np[4] = {1, 1, 1, 1}, nc[4] = {1, 1, 1, 1};
  for (int a = 0; a < f.arity; ++a) {
    l[a] = B(shls[a], 1);
    np[a] = B(shls[a], 2);
    nc[a] = B(shls[a], 3);
  }
  size_t nf = 1;
  for (int a = 0; a < f.arity; ++a) nf *= ncart(l[a]);
  size_t nctr = (size_t)nc[0] * nc[1] * nc[2] * nc[3];
  size_t ncomp = f.ncomp;
  int gbits = f.deriv ? 1 : 0;
  int inc[4] = {0, 0, 0, 0};  // ng[IINC..LINC]
  if (f.dcenter >= 0) inc[f.dcenter] = 1;
  int lic = l[0] + inc[0], ljc = l[1] + inc[1];
  size_t g_size;
  if (f.arity == 2) {
    // int1e_cache_size (cint1e.c:163-183) as CINT2c2e_drv uses it
    // (cint2c2e.c:245): shells (i, k), nf = nfi nfk, g_size over (li, lk)
    int li2 = l[0], lk2 = l[1];
    int nroots = (li2 + lk2) / 2 + 1;
    g_size = (size_t)nroots * (li2 + 1) * (lk2 + 1);
    size_t nf2 = (size_t)ncart(li2) * ncart(lk2);
    size_t nc2 = nf2 * nc[0] * nc[1];
    size_t leng = g_size * 3 * 2;
    size_t lenj = nf2 * nc2, leni = nf2 * nc[0], len0 = nf2;
    size_t pdata = (size_t)np[0] * np[1] * 5 + np[0] * nc[0] + np[1] * nc[1] +
                   (np[0] + np[1]) * 2 + nf2 * 3;
    return (int)std::max(nc2 + leng + lenj + leni + len0 + pdata, nc2 + nf2 * 8 * 2);
  }
  if (f.arity == 4) {
    int lkc = l[2] + inc[2], llc = l[3] + inc[3];
    int nroots = (lic + ljc + lkc + llc) / 2 + 1;
    bool ib = lic > ljc, kb = lkc > llc;
    int dli = ib ? lic + ljc + 1 : lic + 1, dlj = ib ? ljc + 1 : lic + ljc + 1;
    int dlk = kb ? lkc + llc + 1 : lkc + 1, dll = kb ? llc + 1 : lkc + llc + 1;
    g_size = (size_t)nroots * dli * dlk * dll * dlj;
    size_t pdata = (size_t)(np[0] * np[1] + np[2] * np[3]) * 5 + np[0] * nc[0] + np[1] * nc[1] +
                   np[2] * nc[2] + np[3] * nc[3] + (np[0] + np[1] + np[2] + np[3]) * 2 + nf * 3;
    size_t leng = g_size * 3 * ((1 << gbits) + 1);
    size_t len0 = nf * ncomp;
    size_t nc_all = nf * nctr;
    size_t a = leng + len0 + nc_all * ncomp * 3 + pdata;
    size_t b = nc_all * ncomp + nf * 4;
    if (f.spinor) {  // CINT2e_spinor_drv, cint2e.c:886-895
      const size_t ci = (size_t)spinor_len(l[0], B(shls[0], 4)) * nc[0];
      const size_t cj = (size_t)spinor_len(l[1], B(shls[1], 4)) * nc[1];
      const size_t n1 = ci * ncart(l[2]) * nc[2] * ncart(l[3]) * nc[3] * cj;
      b = nc_all * ncomp + n1 * 2 + nf * 32 * 2;
    }
    return (int)std::max(a, b);
  } else {
    int llc = l[2] + inc[2];  // g3c2e.c:68-69: k sits in the l slot
    int nroots = (lic + ljc + llc) / 2 + 1;
    bool ib = lic > ljc;
    int dli = ib ? lic + ljc + 1 : lic + 1, dlj = ib ? ljc + 1 : lic + ljc + 1;
    int dlk = llc + 1;
    g_size = (size_t)nroots * dli * dlk * dlj;
    size_t pdata = (size_t)np[0] * np[1] * 5 + np[0] * nc[0] + np[1] * nc[1] + np[2] * nc[2] +
                   (np[0] + np[1]) * 2 + np[2] + nf * 3 + 16;
    size_t leng = g_size * 3 * ((1 << gbits) + 1);
    size_t len0 = nf * ncomp;
    size_t nc_all = nf * nctr;
    size_t a = leng + len0 + nc_all * ncomp * 3 + pdata;
    size_t b = nc_all * ncomp + nf * 3;
    if (f.spinor) b = nc_all * ncomp + nf * 14 * 2;  // CINT3c2e_spinor_drv, cint3c2e.c:640-645
    return (int)std::max(a, b);
  }
}

// ------------------------------------------------------------------
// Raw (libcint-signature) calls: one shell tuple per call, any number of
// host threads at once.  Each thread has its own stream and device arena
// (created on its first call, kept for the process); the calls take no
// lock.  With an optimizer the pair data come from its device copy (built
// once per basis, like CINTOpt's pair data, optimizer.c:344-425); without,
// the call's own pairs are built and uploaded into the arena.
// ------------------------------------------------------------------
struct ThreadCtx {
  hipStream_t stream = nullptr;
  char* arena = nullptr;
  size_t cap = 0;
  std::vector<char> host;  // upload staging
};
static std::mutex g_threads_mu;
static std::vector<ThreadCtx*> g_threads;  // kept for the process
static thread_local ThreadCtx* t_ctx = nullptr;

static ThreadCtx& thread_ctx() {
  if (t_ctx) return *t_ctx;
  DevCtx& C = ctx();
  HIPCHK(hipSetDevice(C.device));
  ThreadCtx* T = new ThreadCtx();
  if (hipStreamCreateWithFlags(&T->stream, hipStreamNonBlocking) != hipSuccess) {
    delete T;
    fail(CINTX_STATUS_EXECUTION_FAILED, "hipStreamCreate failed");
  }
  {
    std::lock_guard<std::mutex> lk(g_threads_mu);
    g_threads.push_back(T);
  }
  t_ctx = T;
  return *T;
}

// The thread's arena holds at least `bytes` (grows by doubling; the stream
// is idle between calls, so the old buffer is free to go).
static char* arena(ThreadCtx& T, size_t bytes) {
  if (bytes > T.cap) {
    size_t cap = std::max<size_t>(bytes, std::max<size_t>(2 * T.cap, (size_t)1 << 20));
    if (T.arena) HIPCHK(hipFree(T.arena));
    T.arena = nullptr;
    T.cap = 0;
    hipError_t e = hipMalloc(&T.arena, cap);
    if (e != hipSuccess) fail(CINTX_STATUS_ALLOCATION_FAILED, hipGetErrorString(e));
    T.cap = cap;
  }
  return T.arena;
}

}  // namespace cx

// Device-resident optimizer: the basis, every pair of (virtual) shells in
// both orders plus the one-Gaussian pseudo pairs, and their 64-byte
// primitive records, pair offsets/counts and geometry in one device blob.
struct CINTOpt {
  int family = 0;
  int natm = 0, nbas = 0;
  cx::HostBasis basis;
  cx::PairSet pairs;
  char* blob = nullptr;
  cx::PairDev dev{};
};

namespace cx {

// Upload a pair set's device form (records, offsets, counts, geometry) to
// `dst` (device) through the host staging buffer; returns the pointers.
static size_t pairs_bytes(const PairSet& S) {
  return align256(S.aij.size() * 64) + align256(S.off.size() * 4) * 2 +
         align256(S.off.size() * 64);
}

static PairDev pairs_upload(const PairSet& S, char* dst, std::vector<char>& h, hipStream_t st) {
  const size_t npp = S.aij.size(), npair = S.off.size();
  const size_t o_rec = 0, o_off = align256(npp * 64), o_np = o_off + align256(npair * 4),
               o_geo = o_np + align256(npair * 4), total = o_geo + align256(npair * 64);
  h.assign(total, 0);
  double* rec = (double*)(h.data() + o_rec);
  for (size_t k = 0; k < npp; ++k) {
    double* r = rec + 8 * k;
    r[0] = S.aij[k];
    r[1] = S.wi[k];
    r[2] = S.wj[k];
    r[3] = 0.0;
    r[4] = S.kfac[k];
    r[5] = S.hinv[k];
    r[6] = S.cce[k];
    r[7] = S.ai[k];
  }
  if (npair) {
    std::memcpy(h.data() + o_off, S.off.data(), npair * 4);
    std::memcpy(h.data() + o_np, S.np.data(), npair * 4);
    std::memcpy(h.data() + o_geo, S.geo.data(), npair * 64);
  }
  HIPCHK(hipMemcpyAsync(dst, h.data(), total, hipMemcpyHostToDevice, st));
  return PairDev{(const double*)(dst + o_rec), (const int*)(dst + o_off),
                 (const int*)(dst + o_np), (const double*)(dst + o_geo)};
}

static CINTOpt* opt_build(int fam, const int* atm, int natm, const int* bas, int nbas,
                          const double* env) {
  DevCtx& C = ctx();
  CINTOpt* O = new CINTOpt();
  try {
    O->family = fam;
    O->natm = natm;
    O->nbas = nbas;
    O->basis = ingest(atm, natm, bas, nbas, env, 0);
    O->pairs.init(O->basis);
    const int nv = (int)O->basis.v.size();
    for (int a = 0; a < nv; ++a) {
      for (int b = 0; b < nv; ++b) O->pairs.add(O->basis, a, b);
      O->pairs.add_pseudo(O->basis, a);
    }
    HIPCHK(hipMalloc(&O->blob, std::max<size_t>(pairs_bytes(O->pairs), 256)));
    std::vector<char> h;
    O->dev = pairs_upload(O->pairs, O->blob, h, C.stream);
    HIPCHK(hipStreamSynchronize(C.stream));
  } catch (...) {
    if (O->blob) (void)hipFree(O->blob);
    delete O;
    throw;
  }
  return O;
}

// One libcint-signature call.  Returns not0.
static int raw_eval(int fam, double* out, const int* dims, const int* shls, const int* atm,
                    int natm, const int* bas, int nbas, const double* env, size_t env_len,
                    const CINTOpt* opt) {
  const FamilyInfo& f = kFam[fam];
  if (!shls) fail(CINTX_STATUS_NULL_POINTER, "shls is null");
  ThreadCtx& T = thread_ctx();
  HostBasis local_basis;
  PairSet local_pairs;
  const HostBasis* B;
  if (opt) {
    if (opt->natm != natm || opt->nbas != nbas)
      fail(CINTX_STATUS_INVALID_INPUT, "optimizer was built for another basis");
    B = &opt->basis;
  } else {
    local_basis = ingest(atm, natm, bas, nbas, env, env_len);
    local_pairs.init(local_basis);
    B = &local_basis;
  }
  int l[4] = {0, 0, 0, 0}, shl[4] = {0, 0, 0, 0};
  for (int a = 0; a < f.arity; ++a) {
    if (shls[a] < 0 || shls[a] >= B->nbas) fail(CINTX_STATUS_INVALID_INPUT, "shell index out of range");
    shl[a] = shls[a];
    l[a] = B->l[shl[a]];
  }
  int perm[4];
  canon_perm(fam, l, perm);
  int cl[4], cn[4], cs[4];
  for (int c = 0; c < 4; ++c) {
    const int a = perm[c];
    const bool dummy = a >= f.arity;
    cl[c] = dummy ? 0 : l[a];
    cn[c] = dummy ? 1 : B->nctr[shl[a]];
    cs[c] = dummy ? -1 : shl[a];
  }
  // virtual tuples (one per contraction column combination)
  const int nv = (int)B->v.size();
  auto pair_id = [&](int va, int vb) {  // vb < 0: pseudo pair of va
    if (opt) return opt->pairs.find(va, vb < 0 ? nv : vb);
    return vb < 0 ? local_pairs.add_pseudo(*B, va) : local_pairs.add(*B, va, vb);
  };
  std::vector<int2> q;
  std::vector<std::array<int, 4>> cc;
  for (int c3 = 0; c3 < cn[3]; ++c3)
    for (int c2 = 0; c2 < cn[2]; ++c2)
      for (int c1 = 0; c1 < cn[1]; ++c1)
        for (int c0 = 0; c0 < cn[0]; ++c0) {
          const int v0 = B->first_v[cs[0]] + c0;
          const int bra = f.two ? pair_id(v0, -1) : pair_id(v0, B->first_v[cs[1]] + c1);
          const int ket = f.three ? pair_id(B->first_v[cs[2]] + c2, -1)
                                  : pair_id(B->first_v[cs[2]] + c2, B->first_v[cs[3]] + c3);
          if (bra < 0 || ket < 0) fail(CINTX_STATUS_PANIC, "optimizer lacks a pair");
          q.push_back(make_int2(bra, ket));
          cc.push_back({c0, c1, c2, c3});
        }
  const int64_t nq = (int64_t)q.size();
  const int64_t block = (int64_t)f.ncomp * nfun(f, cl[0]) * nfun(f, cl[1]) * nfun(f, cl[2]) *
                        nfun(f, cl[3]);
  // kernel of the class: the autotuned or default variant, or the
  // runtime-class kernel
  DevCtx& C = ctx();
  const KernelInfo* ks = C.reg->k[f.deriv][f.cart ? 2 : 0][cl[0]][cl[1]][cl[2]][cl[3]];
  const bool generic = ks[0].fn == nullptr;
  int variant = 0;
  if (!generic) {
    const int v = C.tuned[f.deriv][f.cart ? 2 : 0][cl[0]][cl[1]][cl[2]][cl[3]];
    if (v > 0 && v < kVariants && ks[v].fn) variant = v;
  }
  // natural extents (complex elements for the spinor families)
  int n[4] = {1, 1, 1, 1};
  for (int a = 0; a < f.arity; ++a) n[a] = out_extent(f, *B, a, shl[a]);
  const int64_t nnat = (int64_t)n[0] * n[1] * n[2] * n[3];
  // spinor families: one transform job per contracted block, into a
  // compact natural block
  std::vector<SpinorJob> jobs;
  int64_t sslice = 1;
  int sgrid = 0;
  if (f.spinor) {
    const int64_t so[4] = {1, n[0], (int64_t)n[0] * n[1], (int64_t)n[0] * n[1] * n[2]};
    for (int64_t v = 0; v < nq; ++v) {
      const int ccv[4] = {cc[v][0], cc[v][1], cc[v][2], cc[v][3]};
      jobs.push_back(spinor_job(f, *B, shl, perm, cl, ccv, v * block, 0, so));
      sslice = std::max(sslice, spinor_job_slice(jobs.back(), f.arity == 4));
    }
    sgrid = spinor_grid(nq, sslice, (size_t)256 << 20);
  }
  // arena: [pairs (no optimizer)] [tuples] [output] [counters] [scratch]
  //        [spinor jobs, output, workspace]
  const size_t pbytes = opt ? 0 : pairs_bytes(local_pairs);
  const size_t o_q = pbytes, o_out = o_q + align256(nq * sizeof(int2));
  const size_t o_cnt = o_out + align256(nq * block * 8);
  const size_t o_scr = o_cnt + 256;
  const size_t scr = generic ? (size_t)std::min<int64_t>(nq, kGenericGrid) *
                                   generic_scratch_per_block(cl, f.deriv) * 8
                             : 0;
  const size_t o_jobs = align256(o_scr + scr);
  const size_t o_sout = o_jobs + align256(jobs.size() * sizeof(SpinorJob));
  const size_t o_sscr = o_sout + align256(f.spinor ? (size_t)nnat * 16 : 0);
  const size_t sscr = f.spinor ? (size_t)sgrid * sslice * 16 : 0;
  char* A = arena(T, f.spinor ? o_sscr + sscr : o_scr + scr);
  const PairDev pd = opt ? opt->dev : pairs_upload(local_pairs, A, T.host, T.stream);
  HIPCHK(hipMemcpyAsync(A + o_q, q.data(), nq * sizeof(int2), hipMemcpyHostToDevice, T.stream));
  HIPCHK(hipMemsetAsync(A + o_cnt, 0, 16, T.stream));
  double* dout = (double*)(A + o_out);
  unsigned long long* dcnt = (unsigned long long*)(A + o_cnt);
  launch_bin_kernels(fam, cl, variant, generic, ks[variant], pd, fam_expcutoff(fam, env),
                     (const int2*)(A + o_q), nq, dout, 0, block,
                     generic ? (double*)(A + o_scr) : nullptr, dcnt, T.stream);
  int d[4] = {n[0], n[1], n[2], n[3]};
  if (dims)
    for (int a = 0; a < f.arity; ++a) d[a] = dims[a];
  for (int a = 0; a < f.arity; ++a)
    if (d[a] < n[a]) fail(CINTX_STATUS_INVALID_INPUT, "dims smaller than the shell extents");
  const int64_t so[4] = {1, d[0], (int64_t)d[0] * d[1], (int64_t)d[0] * d[1] * d[2]};
  const int64_t nout_d = (int64_t)d[0] * d[1] * d[2] * d[3];
  unsigned long long cnt[2] = {0, 0};
  if (f.spinor) {
    // transform on the device, then place the compact block with the
    // caller's dims (cart2sph.c:4676-4703 zcopy_iklj placement)
    HIPCHK(hipMemcpyAsync(A + o_jobs, jobs.data(), jobs.size() * sizeof(SpinorJob),
                          hipMemcpyHostToDevice, T.stream));
    (void)hipGetLastError();
    HIPCHK(launch_spinor((const SpinorJob*)(A + o_jobs), nq, dout, (double*)(A + o_sout),
                         C.d_spinor, C.d_c2s, (double*)(A + o_sscr), sslice, sgrid,
                         f.arity == 4, T.stream));
    std::vector<double> nat((size_t)nnat * 2);
    HIPCHK(hipMemcpyAsync(nat.data(), A + o_sout, nat.size() * 8, hipMemcpyDeviceToHost,
                          T.stream));
    HIPCHK(hipMemcpyAsync(cnt, dcnt, 16, hipMemcpyDeviceToHost, T.stream));
    HIPCHK(hipStreamSynchronize(T.stream));
    const bool nz = cnt[0] != 0;
    int64_t e = 0;
    for (int m3 = 0; m3 < n[3]; ++m3)
      for (int m2 = 0; m2 < n[2]; ++m2)
        for (int m1 = 0; m1 < n[1]; ++m1)
          for (int m0 = 0; m0 < n[0]; ++m0, ++e) {
            double* dp = out + 2 * (m0 + m1 * so[1] + m2 * so[2] + m3 * so[3]);
            dp[0] = nz ? nat[(size_t)(2 * e)] : 0.0;
            dp[1] = nz ? nat[(size_t)(2 * e + 1)] : 0.0;
          }
    return nz ? 1 : 0;
  }
  std::vector<double> canon((size_t)(nq * block));
  HIPCHK(hipMemcpyAsync(canon.data(), dout, canon.size() * 8, hipMemcpyDeviceToHost, T.stream));
  HIPCHK(hipMemcpyAsync(cnt, dcnt, 16, hipMemcpyDeviceToHost, T.stream));
  HIPCHK(hipStreamSynchronize(T.stream));
  // place the canonical blocks into libcint's layout with the caller's dims
  const bool not0 = cnt[0] != 0;
  int ns[4];
  int64_t st[4];
  for (int c = 0; c < 4; ++c) {
    ns[c] = nfun(f, cl[c]);
    st[c] = perm[c] >= f.arity ? 0 : so[perm[c]];
  }
  const int64_t nsall = (int64_t)ns[0] * ns[1] * ns[2] * ns[3];
  for (int64_t v = 0; v < nq; ++v) {
    const double* src = canon.data() + v * block;
    int base[4];
    for (int c = 0; c < 4; ++c) base[c] = cc[v][c] * ns[c];
    for (int comp = 0; comp < f.ncomp; ++comp) {
      const double* sp = src + comp * nsall;
      double* dp = out + comp * nout_d;
      int64_t e = 0;
      for (int m3 = 0; m3 < ns[3]; ++m3)
        for (int m2 = 0; m2 < ns[2]; ++m2)
          for (int m1 = 0; m1 < ns[1]; ++m1)
            for (int m0 = 0; m0 < ns[0]; ++m0, ++e)
              dp[(base[0] + m0) * st[0] + (base[1] + m1) * st[1] + (base[2] + m2) * st[2] +
                 (base[3] + m3) * st[3]] = not0 ? sp[e] : 0.0;
    }
  }
  return not0 ? 1 : 0;
}

static int family_of_api(int api) {
  switch (api) {
    case CINTX_API_INT2E_SPH:
      return 0;
    case CINTX_API_INT2E_IP1_SPH:
      return 1;
    case CINTX_API_INT3C2E_SPH:
      return 2;
    case CINTX_API_INT3C2E_IP1_SPH:
      return 3;
    case CINTX_API_INT2C2E_SPH:
      return 4;
    case CINTX_API_INT2E_CART:
      return 7;
    case CINTX_API_INT2E_SPINOR:
      return 9;
    case CINTX_API_INT3C2E_SPINOR:
      return 10;
    default:
      return -1;
  }
}

// Scoped device allocation.
template <typename T>
struct DevArr {
  T* p = nullptr;
  explicit DevArr(size_t n) {
    HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  }
  ~DevArr() { (void)hipFree(p); }
};

}  // namespace cx

using namespace cx;

// ------------------------------------------------------------------
// 1. libcint raw signature
// ------------------------------------------------------------------
// A failed call leaves the block it would have written zeroed (so the
// caller's buffer is defined) and returns 0; with CINTX_STRICT=1 in the
// environment it aborts the process instead, so that a device failure
// cannot pass for a screened-out block.
static int raw_entry(int fam, double* out, int* dims, int* shls, int* atm, int natm, int* bas,
                     int nbas, double* env, const CINTOpt* opt) {
  if (!out) {
    if (!shls || !bas) return 0;
    return cache_size(fam, shls, bas);
  }
  int not0 = 0;
  int st = guarded(
      [&] { not0 = raw_eval(fam, out, dims, shls, atm, natm, bas, nbas, env, 0, opt); });
  if (st != CINTX_STATUS_SUCCESS) {
    fprintf(stderr, "libcintx_hip %s: %s\n", kFam[fam].name, t_msg.c_str());
    const char* strict = getenv("CINTX_STRICT");
    if (strict && strict[0] == '1') abort();
    const FamilyInfo& f = kFam[fam];
    if (shls && bas && nbas > 0) {
      int64_t n = (int64_t)f.ncomp * out_width(f);
      bool ok = true;
      for (int a = 0; a < f.arity; ++a) {
        if (shls[a] < 0 || shls[a] >= nbas) {
          ok = false;
          break;
        }
        const int* b = bas + 8 * (size_t)shls[a];
        n *= dims ? (int64_t)dims[a] : (int64_t)out_fun(f, a, b[1], b[4]) * b[3];
      }
      if (ok && n > 0) std::memset(out, 0, (size_t)n * sizeof(double));
    }
    return 0;
  }
  return not0;
}

extern "C" {

#define CINTX_RAW_ENTRY(name, fam)                                                        \
  int name(double* out, int* dims, int* shls, int* atm, int natm, int* bas, int nbas,     \
           double* env, CINTOpt* opt, double* cache) {                                    \
    (void)cache;                                                                          \
    return raw_entry(fam, out, dims, shls, atm, natm, bas, nbas, env, opt);               \
  }
CINTX_RAW_ENTRY(int2e_sph, 0)
CINTX_RAW_ENTRY(int2e_ip1_sph, 1)
CINTX_RAW_ENTRY(int3c2e_sph, 2)
CINTX_RAW_ENTRY(int3c2e_ip1_sph, 3)
CINTX_RAW_ENTRY(int2c2e_sph, 4)
CINTX_RAW_ENTRY(int2e_ip2_sph, 5)
CINTX_RAW_ENTRY(int3c2e_ip2_sph, 6)
CINTX_RAW_ENTRY(int2e_cart, 7)
CINTX_RAW_ENTRY(int3c2e_cart, 8)
// out is double complex (re, im interleaved)
CINTX_RAW_ENTRY(int2e_spinor, 9)
CINTX_RAW_ENTRY(int3c2e_spinor, 10)
#undef CINTX_RAW_ENTRY

// optimizer.c:344-425 builds the pair data of a basis once; so does this,
// on the device.  A failure leaves *opt NULL (calls then build their pairs
// per call).
static void make_opt(CINTOpt** opt, int fam, int* atm, int natm, int* bas, int nbas,
                     double* env) {
  if (!opt) return;
  *opt = nullptr;
  int st = guarded([&] { *opt = opt_build(fam, atm, natm, bas, nbas, env); });
  if (st != CINTX_STATUS_SUCCESS)
    fprintf(stderr, "libcintx_hip optimizer: %s\n", t_msg.c_str());
}
#define CINTX_OPTIMIZER(name, fam)                                                     \
  void name(CINTOpt** opt, int* atm, int natm, int* bas, int nbas, double* env) {       \
    make_opt(opt, fam, atm, natm, bas, nbas, env);                                     \
  }
CINTX_OPTIMIZER(int2e_optimizer, 0)
CINTX_OPTIMIZER(int2e_ip1_optimizer, 1)
CINTX_OPTIMIZER(int3c2e_optimizer, 2)
CINTX_OPTIMIZER(int3c2e_ip1_optimizer, 3)
CINTX_OPTIMIZER(int2c2e_optimizer, 4)
CINTX_OPTIMIZER(int2e_ip2_optimizer, 5)
CINTX_OPTIMIZER(int3c2e_ip2_optimizer, 6)
#undef CINTX_OPTIMIZER

void CINTdel_optimizer(CINTOpt** opt) {
  if (opt && *opt) {
    if ((*opt)->blob) (void)hipFree((*opt)->blob);
    delete *opt;
    *opt = nullptr;
  }
}

// ------------------------------------------------------------------
// 2. cintx C shim
// ------------------------------------------------------------------
static void shim_common(int fam, const int* dims, size_t dims_len, const int* shls,
                        size_t shls_len, const int* atm, size_t atm_len, const int* bas,
                        size_t bas_len, const double* env, size_t env_len, int64_t* required) {
  const FamilyInfo& f = kFam[fam];
  if (!shls) fail(CINTX_STATUS_NULL_POINTER, "shls is null");
  if (!atm) fail(CINTX_STATUS_NULL_POINTER, "atm is null");
  if (!bas) fail(CINTX_STATUS_NULL_POINTER, "bas is null");
  if (!env) fail(CINTX_STATUS_NULL_POINTER, "env is null");
  if (shls_len != (size_t)f.arity)
    fail(CINTX_STATUS_INVALID_INPUT, "shls length does not match the operator arity");
  if (atm_len % 6 || bas_len % 8 || atm_len == 0 || bas_len == 0)
    fail(CINTX_STATUS_INVALID_INPUT, "atm/bas lengths must be multiples of 6/8");
  int natm = (int)(atm_len / 6), nbas = (int)(bas_len / 8);
  HostBasis B = ingest(atm, natm, bas, nbas, env, env_len);
  int n[4] = {1, 1, 1, 1};
  for (int a = 0; a < f.arity; ++a) {
    if (shls[a] < 0 || shls[a] >= nbas) fail(CINTX_STATUS_INVALID_INPUT, "shell index out of range");
    n[a] = out_extent(f, B, a, shls[a]);
  }
  if (dims) {
    // crates/cintx-compat/src/layout.rs:25-64: dims must equal the natural extents
    if (dims_len != (size_t)f.arity) fail(CINTX_STATUS_INVALID_INPUT, "invalid dims length");
    for (int a = 0; a < f.arity; ++a)
      if (dims[a] != n[a]) fail(CINTX_STATUS_INVALID_INPUT, "dims differ from natural extents");
  }
  *required = (int64_t)f.ncomp * out_width(f) * n[0] * n[1] * n[2] * n[3];
}

int cintrs_query_workspace(int api, const int* dims, size_t dims_len, const int* shls,
                           size_t shls_len, const int* atm, size_t atm_len, const int* bas,
                           size_t bas_len, const double* env, size_t env_len, const void* opt,
                           CintxWorkspaceQuery* query_out) {
  (void)opt;
  int fam = family_of_api(api);
  if (fam < 0) {
    set_error(CINTX_STATUS_UNSUPPORTED_API, "api id " + std::to_string(api) + " is not served");
    return CINTX_STATUS_UNSUPPORTED_API;
  }
  return guarded([&] {
    if (!query_out) fail(CINTX_STATUS_NULL_POINTER, "query_out is null");
    int64_t req = 0;
    shim_common(fam, dims, dims_len, shls, shls_len, atm, atm_len, bas, bas_len, env, env_len,
                &req);
    CintxWorkspaceQuery q;
    q.bytes = (size_t)cache_size(fam, shls, bas) * 8;
    q.alignment = 64;
    q.required_bytes = (size_t)req * 8;
    q.chunk_count = 1;
    q.work_units = 1;
    *query_out = q;
  });
}

int cintrs_eval(int api, double* out, size_t out_len, const int* dims, size_t dims_len,
                const int* shls, size_t shls_len, const int* atm, size_t atm_len, const int* bas,
                size_t bas_len, const double* env, size_t env_len, const void* opt, double* cache,
                size_t cache_len, CintxEvalSummary* summary_out) {
  int fam = family_of_api(api);
  if (fam < 0) {
    set_error(CINTX_STATUS_UNSUPPORTED_API, "api id " + std::to_string(api) + " is not served");
    return CINTX_STATUS_UNSUPPORTED_API;
  }
  return guarded([&] {
    if (!summary_out) fail(CINTX_STATUS_NULL_POINTER, "summary_out is null");
    if (!out && out_len > 0) fail(CINTX_STATUS_NULL_POINTER, "out is null");
    if (!cache && cache_len > 0) fail(CINTX_STATUS_NULL_POINTER, "cache is null");
    if (out && cache && out == cache)
      fail(CINTX_STATUS_INVALID_INPUT, "`out` and `cache` pointers must not alias");
    int64_t req = 0;
    shim_common(fam, dims, dims_len, shls, shls_len, atm, atm_len, bas, bas_len, env, env_len,
                &req);
    size_t ws = (size_t)cache_size(fam, shls, bas) * 8;
    CintxEvalSummary s{0, 0, ws};
    if (!out) {
      *summary_out = s;
      return;
    }
    if (out_len < (size_t)req)
      fail(CINTX_STATUS_BUFFER_TOO_SMALL, "out holds " + std::to_string(out_len) + " elements, " +
                                              std::to_string(req) + " required");
    if (cache && cache_len * 8 < ws) fail(CINTX_STATUS_BUFFER_TOO_SMALL, "cache too small");
    // evaluate into a staging buffer first: no partial writes on failure
    std::vector<double> staging((size_t)req);
    int not0 = raw_eval(fam, staging.data(), nullptr, shls, atm, (int)(atm_len / 6), bas,
                        (int)(bas_len / 8), env, env_len, (const CINTOpt*)opt);
    std::memcpy(out, staging.data(), (size_t)req * 8);
    s.not0 = not0;
    s.bytes_written = (size_t)req * 8;
    *summary_out = s;
  });
}

int cintrs_last_error_code(void) { return t_status; }
void cintrs_clear_last_error(void) { clear_error(); }
size_t cintrs_copy_last_error_message(char* out, size_t out_len) {
  size_t need = t_msg.size() + 1;
  if (out && out_len) {
    size_t n = std::min(out_len - 1, t_msg.size());
    std::memcpy(out, t_msg.data(), n);
    out[n] = 0;
  }
  return need;
}

// ------------------------------------------------------------------
// 3. Plans
// ------------------------------------------------------------------
// sqrt(max |block|) of each diagonal tuple (4 ints each; family 0 for
// (ij|ij), family 4 for (P|P)), evaluated on the device with the
// coefficients of env_use.
static std::vector<double> diag_schwarz(int fam, const int* atm, int natm, const int* bas,
                                        int nbas, const double* env_use,
                                        const std::vector<int>& diag) {
  const int64_t n = (int64_t)diag.size() / 4;
  std::vector<double> qv((size_t)n, 0.0);
  if (n == 0) return qv;
  cintx_plan* D = plan_new(fam, atm, natm, bas, nbas, env_use, 0);
  struct G {
    cintx_plan* p;
    ~G() { plan_free(p); }
  } g{D};
  plan_add_tuples(D, diag.data(), n, 0);
  plan_check_kernels(D);
  plan_upload(D);
  DevCtx& C = ctx();
  plan_launch_all(D, D->d_out, false);
  std::vector<double> canon((size_t)D->total);
  HIPCHK(hipMemcpyAsync(canon.data(), D->d_out, (size_t)D->total * 8, hipMemcpyDeviceToHost,
                        C.stream));
  HIPCHK(hipStreamSynchronize(C.stream));
  for (int64_t t = 0; t < n; ++t) {
    const TupleRec& tr = D->tuples[t];
    double m = 0;
    for (int64_t k = 0; k < tr.nvq; ++k) {
      const VQ& vq = D->vqs[tr.first_vq + k];
      const Bin& b = D->bins[vq.bin];
      const double* src = canon.data() + b.out_base + vq.idx * b.block;
      for (int64_t e = 0; e < b.block; ++e) m = std::max(m, std::fabs(src[e]));
    }
    qv[t] = std::sqrt(m);
  }
  return qv;
}

// env with every contraction coefficient replaced by its absolute value
// (the Schwarz bound then bounds the size of the terms summed, which is what
// FP32 rounding follows, not the size of their possibly cancelled sum)
static std::vector<double> abs_coef_env(const int* atm, const int* bas, int nbas,
                                        const double* env) {
  size_t len = 0;
  for (int s = 0; s < nbas; ++s) {
    const int* b = bas + 8 * s;
    len = std::max(len, (size_t)b[5] + b[2]);
    len = std::max(len, (size_t)b[6] + (size_t)b[2] * b[3]);
    len = std::max(len, (size_t)atm[6 * b[0] + 1] + 3);
  }
  std::vector<double> eabs(env, env + len);
  for (int s = 0; s < nbas; ++s) {
    const int* b = bas + 8 * s;
    for (int k = 0; k < b[2] * b[3]; ++k)
      eabs[(size_t)b[6] + k] = std::fabs(eabs[(size_t)b[6] + k]);
  }
  return eabs;
}

// Precision of each tuple of an explicit int2e_sph / int3c2e_sph list in a
// mixed plan: FP32 recurrences when the Schwarz bound over the absolute
// coefficients, sqrt((ij|ij)) sqrt((kl|kl)) or sqrt((ij|ij)) sqrt((P|P)),
// is below mixed_threshold.
static std::vector<uint8_t> mixed_selection(int family, const int* atm, int natm,
                                            const int* bas, int nbas, const double* env,
                                            const int* tuples, int64_t nt, double thr) {
  const std::vector<double> eabs = abs_coef_env(atm, bas, nbas, env);
  std::unordered_map<uint64_t, int> pid, aid;
  std::vector<int> dpair, daux;
  auto pair_of = [&](int a, int b) {
    if (a < b) std::swap(a, b);
    if (a < 0 || b < 0 || a >= nbas) fail(CINTX_STATUS_INVALID_INPUT, "shell index out of range");
    const uint64_t key = ((uint64_t)a << 32) | (uint32_t)b;
    auto it = pid.find(key);
    if (it != pid.end()) return it->second;
    const int id = (int)pid.size();
    pid.emplace(key, id);
    dpair.insert(dpair.end(), {a, b, a, b});
    return id;
  };
  auto aux_of = [&](int k) {
    if (k < 0 || k >= nbas) fail(CINTX_STATUS_INVALID_INPUT, "shell index out of range");
    auto it = aid.find((uint64_t)k);
    if (it != aid.end()) return it->second;
    const int id = (int)aid.size();
    aid.emplace((uint64_t)k, id);
    daux.insert(daux.end(), {k, k, 0, 0});
    return id;
  };
  std::vector<int2> ref((size_t)nt);
  for (int64_t t = 0; t < nt; ++t) {
    const int* sh = tuples + 4 * t;
    ref[(size_t)t].x = pair_of(sh[0], sh[1]);
    ref[(size_t)t].y = family == 0 ? pair_of(sh[2], sh[3]) : aux_of(sh[2]);
  }
  const std::vector<double> qp = diag_schwarz(0, atm, natm, bas, nbas, eabs.data(), dpair);
  const std::vector<double> qa = diag_schwarz(4, atm, natm, bas, nbas, eabs.data(), daux);
  std::vector<uint8_t> prec((size_t)nt);
  for (int64_t t = 0; t < nt; ++t) {
    const double b = qp[(size_t)ref[(size_t)t].x] *
                     (family == 0 ? qp[(size_t)ref[(size_t)t].y] : qa[(size_t)ref[(size_t)t].y]);
    prec[(size_t)t] = b < thr ? 1 : 0;
  }
  return prec;
}

int cintx_plan_create(cintx_plan** plan, int family, const int* atm, int natm, const int* bas,
                      int nbas, const double* env, const int* tuples, int64_t ntuples,
                      int precision, double mixed_threshold) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    if (!tuples && ntuples > 0) fail(CINTX_STATUS_NULL_POINTER, "tuples is null");
    if (precision != CINTX_PREC_FP64 && precision != CINTX_PREC_MIXED)
      fail(CINTX_STATUS_INVALID_INPUT, "unknown precision mode");
    if (precision == CINTX_PREC_MIXED && family != CINTX_FAMILY_INT2E_SPH &&
        family != CINTX_FAMILY_INT3C2E_SPH)
      fail(CINTX_STATUS_INVALID_INPUT, "the mixed mode serves int2e_sph and int3c2e_sph");
    cintx_plan* P = plan_new(family, atm, natm, bas, nbas, env, 0);
    try {
      if (precision == CINTX_PREC_MIXED) {
        const std::vector<uint8_t> prec = mixed_selection(family, atm, natm, bas, nbas, env,
                                                          tuples, ntuples, mixed_threshold);
        plan_add_tuples(P, tuples, ntuples, 0, prec.data());
        P->ntuples_fp32 = 0;
        for (uint8_t v : prec) P->ntuples_fp32 += v;
      } else {
        plan_add_tuples(P, tuples, ntuples, 0);
        P->ntuples_fp32 = 0;
      }
      plan_check_kernels(P);
      plan_upload(P);
    } catch (...) {
      plan_free(P);
      throw;
    }
    *plan = P;
  });
}

int cintx_plan_destroy(cintx_plan* plan) {
  return guarded([&] { plan_free(plan); });
}

int cintx_plan_sizes(const cintx_plan* plan, int64_t* nelem, int64_t* ntuples) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    if (nelem) *nelem = plan->total;
    if (ntuples) *ntuples = (int64_t)plan->tuples.size();
  });
}

int cintx_plan_tuples(const cintx_plan* plan, int* tuples, int64_t* offsets) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    for (size_t t = 0; t < plan->tuples.size(); ++t) {
      if (tuples)
        for (int a = 0; a < 4; ++a) tuples[4 * t + a] = plan->tuples[t].shl[a];
      if (offsets) offsets[t] = plan->tuples[t].nat_off;
    }
  });
}

int cintx_plan_execute(cintx_plan* plan, double* out_dev) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    DevCtx& C = ctx();
    HIPCHK(hipEventRecord(plan->ev0, C.stream));
    plan_launch_all(plan, out_dev ? out_dev : plan->d_out, false);
    HIPCHK(hipEventRecord(plan->ev1, C.stream));
    plan->timed = true;
  });
}

int cintx_plan_execute_graph(cintx_plan* plan, double* out_dev) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    DevCtx& C = ctx();
    double* out = out_dev ? out_dev : plan->d_out;
    if (plan->graph_exec && plan->graph_out != out) plan_drop_graph(plan);
    if (!plan->graph_exec) {
      // kChains chains of bins (fork/join on the plan stream), so that
      // kernels of different classes share the CUs (the one-lane kernels
      // leave registers or LDS idle that another class uses) and one bin's
      // tail overlaps the others; runtime-class bins use their chain's
      // workspace slice (only the first scratch_chains chains have one)
      hipEvent_t ev[kChains] = {};  // ev[0]: fork, ev[1..]: joins
      for (int c = 0; c < kChains; ++c)
        if (hipEventCreateWithFlags(&ev[c], hipEventDisableTiming) != hipSuccess) {
          for (int d = 0; d < c; ++d) (void)hipEventDestroy(ev[d]);
          fail(CINTX_STATUS_EXECUTION_FAILED, "hipEventCreate failed");
        }
      auto drop_events = [&] {
        for (int c = 0; c < kChains; ++c) (void)hipEventDestroy(ev[c]);
      };
      if (hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        drop_events();
        fail(CINTX_STATUS_EXECUTION_FAILED, "hipStreamBeginCapture failed");
      }
      try {
        HIPCHK(hipEventRecord(ev[0], C.stream));
        for (int c = 0; c < kChains - 1; ++c) HIPCHK(hipStreamWaitEvent(C.branch[c], ev[0], 0));
        size_t n_other = 0, n_gen = 0;
        for (size_t i = 0; i < plan->bins.size(); ++i) {
          const int chain = plan->bins[i].generic ? (int)(n_gen++ % plan->scratch_chains)
                                                  : (int)(n_other++ % kChains);
          plan_launch_bin(plan, i, out, false, chain ? C.branch[chain - 1] : C.stream, chain);
        }
        for (int c = 0; c < kChains - 1; ++c) {
          HIPCHK(hipEventRecord(ev[c + 1], C.branch[c]));
          HIPCHK(hipStreamWaitEvent(C.stream, ev[c + 1], 0));
        }
        plan_launch_spinor(plan, out, C.stream);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(C.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        drop_events();
        throw;
      }
      const hipError_t ec = hipStreamEndCapture(C.stream, &plan->graph);
      drop_events();
      HIPCHK(ec);
      HIPCHK(hipGraphInstantiate(&plan->graph_exec, plan->graph, nullptr, nullptr, 0));
      plan->graph_out = out;
    }
    HIPCHK(hipEventRecord(plan->ev0, C.stream));
    HIPCHK(hipGraphLaunch(plan->graph_exec, C.stream));
    HIPCHK(hipEventRecord(plan->ev1, C.stream));
    plan->timed = true;
  });
}

int cintx_plan_sync(cintx_plan* plan) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    DevCtx& C = ctx();
    HIPCHK(hipStreamSynchronize(C.stream));
    // the time of the last execute, once (events never recorded would make
    // hipEventElapsedTime fail, and that error would stay the thread's last
    // HIP error and be reported by the next kernel launch)
    if (plan->timed) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, plan->ev0, plan->ev1));
      plan->last_ms = ms;
      plan->timed = false;
    }
  });
}

int cintx_plan_copy_output(cintx_plan* plan, double* host_out) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    if (!host_out) fail(CINTX_STATUS_NULL_POINTER, "host_out is null");
    DevCtx& C = ctx();
    if (kFam[plan->family].spinor) {  // already in the callers' layout
      HIPCHK(hipMemcpyAsync(host_out, plan->d_out, (size_t)plan->total * 8,
                            hipMemcpyDeviceToHost, C.stream));
      HIPCHK(hipStreamSynchronize(C.stream));
      return;
    }
    std::vector<double> canon((size_t)plan->total);
    HIPCHK(hipMemcpyAsync(canon.data(), plan->d_out, (size_t)plan->total * 8,
                          hipMemcpyDeviceToHost, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
    plan_to_natural(plan, canon.data(), host_out);
  });
}

// Algorithmic work per surviving primitive tuple and per tuple (c2s), the
// trip counts of libcint's loops (same model as the oracle's work counter).
static void class_work(int fam, const int* l, double* per_prim, double* per_tuple, int* nroots) {
  const FamilyInfo& f = kFam[fam];
  int Lc[4] = {l[0] + f.deriv, l[1], l[2], l[3]};
  int nr = (Lc[0] + Lc[1] + Lc[2] + Lc[3]) / 2 + 1;
  int nmax = Lc[0] + Lc[1], mmax = Lc[2] + Lc[3];
  double nf = (double)ncart(l[0]) * ncart(l[1]) * ncart(l[2]) * ncart(l[3]);
  double per = 24.0 * nr + 3.0 * nr * (4.0 * (nmax + mmax) + 7.0 * nmax * mmax);
  int A = std::max(Lc[0], Lc[1]), Bv = std::min(Lc[0], Lc[1]);
  int Cv = std::max(Lc[2], Lc[3]), D = std::min(Lc[2], Lc[3]);
  double hb = 0, hk = 0;
  for (int b = 1; b <= Bv; ++b) hb += (A + Bv - b + 1) * (double)(mmax + 1);
  for (int d = 1; d <= D; ++d) hk += (Cv + D - d + 1) * (double)((Lc[0] + 1) * (Lc[1] + 1));
  per += 6.0 * nr * (hb + hk);
  if (f.deriv)
    per += 9.0 * nf * nr + 9.0 * nr * (l[0] + 1) * (Lc[1] + 1) * (Lc[2] + 1) * (Lc[3] + 1);
  else
    per += 3.0 * nf * nr;
  per += 2.0 * nf * f.ncomp;
  double c2s = 0, dim[4] = {(double)ncart(l[0]), (double)ncart(l[1]), (double)ncart(l[2]),
                            (double)ncart(l[3])};
  for (int a = 0; a < 4; ++a) {
    double other = 1;
    for (int b = 0; b < 4; ++b)
      if (b != a) other *= dim[b];
    if (f.cart) continue;  // Cartesian output: no transform
    c2s += 2.0 * other * nsph(l[a]) * ncart(l[a]);
    dim[a] = nsph(l[a]);
  }
  *per_prim = per;
  *per_tuple = c2s * f.ncomp;
  *nroots = nr;
}

int cintx_plan_copy_tuples(cintx_plan* plan, const int64_t* which, int64_t n, double* host_out,
                           int64_t* offsets) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    if (n > 0 && (!which || !host_out)) fail(CINTX_STATUS_NULL_POINTER, "null argument");
    DevCtx& C = ctx();
    const FamilyInfo& f = kFam[plan->family];
    int64_t pos = 0;
    for (int64_t k = 0; k < n; ++k) {
      int64_t t = which[k];
      if (t < 0 || t >= (int64_t)plan->tuples.size())
        fail(CINTX_STATUS_INVALID_INPUT, "tuple index out of range");
      const TupleRec& tr = plan->tuples[t];
      if (offsets) offsets[k] = pos;
      if (f.spinor) {  // d_out holds the callers' layout
        HIPCHK(hipMemcpyAsync(host_out + pos, plan->d_out + tr.nat_off, (size_t)tr.nat_size * 8,
                              hipMemcpyDeviceToHost, C.stream));
        HIPCHK(hipStreamSynchronize(C.stream));
        pos += tr.nat_size;
        continue;
      }
      // gather this tuple's canonical blocks, then place them
      // size the staging buffer first: copies in flight must not see it move
      std::vector<int64_t> where((size_t)tr.nvq);
      size_t total = 0;
      for (int64_t v = 0; v < tr.nvq; ++v) {
        where[(size_t)v] = (int64_t)total;
        total += (size_t)plan->bins[plan->vqs[tr.first_vq + v].bin].block;
      }
      std::vector<double> blocks(total);
      for (int64_t v = 0; v < tr.nvq; ++v) {
        const VQ& vq = plan->vqs[tr.first_vq + v];
        const Bin& b = plan->bins[vq.bin];
        HIPCHK(hipMemcpyAsync(blocks.data() + where[(size_t)v],
                              plan->d_out + b.out_base + vq.idx * b.block, (size_t)b.block * 8,
                              hipMemcpyDeviceToHost, C.stream));
      }
      HIPCHK(hipStreamSynchronize(C.stream));
      int nn[4] = {1, 1, 1, 1};
      for (int a = 0; a < f.arity; ++a)
        nn[a] = nfun(f, plan->basis.l[tr.shl[a]]) * plan->basis.nctr[tr.shl[a]];
      int64_t so[4] = {1, nn[0], (int64_t)nn[0] * nn[1], (int64_t)nn[0] * nn[1] * nn[2]};
      int64_t nout = (int64_t)nn[0] * nn[1] * nn[2] * nn[3];
      double* dst = host_out + pos;
      for (int64_t v = 0; v < tr.nvq; ++v) {
        const VQ& vq = plan->vqs[tr.first_vq + v];
        const Bin& b = plan->bins[vq.bin];
        const double* src = blocks.data() + where[v];
        int ns[4], base[4];
        int64_t st[4];
        for (int c = 0; c < 4; ++c) {
          ns[c] = nfun(f, b.l[c]);
          int a = tr.perm[c];
          base[c] = vq.cc[c] * ns[c];
          st[c] = a >= f.arity ? 0 : so[a];
        }
        int64_t nsall = (int64_t)ns[0] * ns[1] * ns[2] * ns[3];
        for (int comp = 0; comp < f.ncomp; ++comp) {
          int64_t e = 0;
          for (int m3 = 0; m3 < ns[3]; ++m3)
            for (int m2 = 0; m2 < ns[2]; ++m2)
              for (int m1 = 0; m1 < ns[1]; ++m1)
                for (int m0 = 0; m0 < ns[0]; ++m0, ++e)
                  dst[comp * nout + (base[0] + m0) * st[0] + (base[1] + m1) * st[1] +
                      (base[2] + m2) * st[2] + (base[3] + m3) * st[3]] = src[comp * nsall + e];
        }
      }
      pos += tr.nat_size;
    }
  });
}

int cintx_plan_info(const cintx_plan* plan, int64_t* ntuples_fp32, int64_t* nbins) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    if (ntuples_fp32) *ntuples_fp32 = plan->ntuples_fp32;
    if (nbins) *nbins = (int64_t)plan->bins.size();
  });
}

int cintx_plan_execute_timed(cintx_plan* plan, double* out_dev) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    DevCtx& C = ctx();
    std::vector<hipEvent_t> evs(2 * plan->bins.size());
    for (auto& e : evs) HIPCHK(hipEventCreate(&e));
    struct G {
      std::vector<hipEvent_t>& v;
      ~G() {
        for (auto e : v) (void)hipEventDestroy(e);
      }
    } g{evs};
    HIPCHK(hipEventRecord(plan->ev0, C.stream));
    plan_launch_all(plan, out_dev ? out_dev : plan->d_out, false, &evs);
    HIPCHK(hipEventRecord(plan->ev1, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, plan->ev0, plan->ev1));
    plan->last_ms = ms;
    for (size_t i = 0; i < plan->bins.size(); ++i) {
      HIPCHK(hipEventElapsedTime(&ms, evs[2 * i], evs[2 * i + 1]));
      plan->bins[i].last_ms = ms;
    }
  });
}

int cintx_plan_bin_stats(const cintx_plan* plan, int64_t nmax, int* cls, int64_t* ntuples,
                         double* ms, double* prim_tuples, double* flops) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    int64_t n = std::min<int64_t>(nmax, (int64_t)plan->bins.size());
    for (int64_t i = 0; i < n; ++i) {
      const Bin& b = plan->bins[i];
      if (cls) {
        for (int c = 0; c < 4; ++c) cls[6 * i + c] = b.l[c];
        cls[6 * i + 4] = b.prec;
        cls[6 * i + 5] = b.generic ? 1 : 0;
      }
      if (ntuples) ntuples[i] = (int64_t)b.q.size();
      if (ms) ms[i] = b.last_ms;
      if (prim_tuples) prim_tuples[i] = b.prim_tuples;
      if (flops) flops[i] = b.flops;
    }
  });
}

int cintx_plan_autotune(cintx_plan* plan, int reps) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    if (reps < 1) fail(CINTX_STATUS_INVALID_INPUT, "reps must be >= 1");
    DevCtx& C = ctx();
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    struct G {
      hipEvent_t a, b;
      ~G() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
      }
    } guard{e0, e1};
    plan_drop_graph(plan);
    for (size_t i = 0; i < plan->bins.size(); ++i) {
      Bin& b = plan->bins[i];
      if (b.generic || b.q.empty()) continue;
      int best = b.variant;
      double best_ms = 1e300;
      for (int v = 0; v < 2 * kVariants; ++v) {
        if (!variant_info(plan, b, v).fn) continue;
        b.variant = v;
        auto timed = [&]() {
          HIPCHK(hipEventRecord(e0, C.stream));
          plan_launch_bin(plan, i, plan->d_out, false);
          HIPCHK(hipEventRecord(e1, C.stream));
          HIPCHK(hipEventSynchronize(e1));
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, e0, e1));
          return (double)ms;
        };
        // two warm-up launches, timed too (the first may load the code
        // object): a variant still 4x slower than the best so far is not
        // timed further
        double vmin = timed();
        vmin = std::min(vmin, timed());
        if (vmin > 4.0 * best_ms) continue;
        for (int r = 0; r < reps; ++r) vmin = std::min(vmin, timed());
        if (vmin < best_ms) {
          best_ms = vmin;
          best = v;
        }
      }
      b.variant = best;
      b.last_ms = best_ms;
      tuned_slot(plan, b) = (signed char)best;
    }
  });
}

int cintx_plan_set_variant(cintx_plan* plan, int variant) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    if (variant < 0 || variant >= kVariants)
      fail(CINTX_STATUS_INVALID_INPUT, "variant out of range");
    plan_drop_graph(plan);
    for (auto& b : plan->bins) {
      const KernelInfo* ks = bin_kernels(plan, b);
      b.variant = (!b.generic && ks[variant].fn) ? variant : 0;
    }
  });
}

int cintx_plan_bin_variants(const cintx_plan* plan, int64_t nmax, int* variant, int* nvariants,
                            int* lanes, int* prims, int* tuples_per_block) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    int64_t n = std::min<int64_t>(nmax, (int64_t)plan->bins.size());
    for (int64_t i = 0; i < n; ++i) {
      const Bin& b = plan->bins[i];
      int nv = 0;
      for (int v = 0; v < 2 * kVariants; ++v) nv += variant_info(plan, b, v).fn ? 1 : 0;
      const KernelInfo& k = variant_info(plan, b, b.variant);
      if (variant) variant[i] = b.generic ? -1 : b.variant;
      if (nvariants) nvariants[i] = nv;
      if (lanes) lanes[i] = b.generic ? 256 : k.lanes_per_tuple;
      if (prims) prims[i] = b.generic ? 1 : k.prims_per_step;
      if (tuples_per_block) tuples_per_block[i] = b.generic ? 1 : k.tuples_per_block;
    }
  });
}

int cintx_plan_last_time_ms(const cintx_plan* plan, double* ms) {
  return guarded([&] {
    if (!plan || !ms) fail(CINTX_STATUS_NULL_POINTER, "null argument");
    *ms = plan->last_ms;
  });
}

int cintx_plan_work(cintx_plan* plan, double* prim_tuples, double* root_evals, double* flops) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    DevCtx& C = ctx();
    HIPCHK(hipMemsetAsync(plan->d_counters, 0, plan->bins.size() * 16 + 8, C.stream));
    plan_launch_all(plan, plan->d_out, true);
    std::vector<unsigned long long> cnt(plan->bins.size() * 2);
    HIPCHK(hipMemcpyAsync(cnt.data(), plan->d_counters, cnt.size() * 8, hipMemcpyDeviceToHost,
                          C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
    double np = 0, nr = 0, fl = 0;
    for (size_t i = 0; i < plan->bins.size(); ++i) {
      double per, pt;
      int r;
      class_work(plan->family, plan->bins[i].l, &per, &pt, &r);
      np += (double)cnt[2 * i];
      nr += (double)cnt[2 * i] * r;
      fl += (double)cnt[2 * i] * per + (double)cnt[2 * i + 1] * pt;
      plan->bins[i].prim_tuples = (double)cnt[2 * i];
      plan->bins[i].flops = (double)cnt[2 * i] * per + (double)cnt[2 * i + 1] * pt;
    }
    if (prim_tuples) *prim_tuples = np;
    if (root_evals) *root_evals = nr;
    if (flops) *flops = fl;
  });
}

int cintx_plan_create_screened(cintx_plan** plan, const int* atm, int natm, const int* bas,
                               int nbas, const double* env, double threshold, int bra_stride,
                               int precision, double mixed_threshold) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    if (bra_stride < 1) fail(CINTX_STATUS_INVALID_INPUT, "bra_stride must be >= 1");
    if (precision != CINTX_PREC_FP64 && precision != CINTX_PREC_MIXED)
      fail(CINTX_STATUS_INVALID_INPUT, "unknown precision mode");
    // Schwarz bounds Q_ij = sqrt(max |(ij|ij)|) over i >= j, from a diagonal plan.
    std::vector<int> diag;
    for (int i = 0; i < nbas; ++i)
      for (int j = 0; j <= i; ++j) {
        diag.push_back(i);
        diag.push_back(j);
        diag.push_back(i);
        diag.push_back(j);
      }
    int64_t npair_all = (int64_t)diag.size() / 4;
    auto diag_bounds = [&](const double* env_use) {
      std::vector<double> qv((size_t)npair_all, 0.0);
      cintx_plan* D = plan_new(0, atm, natm, bas, nbas, env_use, 0);
      struct G {
        cintx_plan* p;
        ~G() { plan_free(p); }
      } g{D};
      for (int s = 0; s < nbas; ++s)
        if (D->basis.nctr[s] != 1)
          fail(CINTX_STATUS_UNSUPPORTED_API, "screened plans need segmented shells (nctr == 1)");
      plan_add_tuples(D, diag.data(), npair_all, 0);
      plan_check_kernels(D);
      plan_upload(D);
      DevCtx& C = ctx();
      plan_launch_all(D, D->d_out, false);
      std::vector<double> canon((size_t)D->total);
      HIPCHK(hipMemcpyAsync(canon.data(), D->d_out, (size_t)D->total * 8, hipMemcpyDeviceToHost,
                            C.stream));
      HIPCHK(hipStreamSynchronize(C.stream));
      for (int64_t t = 0; t < npair_all; ++t) {
        const TupleRec& tr = D->tuples[t];
        double m = 0;
        for (int64_t k = 0; k < tr.nvq; ++k) {
          const VQ& vq = D->vqs[tr.first_vq + k];
          const Bin& b = D->bins[vq.bin];
          const double* src = canon.data() + b.out_base + vq.idx * b.block;
          for (int64_t e = 0; e < b.block; ++e) m = std::max(m, std::fabs(src[e]));
        }
        qv[t] = std::sqrt(m);
      }
      return qv;
    };
    const std::vector<double> qv = diag_bounds(env);
    // Mixed plans select FP32 tuples with the same bound taken over the
    // absolute values of the contraction coefficients: primitive
    // contributions of mixed sign can cancel far below their own size, and
    // the FP32 rounding follows the size of the terms, not of the sum.
    std::vector<double> qabs;
    if (precision == CINTX_PREC_MIXED) {
      size_t len = 0;
      for (int s = 0; s < nbas; ++s) {
        const int* b = bas + 8 * s;
        len = std::max(len, (size_t)b[5] + b[2]);
        len = std::max(len, (size_t)b[6] + (size_t)b[2] * b[3]);
        len = std::max(len, (size_t)atm[6 * b[0] + 1] + 3);
      }
      std::vector<double> eabs(env, env + len);
      for (int s = 0; s < nbas; ++s) {
        const int* b = bas + 8 * s;
        for (int k = 0; k < b[2] * b[3]; ++k) eabs[(size_t)b[6] + k] = std::fabs(eabs[(size_t)b[6] + k]);
      }
      qabs = diag_bounds(eabs.data());
    }
    cintx_plan* P = plan_new(0, atm, natm, bas, nbas, env, 0);
    try {
      const HostBasis& B = P->basis;
      double qmax = 0;
      for (double q : qv) qmax = std::max(qmax, q);
      struct SP {
        int a, b;  // oriented: l[a] >= l[b]
        double q;
        int64_t id;
      };
      std::vector<SP> sp;
      for (int64_t t = 0; t < npair_all; ++t) {
        if (qv[t] * qmax < threshold) continue;
        int i = diag[4 * t], j = diag[4 * t + 1];
        SP s;
        s.a = B.l[i] >= B.l[j] ? i : j;
        s.b = B.l[i] >= B.l[j] ? j : i;
        s.q = qv[t];
        s.id = t;
        sp.push_back(s);
      }
      std::vector<int> order(sp.size());
      for (size_t k = 0; k < sp.size(); ++k) order[k] = (int)k;
      std::sort(order.begin(), order.end(), [&](int x, int y) { return sp[x].q > sp[y].q; });
      std::vector<int> tuples;
      for (size_t ip = 0; ip < sp.size(); ++ip) {
        const SP& p = sp[ip];
        if (p.id % bra_stride) continue;
        double need = threshold / std::max(p.q, 1e-300);
        for (int oq : order) {
          const SP& q = sp[oq];
          if (q.q < need) break;
          if (q.id > p.id) continue;
          tuples.push_back(p.a);
          tuples.push_back(p.b);
          tuples.push_back(q.a);
          tuples.push_back(q.b);
        }
      }
      int64_t nt = (int64_t)tuples.size() / 4;
      if (precision == CINTX_PREC_MIXED) {
        // split by bound: FP32 bins for tuples below mixed_threshold
        std::vector<int> lo, hi;
        std::unordered_map<uint64_t, double> qof;
        for (const SP& s : sp) qof[((uint64_t)s.a << 32) | (uint32_t)s.b] = qabs[(size_t)s.id];
        // FP32 without packed math issues at the FP64 rate on CDNA4: the FP32
        // recurrences pay only in classes that must share g through LDS (half
        // the LDS bytes and registers); classes with a one-lane register
        // kernel are faster in FP64, so they stay FP64
        auto fp32_pays = [&](const int* sh) {
          int la = B.l[sh[0]], lb = B.l[sh[1]], lc = B.l[sh[2]], ld = B.l[sh[3]];
          if (lc > la || (lc == la && ld > lb)) {
            std::swap(la, lc);
            std::swap(lb, ld);
          }
          if (la > kMaxL || lb > kMaxL || lc > kMaxL || ld > kMaxL) return false;
          const KernelInfo* k64 = ctx().reg->k[0][0][la][lb][lc][ld];
          if (!ctx().reg->k[0][1][la][lb][lc][ld][0].fn) return false;  // no FP32 build
          for (int v = 0; v < kVariants; ++v)
            if (k64[v].fn && k64[v].lanes_per_tuple == 1) return false;
          return true;
        };
        for (int64_t t = 0; t < nt; ++t) {
          double b = qof[((uint64_t)tuples[4 * t] << 32) | (uint32_t)tuples[4 * t + 1]] *
                     qof[((uint64_t)tuples[4 * t + 2] << 32) | (uint32_t)tuples[4 * t + 3]];
          auto& dst = (b < mixed_threshold && fp32_pays(&tuples[4 * t])) ? lo : hi;
          dst.insert(dst.end(), tuples.begin() + 4 * t, tuples.begin() + 4 * t + 4);
        }
        plan_add_tuples(P, hi.data(), (int64_t)hi.size() / 4, 0);
        plan_add_tuples(P, lo.data(), (int64_t)lo.size() / 4, 1);
        P->ntuples_fp32 = (int64_t)lo.size() / 4;
      } else {
        plan_add_tuples(P, tuples.data(), nt, 0);
      }
      plan_check_kernels(P);
      plan_upload(P);
    } catch (...) {
      plan_free(P);
      throw;
    }
    *plan = P;
  });
}

// Rys rule from the Chebyshev table, evaluated on the host (same table and
// Clenshaw recurrence as rys_eval in eri_device.h).
int cintx_rys_eval_host(int n, const double* x, int64_t nx, double* u, double* w) {
  return guarded([&] {
    if (n < 1 || n > kMaxRoots) fail(CINTX_STATUS_INVALID_INPUT, "nroots out of range");
    if (!x || !u || !w) fail(CINTX_STATUS_NULL_POINTER, "null argument");
    static std::mutex mu;
    static std::vector<double> coef, large;
    static int off[kMaxRoots + 1];
    {
      std::lock_guard<std::mutex> lk(mu);
      if (coef.empty()) rys_table_build(coef, large, off);
    }
    for (int64_t i = 0; i < nx; ++i) {
      double xi = x[i];
      for (int r = 0; r < n; ++r) {
        if (!(xi < rys_largex(n))) {
          const double* L = large.data() + n * (n - 1);
          double sr = L[r] / xi;
          u[i * n + r] = sr / (1.0 - sr);  // the asymptotic s is small
          w[i * n + r] = L[n + r] / std::sqrt(xi);
          continue;
        }
        int iv = (int)xi;
        double t = 2.0 * (xi - iv) - 1.0, t2 = t + t;
        const double* cu = coef.data() + off[n] + ((size_t)iv * 2 * n + r) * kRysNc;
        const double* cw = cu + (size_t)n * kRysNc;
        double bu1 = 0, bu2 = 0, bw1 = 0, bw2 = 0;
        for (int k = kRysDeg; k >= 1; --k) {
          double bu0 = std::fma(t2, bu1, cu[k] - bu2);
          double bw0 = std::fma(t2, bw1, cw[k] - bw2);
          bu2 = bu1;
          bu1 = bu0;
          bw2 = bw1;
          bw1 = bw0;
        }
        double v = std::fma(t, bu1, cu[0] - bu2);
        double small = std::fabs(v);
        double sr = v < 0 ? 1.0 - small : small, qr = v < 0 ? small : 1.0 - small;
        u[i * n + r] = sr / qr;
        w[i * n + r] = std::fma(t, bw1, cw[0] - bw2);
      }
    }
  });
}

// The same rule evaluated by the device (rys_eval of the class kernels).
int cintx_rys_eval_device(int n, const double* x, int64_t nx, double* u, double* w) {
  return guarded([&] {
    if (n < 1 || n > kMaxRoots) fail(CINTX_STATUS_INVALID_INPUT, "nroots out of range");
    if (!x || !u || !w) fail(CINTX_STATUS_NULL_POINTER, "null argument");
    DevCtx& C = ctx();
    double *dx = nullptr, *du = nullptr, *dw = nullptr;
    size_t bx = (size_t)std::max<int64_t>(nx, 1) * 8, bu = bx * n;
    HIPCHK(hipMalloc(&dx, bx));
    HIPCHK(hipMalloc(&du, bu));
    HIPCHK(hipMalloc(&dw, bu));
    struct F {
      double *a, *b, *c;
      ~F() {
        (void)hipFree(a);
        (void)hipFree(b);
        (void)hipFree(c);
      }
    } f{dx, du, dw};
    HIPCHK(hipMemcpy(dx, x, (size_t)nx * 8, hipMemcpyHostToDevice));
    HIPCHK(launch_rys_eval(C.rys, n, dx, nx, du, dw, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
    HIPCHK(hipMemcpy(u, du, (size_t)nx * n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(w, dw, (size_t)nx * n * 8, hipMemcpyDeviceToHost));
  });
}

// ------------------------------------------------------------------
// Learned Rys initializer (rys_mlp.hip)
// ------------------------------------------------------------------
int cintx_rys_mlp_nparam(void) { return mlp_nparam(); }

int cintx_rys_mlp_train(float* params, int64_t nsamples, int epochs, int batch, float lr,
                        uint64_t seed, double* final_loss) {
  return guarded([&] {
    if (!params) fail(CINTX_STATUS_NULL_POINTER, "params is null");
    if (nsamples <= 0 || epochs <= 0 || batch <= 0)
      fail(CINTX_STATUS_INVALID_INPUT, "nsamples, epochs and batch must be positive");
    DevCtx& C = ctx();
    const int np = mlp_nparam();
    DevArr<float> d(np);
    HIPCHK(hipMemcpy(d.p, params, np * 4, hipMemcpyHostToDevice));
    HIPCHK(mlp_train_device(C.rys, d.p, nsamples, epochs, batch, lr, seed, final_loss, C.stream));
    HIPCHK(hipMemcpy(params, d.p, np * 4, hipMemcpyDeviceToHost));
  });
}

int cintx_rys_mlp_grad(const float* params, const float* z, const int* nroots,
                       const float* targets, int batch, float* grad, float* loss) {
  return guarded([&] {
    if (!params || !z || !nroots || !targets || !grad || !loss)
      fail(CINTX_STATUS_NULL_POINTER, "null argument");
    DevCtx& C = ctx();
    const int np = mlp_nparam();
    DevArr<float> dp(np), dz((size_t)batch * 2), dt((size_t)batch * 26), dg(np), dl(batch);
    DevArr<int> dn(batch);
    HIPCHK(hipMemcpy(dp.p, params, np * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dz.p, z, (size_t)batch * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dn.p, nroots, (size_t)batch * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dt.p, targets, (size_t)batch * 26 * 4, hipMemcpyHostToDevice));
    HIPCHK(mlp_grad_device(dp.p, dz.p, dn.p, dt.p, batch, dg.p, dl.p, C.stream));
    HIPCHK(hipMemcpy(grad, dg.p, np * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(loss, dl.p, (size_t)batch * 4, hipMemcpyDeviceToHost));
  });
}

int cintx_rys_mlp_eval(const float* params, int n, const double* x, int64_t nx, int newton_steps,
                       int init_mode, double* u, double* w, int* status, double* polish_err,
                       double* ms) {
  return guarded([&] {
    if (n < 1 || n > 13) fail(CINTX_STATUS_INVALID_INPUT, "nroots must be in 1..13");
    if (init_mode < 0 || init_mode > 3) fail(CINTX_STATUS_INVALID_INPUT, "init_mode must be 0..3");
    if (!params || !x || !u || !w || !status) fail(CINTX_STATUS_NULL_POINTER, "null argument");
    DevCtx& C = ctx();
    const int np = mlp_nparam();
    DevArr<float> dp(np);
    DevArr<double> dx(nx), du((size_t)nx * n), dw((size_t)nx * n);
    DevArr<int> ds(nx);
    DevArr<double> de(polish_err ? nx : 0);
    HIPCHK(hipMemcpy(dp.p, params, np * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dx.p, x, (size_t)nx * 8, hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, C.stream));
    HIPCHK(mlp_eval_device(dp.p, C.rys, n, dx.p, nx, newton_steps, init_mode, du.p, dw.p, ds.p,
                           polish_err ? de.p : nullptr, C.stream));
    HIPCHK(hipEventRecord(e1, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipMemcpy(u, du.p, (size_t)nx * n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(w, dw.p, (size_t)nx * n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(status, ds.p, (size_t)nx * 4, hipMemcpyDeviceToHost));
    if (polish_err)
      HIPCHK(hipMemcpy(polish_err, de.p, (size_t)nx * 8, hipMemcpyDeviceToHost));
  });
}

// J/K from an int2e_sph plan over unique (8-fold) quartets of segmented
// shells: checks shared by both entry points.
static void jk_check(const cintx_plan* plan, const void* D, const void* J, const void* K) {
  if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
  if (!D || !J || !K) fail(CINTX_STATUS_NULL_POINTER, "null argument");
  if (plan->family != 0) fail(CINTX_STATUS_UNSUPPORTED_API, "J/K needs an int2e_sph plan");
  const HostBasis& B = plan->basis;
  for (int sh = 0; sh < B.nbas; ++sh)
    if (B.nctr[sh] != 1) fail(CINTX_STATUS_UNSUPPORTED_API, "J/K needs segmented shells");
}

// Enqueue the J/K build on the plan stream: dJ and dK are cleared with a
// fill kernel (graph-safe), every bin's class kernel is followed by its
// digest on the same stream, and J = A + A^T, K = B + B^T at the end.  Dsym
// is the symmetric density in device memory.  With chains, the bins are
// spread over the context's chains of streams (fork/join through `ev`, the
// caller's kChains events); without, everything runs on the plan stream.
static void jk_enqueue(cintx_plan* plan, const double* Dsym, double* dJ, double* dK,
                       hipEvent_t* ev) {
  DevCtx& C = ctx();
  const int nao = plan->nao;
  const int64_t nn = (int64_t)nao * nao;
  HIPCHK(launch_fill(dJ, nn, 0.0, C.stream));
  HIPCHK(launch_fill(dK, nn, 0.0, C.stream));
  const int nchain = ev ? kChains : 1;
  if (ev) {
    HIPCHK(hipEventRecord(ev[0], C.stream));
    for (int c = 0; c < kChains - 1; ++c) HIPCHK(hipStreamWaitEvent(C.branch[c], ev[0], 0));
  }
  // Every compiled class kernel digests its own blocks (fused J/K: no ERI
  // reaches memory); the runtime-class kernel writes its bins' blocks and a
  // digest kernel reads them back.  Bins of different classes overlap across
  // chains (J and K take atomics only); runtime-class bins use their
  // chain's workspace slice.
  const JkFuse fuse{Dsym, dJ, dK, plan->d_pair_sh, plan->d_ao_off, nao};
  size_t n_other = 0, n_gen = 0;
  for (size_t i = 0; i < plan->bins.size(); ++i) {
    const Bin& b = plan->bins[i];
    const int chain = b.generic ? (int)(n_gen++ % std::min(plan->scratch_chains, nchain))
                                : (int)(n_other++ % nchain);
    hipStream_t st = chain ? C.branch[chain - 1] : C.stream;
    if (variant_info_jk(plan, b, b.variant).fn) {
      plan_launch_bin(plan, i, plan->d_out, false, st, chain, &fuse);
      continue;
    }
    plan_launch_bin(plan, i, plan->d_out, false, st, chain);
    HIPCHK(launch_jk_digest(plan->d_out, b.out_base, b.block, b.d_q, (int64_t)b.q.size(),
                            plan->d_pair_sh, plan->d_ao_off, plan->d_shl_l, Dsym, dJ, dK, nao,
                            st));
  }
  if (ev) {
    for (int c = 0; c < kChains - 1; ++c) {
      HIPCHK(hipEventRecord(ev[c + 1], C.branch[c]));
      HIPCHK(hipStreamWaitEvent(C.stream, ev[c + 1], 0));
    }
  }
  // the digest accumulates one triangle's worth: J = A + A^T, K = B + B^T
  HIPCHK(launch_jk_symmetrize(dJ, nao, 1.0, C.stream));
  HIPCHK(launch_jk_symmetrize(dK, nao, 1.0, C.stream));
}

// J/K build; D, J, K are nao x nao row-major on the host.
int cintx_plan_jk(cintx_plan* plan, const double* D, double* J, double* K) {
  return guarded([&] {
    jk_check(plan, D, J, K);
    DevCtx& C = ctx();
    const size_t nn = (size_t)plan->nao * plan->nao;
    DevArr<double> dD(nn), dJ(nn), dK(nn);
    hipEvent_t ev[kChains] = {};  // ev[0]: fork, ev[1..]: joins
    // (declared after the device buffers, so on an early exit it waits for
    // every chain before those are freed)
    struct EvGuard {
      hipEvent_t* e;
      DevCtx& c;
      ~EvGuard() {
        for (int b = 0; b < kChains - 1; ++b) (void)hipStreamSynchronize(c.branch[b]);
        (void)hipStreamSynchronize(c.stream);
        for (int b = 0; b < kChains; ++b)
          if (e[b]) (void)hipEventDestroy(e[b]);
      }
    } evguard{ev, C};
    for (int c = 0; c < kChains; ++c)
      HIPCHK(hipEventCreateWithFlags(&ev[c], hipEventDisableTiming));
    HIPCHK(hipMemcpyAsync(dD.p, D, nn * 8, hipMemcpyHostToDevice, C.stream));
    // D is taken as symmetric: (D + D^T) / 2 (the digest relies on it)
    HIPCHK(launch_jk_symmetrize(dD.p, plan->nao, 0.5, C.stream));
    jk_enqueue(plan, dD.p, dJ.p, dK.p, ev);
    HIPCHK(hipMemcpyAsync(J, dJ.p, nn * 8, hipMemcpyDeviceToHost, C.stream));
    HIPCHK(hipMemcpyAsync(K, dK.p, nn * 8, hipMemcpyDeviceToHost, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
  });
}

// The same J/K build with device-resident D, J, K and no allocation: the
// symmetrised density goes to the library's static scratch (one per
// device), everything is enqueued on the plan stream (no chains, so no
// events are created), and the call returns without waiting.
int cintx_plan_jk_noalloc(cintx_plan* plan, const double* D_dev, double* J_dev, double* K_dev) {
  return guarded([&] {
    jk_check(plan, D_dev, J_dev, K_dev);
    DevCtx& C = ctx();
    const size_t need = (size_t)plan->nao * plan->nao * sizeof(double);
    if (need > jk_scratch_bytes())
      fail(CINTX_STATUS_MEMORY_LIMIT_EXCEEDED,
           "the density (" + std::to_string(plan->nao) + " functions, " + std::to_string(need) +
               " bytes) exceeds the " + std::to_string(jk_scratch_bytes()) +
               "-byte scratch of the allocation-free J/K");
    HIPCHK(launch_jk_sym_copy(D_dev, C.d_jk_scratch, plan->nao, C.stream));
    jk_enqueue(plan, C.d_jk_scratch, J_dev, K_dev, C.jk_ev);
  });
}

int cintx_spinor_coeffs(int l, int kappa, double* re, double* im) {
  if (l < 0 || l > kMaxL) return -1;
  static const std::vector<double> tab = [] {
    std::vector<double> c2s, t;
    build_c2s_table(c2s);
    build_spinor_table(c2s, t);
    return t;
  }();
  const int nd = spinor_len(l, kappa), row0 = spinor_row0(l, kappa), n2 = 2 * ncart(l);
  const double* src = tab.data() + 2 * ((size_t)spinor_tab_off(l) + (size_t)row0 * n2);
  for (int k = 0; k < nd * n2; ++k) {
    if (re) re[k] = src[2 * k];
    if (im) im[k] = src[2 * k + 1];
  }
  return nd;
}

int cintx_plan_fill_output(cintx_plan* plan, double value) {
  return guarded([&] {
    if (!plan) fail(CINTX_STATUS_NULL_POINTER, "plan is null");
    DevCtx& C = ctx();
    HIPCHK(launch_fill(plan->d_out, plan->total, value, C.stream));
  });
}

int cintx_device_alloc(double** ptr, int64_t nelem) {
  return guarded([&] {
    if (!ptr) fail(CINTX_STATUS_NULL_POINTER, "ptr is null");
    ctx();
    hipError_t e = hipMalloc(ptr, (size_t)std::max<int64_t>(nelem, 1) * 8);
    if (e != hipSuccess) fail(CINTX_STATUS_ALLOCATION_FAILED, hipGetErrorString(e));
  });
}
int cintx_device_free(double* ptr) {
  return guarded([&] { HIPCHK(hipFree(ptr)); });
}
int cintx_device_upload(double* dst_dev, const double* src_host, int64_t nelem) {
  return guarded([&] {
    if (nelem < 0) fail(CINTX_STATUS_INVALID_INPUT, "negative size");
    if (nelem == 0) return;
    if (!dst_dev || !src_host) fail(CINTX_STATUS_NULL_POINTER, "null argument");
    ctx();
    HIPCHK(hipMemcpy(dst_dev, src_host, (size_t)nelem * 8, hipMemcpyHostToDevice));
  });
}
int cintx_device_download(double* dst_host, const double* src_dev, int64_t nelem) {
  return guarded([&] {
    if (nelem < 0) fail(CINTX_STATUS_INVALID_INPUT, "negative size");
    if (nelem == 0) return;
    if (!dst_host || !src_dev) fail(CINTX_STATUS_NULL_POINTER, "null argument");
    ctx();
    HIPCHK(hipMemcpy(dst_host, src_dev, (size_t)nelem * 8, hipMemcpyDeviceToHost));
  });
}
int cintx_device_count(int* n) {
  return guarded([&] {
    if (!n) fail(CINTX_STATUS_NULL_POINTER, "n is null");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
  });
}

}  // extern "C"
