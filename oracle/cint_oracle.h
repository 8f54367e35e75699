/*
 * cint_oracle.h -- CPU oracle for the Rys two-electron path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the libcint 6.1.3
 * algorithm that cintx's hot path reproduces (int2e_sph, int2e_ip1_sph,
 * int3c2e_sph, int3c2e_ip1_sph).  It is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker;
 * the product library (libcintx_hip.so) never links, loads or calls it.
 *
 * Built from one source as liboracle_cint.so (FP64, libcint's rounding),
 * liboracle_cint_fma.so (the same with fused multiply-add) and
 * liboracle_cint_ext.so (long double with well-conditioned transfers: the
 * judge of FP64 rounding; see cint_oracle.c).
 *
 * Every entry point keeps libcint's raw signature (out, dims, shls, atm, natm,
 * bas, nbas, env, opt, cache) with an `orc_` prefix so it can sit in the same
 * process as the product library without symbol clashes.
 */
#ifndef CINT_ORACLE_H
#define CINT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* libcint slot layout (libcint-master/include/cint.h.in:28-67) */
#define ORC_PTR_EXPCUTOFF 0
#define ORC_PTR_ENV_START 20
#define ORC_ATM_SLOTS 6
#define ORC_BAS_SLOTS 8
#define ORC_MAX_ROOTS 16

/* Boys function F_m(x), m = 0..mmax (fmt.c:186-226). */
void orc_boys(int mmax, double x, double *f);

/* n Rys roots u (= t^2/(1-t^2)) and weights w for argument x
 * (rys_roots.c:57-123).  Roots are returned in ascending order.  Returns 0. */
int orc_rys_roots(int nroots, double x, double *u, double *w);
/* Same, with the number of discretisation nodes forced (convergence tests). */
int orc_rys_roots_m(int nroots, double x, double *u, double *w, int nodes);

/* The same rule unrounded: s = t^2 = u/(1+u) ascending and w, in long double
 * (80-bit on x86-64). */
void orc_rys_rule_ld(int nroots, double x, long double *s, long double *w);

/* Cartesian->real-spherical coefficient C[m][c] for angular momentum l
 * (cart2sph.c:21 table; p is the identity, cart2sph.c:33-43). */
double orc_c2s_coef(int l, int m, int c);

/* One primitive pair's data (CINTset_pairdata, optimizer.c:288-342):
 * out[0..2] = product centre P, out[3] = exp(-ai aj/(ai+aj)|ri-rj|^2),
 * out[4] = the screening key. */
void orc_pairdata_one(double ai, double aj, const double *ri, const double *rj, double *out);

/* libcint-signature integrals (cint2e.c:1186, grad2.c:81, cint3c2e.c:693,
 * autocode/int3c2e.c:80).  opt and cache are accepted and ignored. */
int orc_int2e_sph(double *out, int *dims, int *shls, int *atm, int natm,
                  int *bas, int nbas, double *env, void *opt, double *cache);
int orc_int2e_cart(double *out, int *dims, int *shls, int *atm, int natm,
                   int *bas, int nbas, double *env, void *opt, double *cache);
int orc_int2e_ip1_sph(double *out, int *dims, int *shls, int *atm, int natm,
                      int *bas, int nbas, double *env, void *opt, double *cache);
int orc_int3c2e_sph(double *out, int *dims, int *shls, int *atm, int natm,
                    int *bas, int nbas, double *env, void *opt, double *cache);
int orc_int3c2e_ip1_sph(double *out, int *dims, int *shls, int *atm, int natm,
                        int *bas, int nbas, double *env, void *opt, double *cache);

/* Batched helpers used by the tests (OpenMP over the list).
 * api: 0 int2e_sph, 1 int2e_ip1_sph, 2 int3c2e_sph, 3 int3c2e_ip1_sph,
 *      4 int2e_cart.
 * shls_list holds nq tuples of 4 ints (3c2e uses the first 3); block q is
 * written at out + offsets[q] with natural dims.  Returns the number of
 * tuples whose not0 flag was set. */
long orc_eval_list(int api, int nq, const int *shls_list, const long *offsets,
                   double *out, int *atm, int natm, int *bas, int nbas,
                   double *env, int nthreads);

/* Algorithmic work model of one shell tuple (libcint loop trip counts):
 * counts[0] = surviving primitive tuples, counts[1] = root evaluations,
 * counts[2] = FLOPs (root coefficients, VRR, HRR, gout, contraction, c2s).
 * api as for orc_eval_list. */
void orc_work_model(int api, int *shls, int *atm, int natm, int *bas, int nbas,
                    double *env, double *counts);

#ifdef __cplusplus
}
#endif
#endif
