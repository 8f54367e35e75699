/*
 * cintx_hip.h -- C ABI of libcintx_hip.so, the MI355X (gfx950) Rys
 * two-electron engine that replaces cintx's CubeCL backend for the families
 * int2e_sph, int2e_ip1_sph, int3c2e_sph and int3c2e_ip1_sph.
 *
 * Three layers, each a drop-in for a reference interface:
 *
 *   1. libcint raw signature (pyscf-style callers, cintx-compat legacy
 *      wrappers): int2e_sph & co.  Same arguments, output layout, return value
 *      (not0) and out == NULL cache-size query as libcint 6.1.3.
 *   2. cintx C shim: cintrs_eval / cintrs_query_workspace / last-error
 *      accessors with the same status codes and api ids as
 *      crates/cintx-capi/src/shim.rs and errors.rs.
 *   3. Batched plans (the executor boundary, crates/cintx-runtime/src/
 *      dispatch.rs:187-195 BackendExecutor): device-resident basis + pair
 *      data (CINTOpt analogue), class-binned tuple lists, device output.
 *
 * Integer types follow libcint's FINT = int build (crates/cintx-oracle/
 * build.rs:184-193) and CACHE_SIZE_T = int.
 */
#ifndef CINTX_HIP_H
#define CINTX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* 1. libcint raw signature                                            */
/* ------------------------------------------------------------------ */

/* Opaque optimizer handle; replaces CINTOpt (libcint-master/include/
 * cint.h.in:145-152) and cintx's RawOptimizerHandle
 * (crates/cintx-compat/src/optimizer.rs:15).  Built once per basis by any
 * *_optimizer below: the basis and the primitive-pair data of every shell
 * pair, resident on the device, which every raw call that passes it reuses
 * (optimizer.c:344-425).  A call without one builds its two pairs itself.
 *
 * Threading: any number of host threads may call the raw entry points at
 * once (sharing an optimizer is fine; it is immutable).  Each thread gets
 * its own stream and device arena on its first call; calls take no lock.
 * A call that fails writes zeros into its block and returns 0 (a message
 * goes to stderr); with CINTX_STRICT=1 in the environment it aborts. */
typedef struct CINTOpt CINTOpt;

/* libcint-master/src/cint2e.c:1186-1193 */
int int2e_sph(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
              double *env, CINTOpt *opt, double *cache);
/* libcint-master/src/autocode/grad2.c:81-88 */
int int2e_ip1_sph(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
                  double *env, CINTOpt *opt, double *cache);
/* libcint-master/src/cint3c2e.c:693-700 */
int int3c2e_sph(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
                double *env, CINTOpt *opt, double *cache);
/* libcint-master/src/autocode/int3c2e.c:80-87 */
int int3c2e_ip1_sph(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
                    double *env, CINTOpt *opt, double *cache);
/* libcint-master/src/cint2c2e.c:351-359 (the density-fitting metric (i|k);
 * shls holds two shells, out is [dk][di]) */
int int2c2e_sph(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
                double *env, CINTOpt *opt, double *cache);
/* libcint-master/src/cint2e.c:1202-1210, cint3c2e.c:709-717: Cartesian
 * output (ncart per shell, libcint's xyz order), no cart -> sph; they share
 * int2e_optimizer / int3c2e_optimizer */
int int2e_cart(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
               double *env, CINTOpt *opt, double *cache);
int int3c2e_cart(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
                 double *env, CINTOpt *opt, double *cache);
/* libcint-master/src/autocode/grad2.c:163-170 ((ij|nabla k l), out is
 * [3][dl][dk][dj][di]) */
int int2e_ip2_sph(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
                  double *env, CINTOpt *opt, double *cache);
/* libcint-master/src/autocode/int3c2e.c:161-168 ((ij|nabla k), out is
 * [3][dk][dj][di]) */
int int3c2e_ip2_sph(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
                    double *env, CINTOpt *opt, double *cache);

/* libcint-master/src/cint2e.c:1215-1224, cint3c2e.c:719-727: spinor
 * (two-component, spin-free) output.  out is double complex (re, im
 * interleaved) in libcint's layout: [dl][dk][dj][di] with
 * CINTcgto_spinor(shell) per slot (4l+2 for KAPPA_OF 0, 2l+2 for kappa < 0,
 * 2l for kappa > 0, times NCTR_OF); int3c2e_spinor's k stays spherical
 * (2l+1).  Bra spinors (i, k) are conjugated and the spin is summed per
 * electron, as c2s_sf_2e1/c2s_sf_2e2/c2s_sf_3c2e1 (cart2sph.c:5385-5528,
 * 5997-6050).  They share int2e_optimizer / int3c2e_optimizer. */
int int2e_spinor(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
                 double *env, CINTOpt *opt, double *cache);
int int3c2e_spinor(double *out, int *dims, int *shls, int *atm, int natm, int *bas, int nbas,
                   double *env, CINTOpt *opt, double *cache);

/* libcint-master/src/cint2e.c:1195, autocode/grad2.c:69,
 * cint3c2e.c:702, autocode/int3c2e.c:68 */
void int2e_optimizer(CINTOpt **opt, int *atm, int natm, int *bas, int nbas, double *env);
void int2e_ip1_optimizer(CINTOpt **opt, int *atm, int natm, int *bas, int nbas, double *env);
void int3c2e_optimizer(CINTOpt **opt, int *atm, int natm, int *bas, int nbas, double *env);
void int3c2e_ip1_optimizer(CINTOpt **opt, int *atm, int natm, int *bas, int nbas, double *env);
/* libcint-master/src/cint2c2e.c:360 */
void int2c2e_optimizer(CINTOpt **opt, int *atm, int natm, int *bas, int nbas, double *env);
/* libcint-master/src/autocode/grad2.c:145, autocode/int3c2e.c:143 */
void int2e_ip2_optimizer(CINTOpt **opt, int *atm, int natm, int *bas, int nbas, double *env);
void int3c2e_ip2_optimizer(CINTOpt **opt, int *atm, int natm, int *bas, int nbas, double *env);
/* libcint-master/src/optimizer.c:69 */
void CINTdel_optimizer(CINTOpt **opt);

/* ------------------------------------------------------------------ */
/* 2. cintx C shim (crates/cintx-capi/src/shim.rs, errors.rs)           */
/* ------------------------------------------------------------------ */

/* api ids: crates/cintx-capi/src/shim.rs:8-49.  The reference enum has no
 * int3c2e_sph or int3c2e_spinor entry; CINTX_API_INT3C2E_SPH and
 * CINTX_API_INT3C2E_SPINOR are this library's extensions.  For the spinor
 * ids out_len counts doubles (two per complex element). */
#define CINTX_API_INT2E_SPH 10
#define CINTX_API_INT2E_SPINOR 11
#define CINTX_API_INT3C2E_SPINOR 41
#define CINTX_API_INT3C2E_IP1_SPH 19
#define CINTX_API_INT2E_IP1_SPH 36
#define CINTX_API_INT3C2E_SPH 40
#define CINTX_API_INT2C2E_SPH 13
#define CINTX_API_INT2E_CART 9

/* status codes: crates/cintx-capi/src/errors.rs:9-21 */
#define CINTX_STATUS_SUCCESS 0
#define CINTX_STATUS_INVALID_INPUT 1
#define CINTX_STATUS_UNSUPPORTED_API 2
#define CINTX_STATUS_UNSUPPORTED_REPRESENTATION 3
#define CINTX_STATUS_BUFFER_TOO_SMALL 4
#define CINTX_STATUS_MEMORY_LIMIT_EXCEEDED 5
#define CINTX_STATUS_ALLOCATION_FAILED 6
#define CINTX_STATUS_EXECUTION_FAILED 7
#define CINTX_STATUS_NULL_POINTER 8
#define CINTX_STATUS_PANIC 9
#define CINTX_STATUS_BACKEND_NOT_COMPILED 10

/* crates/cintx-capi/src/shim.rs:51-66 */
typedef struct {
  size_t bytes;
  size_t alignment;
  size_t required_bytes;
  size_t chunk_count;
  size_t work_units;
} CintxWorkspaceQuery;

typedef struct {
  int32_t not0;
  size_t bytes_written;
  size_t workspace_bytes;
} CintxEvalSummary;

/* crates/cintx-capi/src/shim.rs:313-327 */
int cintrs_query_workspace(int api, const int *dims, size_t dims_len, const int *shls,
                           size_t shls_len, const int *atm, size_t atm_len, const int *bas,
                           size_t bas_len, const double *env, size_t env_len, const void *opt,
                           CintxWorkspaceQuery *query_out);
/* crates/cintx-capi/src/shim.rs:371-389; opt: NULL or a CINTOpt built for
 * the same atm/bas (see above) */
int cintrs_eval(int api, double *out, size_t out_len, const int *dims, size_t dims_len,
                const int *shls, size_t shls_len, const int *atm, size_t atm_len, const int *bas,
                size_t bas_len, const double *env, size_t env_len, const void *opt, double *cache,
                size_t cache_len, CintxEvalSummary *summary_out);
/* crates/cintx-capi/src/errors.rs:211-245 */
int cintrs_last_error_code(void);
void cintrs_clear_last_error(void);
size_t cintrs_copy_last_error_message(char *out, size_t out_len);

/* ------------------------------------------------------------------ */
/* 3. Batched plans (executor boundary)                                 */
/* ------------------------------------------------------------------ */

/* Family ids of a plan (the hot-path rows of crates/cintx-ops/
 * src/generated/api_manifest.csv:12, 22, 25, 225). */
#define CINTX_FAMILY_INT2E_SPH 0
#define CINTX_FAMILY_INT2E_IP1_SPH 1
#define CINTX_FAMILY_INT3C2E_SPH 2
#define CINTX_FAMILY_INT3C2E_IP1_SPH 3
/* further families served by the same kernels: the DF metric, the nabla-k
 * derivatives, Cartesian and spinor output (api_manifest.csv:11, 13, 15,
 * 24, 228, 276 and int3c2e_spinor).
 * A spinor plan's output (execute's out_dev, cintx_plan_copy_output) is
 * the complex libcint layout of every tuple directly, re/im interleaved;
 * cintx_plan_sizes counts doubles. */
#define CINTX_FAMILY_INT2C2E_SPH 4
#define CINTX_FAMILY_INT2E_IP2_SPH 5
#define CINTX_FAMILY_INT3C2E_IP2_SPH 6
#define CINTX_FAMILY_INT2E_CART 7
#define CINTX_FAMILY_INT3C2E_CART 8
#define CINTX_FAMILY_INT2E_SPINOR 9
#define CINTX_FAMILY_INT3C2E_SPINOR 10

/* precision modes: FP64 everywhere, or the mixed mode (FP32 recurrences and
 * primitive sums for tuples whose Schwarz bound is below the plan's
 * mixed_threshold, FP64 for the rest and for roots/weights/accumulation). */
#define CINTX_PREC_FP64 0
#define CINTX_PREC_MIXED 1

typedef struct cintx_plan cintx_plan;

/* Explicit tuple list.  tuples holds ntuples x 4 ints (3c2e: the 4th is
 * ignored).  Block t of the output holds the tuple's libcint-layout result
 * (natural dims, components leading) at element offset offsets[t]
 * (cintx_plan_offsets).  Replaces cintx-cubecl's per-tuple launch
 * (crates/cintx-cubecl/src/kernels/two_electron.rs:1309-1430).
 * CINTX_PREC_MIXED (int2e_sph and int3c2e_sph only) runs a tuple with the
 * FP32 recurrences when its Schwarz bound over the absolute contraction
 * coefficients, sqrt((ij|ij)) sqrt((kl|kl)) or sqrt((ij|ij)) sqrt((P|P)),
 * is below mixed_threshold (computed on the device at creation); the output
 * order stays the caller's. */
int cintx_plan_create(cintx_plan **plan, int family, const int *atm, int natm, const int *bas,
                      int nbas, const double *env, const int *tuples, int64_t ntuples,
                      int precision, double mixed_threshold);
/* Unique 8-fold int2e_sph tuples (i>=j, k>=l, ij>=kl) whose Schwarz bound
 * sqrt((ij|ij)) sqrt((kl|kl)) is at least `threshold`; bra pairs restricted to
 * every `bra_stride`-th screened pair (1 = all).  Tuples are generated on the
 * host and kept in HBM.  CINTX_PREC_MIXED puts a tuple in FP32 when its bound
 * taken over the absolute contraction coefficients (the size of the terms
 * the FP32 rounding follows) is below mixed_threshold and its class has no
 * one-lane register kernel (FP32 pays only where g is shared through LDS on
 * CDNA4). */
int cintx_plan_create_screened(cintx_plan **plan, const int *atm, int natm, const int *bas,
                               int nbas, const double *env, double threshold, int bra_stride,
                               int precision, double mixed_threshold);
int cintx_plan_destroy(cintx_plan *plan);
/* total output elements and tuples */
int cintx_plan_sizes(const cintx_plan *plan, int64_t *nelem, int64_t *ntuples);
/* tuples evaluated in FP32 (mixed plans keep them after the FP64 ones) and
 * the number of class bins (kernel launches per execute) */
int cintx_plan_info(const cintx_plan *plan, int64_t *ntuples_fp32, int64_t *nbins);
/* host copies of the tuple list (ntuples x 4) and offsets (ntuples) */
int cintx_plan_tuples(const cintx_plan *plan, int *tuples, int64_t *offsets);
/* run all class bins on the plan's stream into device memory `out_dev`
 * (NULL: the plan's own buffer); asynchronous */
int cintx_plan_execute(cintx_plan *plan, double *out_dev);
int cintx_plan_sync(cintx_plan *plan);
int cintx_plan_copy_output(cintx_plan *plan, double *host_out);
/* libcint-layout blocks of the selected tuples only, packed in the order of
 * `which`; offsets[k] receives block k's element offset. */
int cintx_plan_copy_tuples(cintx_plan *plan, const int64_t *which, int64_t n, double *host_out,
                           int64_t *offsets);
/* Record the plan's launches into a hipGraph once and replay it. */
int cintx_plan_execute_graph(cintx_plan *plan, double *out_dev);
/* Timing of the last execute (HIP events on the plan's stream), the
 * per-kernel-launch records, and the algorithmic work of the plan
 * (surviving primitive tuples, root evaluations, FLOPs). */
int cintx_plan_last_time_ms(const cintx_plan *plan, double *ms);
/* Execute with HIP events around every class bin (no graph), synchronously;
 * then cintx_plan_bin_stats reports per bin: class (li,lj,lk,ll, precision,
 * runtime-class kernel flag; 6 ints), tuples, milliseconds of the last timed
 * execute, and the surviving primitive tuples / algorithmic FLOPs of the last
 * cintx_plan_work. */
int cintx_plan_execute_timed(cintx_plan *plan, double *out_dev);
int cintx_plan_bin_stats(const cintx_plan *plan, int64_t nmax, int *cls, int64_t *ntuples,
                         double *ms, double *prim_tuples, double *flops);
int cintx_plan_work(cintx_plan *plan, double *prim_tuples, double *root_evals, double *flops);
/* Launch-shape autotuner (no counterpart in the reference, whose executor
 * has a single launch shape; crates/cintx-runtime/src/dispatch.rs:187-195).
 * Every class has up to six compiled variants (lanes per tuple, primitive
 * tuples per step); cintx_plan_autotune times each on the plan's own bin
 * (min of `reps` launches after one warm-up, HIP events), keeps the fastest
 * per bin and records it for the process, so later plans of the same class
 * start from it.  Results of all variants agree to rounding (summation order
 * of the primitive loop differs).  The FP32 bins of a mixed plan also try the
 * FP64 instantiations (variants 6..11 = FP64 variant 0..5): FP32 without packed
 * math issues at the FP64 rate on CDNA4 and wins only where it saves
 * registers or LDS bytes.  cintx_plan_set_variant forces one variant (0..5)
 * on every bin that has it (variant 0 otherwise); cintx_plan_bin_variants
 * reports per bin: variant in use (-1: runtime-class kernel), variants
 * available, lanes per tuple, primitive tuples per step (for one lane per
 * tuple: row chunks), tuples per workgroup. */
int cintx_plan_autotune(cintx_plan *plan, int reps);
int cintx_plan_set_variant(cintx_plan *plan, int variant);
int cintx_plan_bin_variants(const cintx_plan *plan, int64_t nmax, int *variant, int *nvariants,
                            int *lanes, int *prims, int *tuples_per_block);

/* The n-point Rys rule (roots u = t^2/(1-t^2), weights) that the kernels use,
 * evaluated for nx arguments on the host or on the device; replaces
 * CINTrys_roots (libcint-master/src/rys_roots.c:57-123).  u, w: nx x n. */
int cintx_rys_eval_host(int n, const double *x, int64_t nx, double *u, double *w);
int cintx_rys_eval_device(int n, const double *x, int64_t nx, double *u, double *w);

/* Coulomb and exchange matrices from an int2e_sph plan whose tuples are
 * unique 8-fold shell quartets (each unordered quartet once) of segmented
 * shells: J_pq = sum_rs (pq|rs) D_rs, K_pr = sum_qs (pq|rs) D_qs for a
 * symmetric density D (nao x nao, row-major, host memory; a non-symmetric D
 * is replaced by (D + D^T) / 2).  The class kernels digest their blocks
 * into J and K in their epilogue (no ERI block of a compiled class reaches
 * memory; runtime-class bins are written and read back).  No counterpart in
 * the reference's API: the consumer of the unique-quartet workload whose
 * loop libcint-master/examples/time_c60.c:196-227 shows. */
int cintx_plan_jk(cintx_plan *plan, const double *D, double *J, double *K);

/* The same J/K build for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc);
 * no counterpart in the reference.  D_dev, J_dev, K_dev are nao x nao
 * row-major device buffers; the call allocates nothing: the symmetrised D is
 * kept in a static device scratch of CINTX_JK_NOALLOC_MAX_BYTES per device
 * (nao^2 * 8 bytes above it: CINTX_STATUS_MEMORY_LIMIT_EXCEEDED), J_dev and
 * K_dev are cleared by a kernel, and the plan's own ERI buffer and workspace
 * (allocated when the plan was created) are reused.  Enqueued on the plan's
 * stream; returns without waiting (cintx_plan_sync waits).  Calls that
 * overlap on one GPU share the scratch and race. */
#define CINTX_JK_NOALLOC_MAX_BYTES ((size_t)64 << 20)
int cintx_plan_jk_noalloc(cintx_plan *plan, const double *D_dev, double *J_dev, double *K_dev);

/* Learned Rys initializer (no counterpart in the reference): a 2x64 tanh
 * MLP mapping (log x, nroots) to initial roots/weights, trained on the device
 * by this library's forward/backward/Adam kernels on rules produced by the
 * table solver, polished by FP64 Newton steps on the moment equations, with
 * the table rule as fallback.  params: cintx_rys_mlp_nparam() floats. */
int cintx_rys_mlp_nparam(void);
int cintx_rys_mlp_train(float *params, int64_t nsamples, int epochs, int batch, float lr,
                        uint64_t seed, double *final_loss);
/* batch-mean loss gradient for given inputs z (batch x 2), nroots and
 * targets (batch x 26: ln u_r then ln w_r); loss per sample */
int cintx_rys_mlp_grad(const float *params, const float *z, const int *nroots,
                       const float *targets, int batch, float *grad, float *loss);
/* init_mode 0: network, 1: asymptotic rule (the initializer without the
 * network), 2: the table rule alone (no Newton; timing reference), 3: the
 * table rule as the start (the floor of the Newton polish itself).  After
 * newton_steps FP64 Newton steps on the moment equations the table rule
 * (the plain solver) judges the polished rule: status[i] = 1 when every
 * root u_r and weight w_r is within 1e-14 relative of it (the polished rule
 * is returned), 0 when the table rule was substituted.  polish_err
 * (nullable): that largest relative difference (inf when Newton left the
 * domain, -1 for init_mode 2).  ms: device time of the evaluation. */
int cintx_rys_mlp_eval(const float *params, int n, const double *x, int64_t nx,
                       int newton_steps, int init_mode, double *u, double *w, int *status,
                       double *polish_err, double *ms);

/* The Cartesian -> spinor coupling coefficients this library applies for
 * shell angular momentum l (0..4) and KAPPA_OF kappa: CINTlen_spinor rows
 * (m_j = -j..j; kappa = 0: the j = l-1/2 rows, then the j = l+1/2 rows),
 * each 2 ncart(l) values, the alpha then the beta component; re/im may be
 * NULL.  Returns the row count, -1 for l outside 0..4.  Host-only (no
 * device use).  Same values as libcint's g_trans_cart2jR/g_trans_cart2jI
 * (cart2sph.c:3561) and crates/cintx-cubecl/src/transform/
 * c2spinor_coeffs.rs, generated here from the real solid harmonics and the
 * Clebsch-Gordan coefficients. */
int cintx_spinor_coeffs(int l, int kappa, double *re, double *im);

/* Fill the plan's own output buffer with `value` on the plan's stream
 * (asynchronous; a kernel, so it may be captured in a graph).  No counterpart
 * in the reference: lets a caller poison the buffer (NaN) before a replay to
 * see that every block is rewritten. */
int cintx_plan_fill_output(cintx_plan *plan, double value);

/* Device memory helpers for callers without a runtime of their own. */
int cintx_device_alloc(double **ptr, int64_t nelem);
int cintx_device_free(double *ptr);
/* synchronous copies of nelem doubles host -> device and device -> host */
int cintx_device_upload(double *dst_dev, const double *src_host, int64_t nelem);
int cintx_device_download(double *dst_host, const double *src_dev, int64_t nelem);
int cintx_device_count(int *n);

#ifdef __cplusplus
}
#endif
#endif
