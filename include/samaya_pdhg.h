/* C ABI of the MI355X (gfx950) PDHG device layer: the drop-in for samaya's gpu::PdhgDevice seam
 * and its CSR SpMV. Plain C: opaque handles, pointers and sizes; every call returns an int
 * status (SMY_OK = 0) and never unwinds. smy_last_error() returns the message of the calling
 * thread's last failure.
 *
 * Reference interfaces each entry point replaces (paths relative to the samaya tree):
 *   class gpu::PdhgDevice ........ src/gpu/pdhg_core.hpp:24-79 (implementation pdhg_core.cu)
 *   gpu::csr_spmv ................ src/gpu/spmv.hpp:16-21
 *   gpu::spmv_threads_per_row .... src/gpu/spmv.hpp:23
 *   gpu::gpu_available/device_name src/gpu/device.hpp:34-63
 */
#ifndef SAMAYA_PDHG_H
#define SAMAYA_PDHG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SMY_OK = 0,
  SMY_NO_DEVICE = 1,    /* no usable gfx950 GPU: PdhgDevice::create returned null */
  SMY_TOO_LARGE = 2,    /* nnz >= INT_MAX (32-bit CSR offsets): create returned null */
  SMY_ERROR = 3,        /* a HIP failure (driver, out of memory): the reference throws */
  SMY_INVALID = 4,      /* bad argument */
  SMY_STATE = 5         /* call out of order (e.g. set_original missing before evaluate) */
};

enum { SMY_FP64 = 0, SMY_FP32 = 1 }; /* iterate and matrix-value precision of the step kernels */
enum { SMY_TILE_BINS = 3 };           /* size bins of smy_pdhg_get_bin_tile / set_bin_tile */
enum { SMY_CURRENT = 0, SMY_AVERAGE = 1 }; /* PdhgDevice::Iterate (pdhg_core.hpp:53) */

/* A compressed sparse column matrix as samaya's SparseMatrix holds it
 * (include/samaya/sparse_matrix.hpp:36-69): start[cols + 1] (64-bit), index[nnz], value[nnz],
 * row indices sorted within each column. */
typedef struct smy_csc {
  int32_t rows;
  int32_t cols;
  const int64_t* start;
  const int32_t* index;
  const double* value;
} smy_csc;

typedef struct smy_pdhg smy_pdhg;

/* Replaces gpu::gpu_available (device.hpp:34-48): 1 when a gfx950 device is present. */
int smy_gpu_available(void);
/* Replaces gpu::device_name (device.hpp:51-63): writes the device name (empty if none). */
int smy_gpu_device_name(char* buf, size_t len);
/* Message of this thread's last failed call ("" if none). */
const char* smy_last_error(void);

/* Replaces PdhgDevice::create (pdhg_core.hpp:27-30, pdhg_core.cu:242-273): uploads A (m x n,
 * CSC), its transpose At (CSC, n x m) and cost/lower/upper (n + m each, equilibrated space).
 * precision SMY_FP32 also keeps FP32 copies of the values and iterates for the step kernels;
 * evaluation sums stay FP64 either way. *out is NULL unless SMY_OK is returned. */
int smy_pdhg_create(const smy_csc* A, const smy_csc* At, const double* cost, const double* lower,
                    const double* upper, int precision, smy_pdhg** out);
/* Replaces ~PdhgDevice. */
void smy_pdhg_destroy(smy_pdhg* h);
/* Replaces PdhgDevice::set_original (pdhg_core.hpp:49-51): equilibration scales and the cost
 * and bounds before equilibration (n + m), used by evaluate. */
int smy_pdhg_set_original(smy_pdhg* h, const double* col_scale, const double* row_scale,
                          const double* cost, const double* lower, const double* upper);
/* Replaces PdhgDevice::step (pdhg_core.hpp:36): v_new, y_new and the averages with weight
 * 1 / (iter + 1). Asynchronous. */
int smy_pdhg_step(smy_pdhg* h, double tau, double sigma, double theta, long long iter);
/* Replaces PdhgDevice::advance (pdhg_core.hpp:38): v_prev <- v <- v_new, y <- y_new. */
int smy_pdhg_advance(smy_pdhg* h);
/* Replaces PdhgDevice::restart (pdhg_core.hpp:40): v_prev <- v. */
int smy_pdhg_restart(smy_pdhg* h);
/* Replaces PdhgDevice::download (pdhg_core.hpp:42-44): copies iterates to host arrays (n + m or
 * m doubles each); any pointer may be NULL. Synchronizes. */
int smy_pdhg_download(smy_pdhg* h, double* v, double* v_new, double* y, double* y_new,
                      double* v_erg, double* y_erg);
/* Replaces PdhgDevice::evaluate (pdhg_core.hpp:67): out[7] = ||Ax-r||^2, dual residual^2,
 * primal objective, dual objective, ||x||^2, ||r||^2, ||y||^2 of the current or average iterate,
 * summed in FP64 in a fixed order. Synchronizes. */
int smy_pdhg_evaluate(smy_pdhg* h, int which, double out[7]);
/* Replaces PdhgDevice::restart_distance (pdhg_core.hpp:70). Synchronizes. */
int smy_pdhg_restart_distance(smy_pdhg* h, int which, double* dv_sq, double* dy_sq);
/* Replaces PdhgDevice::restart_to (pdhg_core.hpp:73). */
int smy_pdhg_restart_to(smy_pdhg* h, int which);

/* Additions without a reference counterpart (the host solver's fast path):
 * evaluate both iterates in one pass over A and A' (out[0..6] current, out[7..13] average);
 * run `count` steps, advancing after each but the last (iterations iter0 ... iter0+count-1
 * since the last restart), from a captured HIP graph when count == the graph length. */
int smy_pdhg_evaluate_both(smy_pdhg* h, double out[14]);
int smy_pdhg_run(smy_pdhg* h, double tau, double sigma, double theta, long long iter0,
                 int count);
/* Switches the step precision (SMY_FP64 / SMY_FP32), converting the device state. */
int smy_pdhg_set_precision(smy_pdhg* h, int precision);
/* Uploads a primal-dual starting point (v: n + m, y: m; scaled space) as current iterate, restart
 * point and average (warm start). */
int smy_pdhg_set_start(smy_pdhg* h, const double* v, const double* y);
/* Kernel time of the last smy_pdhg_run, in milliseconds (HIP events on the handle's stream). */
int smy_pdhg_last_run_ms(smy_pdhg* h, double* ms);

/* The enhanced algorithm's preconditioner computed and applied on the device (no reference
 * counterpart: the reference equilibrates on the host, src/lp/pdlp.cpp:109-209,284-318). For a
 * single-LP handle created from the unscaled LpProblem, before its first step: `passes`
 * infinity-norm Ruiz rounds (every row and column divided by the square root of its largest
 * magnitude, from the same current scaling) and, with pock_chambolle, one Pock-Chambolle step
 * (alpha = 1: square roots of the 1-norms); then the handle's matrices, cost and bounds are
 * scaled in place (FP32 copies included) and 20 power iterations from the reference's seeded start
 * bound ||[A_eq -I]||. Writes col_scale (n), row_scale (m), out = {that bound, max column 2-norm
 * of the unscaled A, max row 2-norm}. The factors and the norms equal smy_pdlp_equilibrate's bit
 * for bit (the same per-row sums in the same order); the bound agrees to rounding. */
int smy_pdhg_equilibrate(smy_pdhg* h, int passes, int pock_chambolle, double* col_scale,
                         double* row_scale, double out[3]);
/* The same on the device for the reference's own preconditioner, ruiz_equilibration
 * (src/lp/pdlp.cpp:109-187, applied as :190-209,284-299): up to `passes` full-step updates of every
 * column and row by the inverse of its current 2-norm (both norms from the same factors), clamped
 * to [1e-8, 1e8], stopping after the first pass that moves no factor by more than 10%; then the
 * handle is scaled and the operator norm bounded as above. The factors and norms equal
 * smy_pdlp_equilibrate's with SMY_ALGORITHM_REFERENCE bit for bit. */
int smy_pdhg_equilibrate_reference(smy_pdhg* h, int passes, double* col_scale, double* row_scale,
                                   double out[3]);

/* Test hook for the reference's GPU-failure path (src/lp/pdlp.cpp:228-246,341-360): once
 * the handle's step entry points (smy_pdhg_step, _run, _run_batch, _run_adaptive) would run past
 * step `step_no` (counted over the handle's life), they return SMY_ERROR with the message
 * "injected fault at step N (test hook)" and launch nothing; the handle stays valid and can be
 * destroyed. A host-side failure, not a GPU fault. step_no < 0 turns it off. */
int smy_pdhg_inject_fault(smy_pdhg* h, long long step_no);

/* Measurement helper: average milliseconds of the dual (rows of A) and primal (rows of A') step
 * kernels over `reps` back-to-back launches on the handle's stream. Changes the iterates. */
int smy_pdhg_time_kernels(smy_pdhg* h, int reps, double* dual_ms, double* primal_ms);

/* ---- Batches: many LPs in one handle (no reference counterpart; the reference solves one LP
 * per PdhgDevice). The `count` instances (at most 16384) are stacked block-diagonally and step
 * together, two kernel launches per step and precision for the whole batch; every instance keeps
 * its own step sizes, step count since restart, restart point, precision, and can be frozen once
 * its solve has ended. An instance's arithmetic is the same as in a handle of its own: a batch
 * of K gives each LP the bits smy_pdhg_create would give it. The single-LP functions above take
 * only handles of one instance; advance, last_run_ms, time_kernels and destroy take both. ---- */

/* A[i] (m_i x n_i CSC), At[i] its transpose, cost[i] / lower[i] / upper[i] (n_i + m_i). */
int smy_pdhg_create_batch(int count, const smy_csc* A, const smy_csc* At,
                          const double* const* cost, const double* const* lower,
                          const double* const* upper, int precision, smy_pdhg** out);
/* The shared-matrix batch: `count` LPs with one matrix A (m x n CSC, At its transpose) and one
 * cost (n + m), each with its own bounds lower[i] / upper[i] (n + m; the m row entries must be
 * equal for all instances: SMY_INVALID otherwise). Branch-and-bound node relaxations of one LP.
 * Every iterate is stored instance-minor (row-major variables x instances), so each product with
 * A or A' is one SpMM for the whole batch: the matrix is read once per step, and every gather
 * reads the contiguous values of all instances. Driven by the batch functions below; the
 * instances' sums are taken in another order than by smy_pdhg_create_batch (equal up to
 * rounding, bit-identical for identical instances). */
int smy_pdhg_create_shared(const smy_csc* A, const smy_csc* At, int count, const double* cost,
                           const double* const* lower, const double* const* upper, int precision,
                           smy_pdhg** out);
/* Number of instances (0 for NULL). */
int smy_pdhg_batch_size(smy_pdhg* h);
/* set_original / set_start / download for instance i (vectors of that instance's sizes). */
int smy_pdhg_set_original_at(smy_pdhg* h, int instance, const double* col_scale,
                             const double* row_scale, const double* cost, const double* lower,
                             const double* upper);
int smy_pdhg_set_start_at(smy_pdhg* h, int instance, const double* v, const double* y);
int smy_pdhg_download_at(smy_pdhg* h, int instance, double* v, double* v_new, double* y,
                         double* y_new, double* v_erg, double* y_erg);
/* `count` steps of every instance that is not frozen, instance i with tau[i], sigma[i] and its
 * iterations since restart starting at iter0[i] (arrays of batch size). */
int smy_pdhg_run_batch(smy_pdhg* h, const double* tau, const double* sigma, double theta,
                       const long long* iter0, int count);
/* out[14 i + q]: smy_pdhg_evaluate_both of instance i (meaningless for frozen instances). */
int smy_pdhg_evaluate_batch(smy_pdhg* h, double* out);
/* which[i]: -1 (skip), SMY_CURRENT or SMY_AVERAGE; out[2 i], out[2 i + 1]: squared distances of
 * that candidate from instance i's restart point (v, then y). */
int smy_pdhg_restart_distance_batch(smy_pdhg* h, const int* which, double* out);
/* Restarts instance i at its candidate which[i] (-1: no restart). */
int smy_pdhg_restart_to_batch(smy_pdhg* h, const int* which);
/* precision[i]: SMY_FP64 / SMY_FP32 (state converted when it changes) or -1 (unchanged). */
int smy_pdhg_set_precision_batch(smy_pdhg* h, const int* precision);
/* Instance i takes no further steps and is not evaluated; its state stays downloadable. */
int smy_pdhg_freeze(smy_pdhg* h, int instance);

/* Device-side infeasibility and unboundedness tests (the reference downloads the iterates and
 * runs them on the host: src/lp/pdlp.cpp:715-1043) for the instances with flags[i] > 0, in two
 * device phases. out[11 i + q] for those instances: 0 sum |w| and 1 max |yf| (yf = row-scaled
 * average dual, w = (A' yf, -yf)); 2 f_min and 3 f_max of the Farkas bound test, finite parts, with
 * 4 and 5 the counts of their infinite terms; 6 scale (max |t bound|, |t| > 1e-9, t = +-w/y_max);
 * 7 d_max = max |v_new - v|; 8 ||A x - r||^2 of the normalized difference; 9 its cost slope; 10 the
 * number of bound failures of the projected ray and of A times it. The host decides as the
 * reference does from these. Other Farkas candidates (no reference counterpart; a y that passes
 * the test is a Farkas certificate all the same): flags[i] = 2 or 3, the sums of the dual's sign
 * projection, its entries that make a row's term of f_min (2) or f_max (3) infinite set to 0;
 * flags 5, 6, 7: the same for the dual's change since the restart point, y_erg - y_rst, as it is
 * (5) or projected (6, 7). The ray test is skipped, its slots 7-10 left 0, when no instance with
 * flag 1 has a structural variable with an infinite bound (a bounded LP has no improving ray).
 * Flags 4 and above 7 are invalid. Synchronizes. */
int smy_pdhg_detect_batch(smy_pdhg* h, const int* flags, double* out);
/* The projected ray (n_i structural entries) of instance i's last detection. */
int smy_pdhg_download_ray(smy_pdhg* h, int instance, double* ray);
/* The dual (m_i entries, the y_erg space) that the last detection with a flag above 1 tested for
 * instance i, when that detection flagged instance i (its contents are unspecified otherwise): the
 * Farkas certificate when its test passed. No reference counterpart. */
int smy_pdhg_download_farkas(smy_pdhg* h, int instance, double* y);

/* Adaptive step sizes (no reference counterpart: the reference's adaptive step is disabled code,
 * src/lp/pdlp.cpp:1051-1067). PDLP's rule (Applegate et al., NeurIPS 2021, sec. 3.1): an attempt
 * is one PDHG step with tau = eta / omega, sigma = eta * omega in the primal-first order; with the
 * moves dv, dy it is accepted when eta <= (omega ||dv||^2 + ||dy||^2 / omega) / (2 |dy' [A -I] dv|),
 * else retried from the same point; either way eta becomes
 * min((1 - (k+1)^-0.3) * that limit, (1 + (k+1)^-0.6) * eta), k = attempts so far. The averages
 * weight each accepted iterate by its eta. The decisions are taken on the device. */
typedef struct smy_pdhg_adaptive {
  double eta;          /* step size of the next attempt */
  double omega;        /* primal weight */
  double weight;       /* sum of the accepted step sizes since the restart (0 after a restart:
                          the next accepted iterate becomes the average) */
  long long attempts;  /* attempts since the start of the solve */
  long long rejected;  /* of which rejected */
} smy_pdhg_adaptive;

/* `count` attempts of every instance that is not frozen, from state[i] (batch size), which is
 * updated. Afterwards the handle is laid out as after `count` smy_pdhg_run_batch steps: v_new /
 * y_new hold the current (last accepted) iterate, v / y the point before it (or the last
 * rejected attempt), the averages their step-size weighted means; smy_pdhg_advance and the other
 * calls apply unchanged. Not available for the shared-matrix batch. Synchronizes. */
int smy_pdhg_run_adaptive(smy_pdhg* h, smy_pdhg_adaptive* state, int count);
/* Measurement helper: average milliseconds of the adaptive attempt's dual (rows of A), primal
 * (rows of A') and decision kernels over `reps` launches each. Changes the state. */
int smy_pdhg_time_adaptive_kernels(smy_pdhg* h, int reps, double* dual_ms, double* primal_ms,
                                   double* decide_ms);

/* Restarted Halpern PDHG with reflection (no reference counterpart: the reference's PDLP is
 * restarted PDHG with averages, src/lp/pdlp.cpp:574-641). With on != 0, smy_pdhg_run_batch and
 * smy_pdhg_step take, from z = (v, y) and the restart point z0 as anchor, the steps
 *   T(z): v+ = clamp(v - tau (c + [A -I]'y)), y+ = y + sigma [A -I] (2 v+ - v)
 *   z'   = lambda ((1 + reflection) T(z) - reflection z) + (1 - lambda) z0,
 * lambda = (k + 1) / (k + 2), k = iter0 + the step's index (steps since the restart); theta is
 * ignored. After a run, v_new / y_new hold the last z', v / y the z it was taken from, and the
 * averages (v_erg / y_erg, SMY_AVERAGE) hold T(z) of that last step: smy_pdhg_evaluate_batch
 * measures T(z) into both halves (z' may lie outside the bounds and is no answer),
 * smy_pdhg_restart_distance_batch with SMY_AVERAGE gives ||T(z) - z0||^2, and
 * smy_pdhg_restart_to_batch with SMY_AVERAGE restarts at T(z), which becomes the anchor.
 * reflection in [0, 1]. Both batch layouts (stacked and shared-matrix). */
int smy_pdhg_set_halpern(smy_pdhg* h, int on, double reflection);
/* After a run of Halpern steps, per instance: out[2i] = ||v+ - v||^2, out[2i + 1] = ||y+ - y||^2
 * of the run's last step (the fixed-point residual z - T(z) of the restart test). Synchronizes. */
int smy_pdhg_fixed_point_batch(smy_pdhg* h, double* out);

/* The restart rule on the device (the reference decides on the host from downloaded sums:
 * src/lp/pdlp.cpp:1075-1132). smy_pdhg_rule_init sets each instance's normalizers (the measures
 * of src/lp/pdlp.cpp:493-519: rel primal = sqrt(||Ax-r||^2) / (1 + a_colmax ||x|| + ||r||), rel
 * dual = sqrt(dual residual^2) / (1 + at_rowmax ||y|| + c_norm), both 0 without rows) and its
 * starting primal weight. After smy_pdhg_evaluate_batch, smy_pdhg_rule_decide takes the decision
 * of every instance with in[i].due from the device-resident sums (candidate = the current or
 * average iterate with the smaller weighted KKT error; restart on 0.2 / 0.8-and-rising / 36% of
 * iter1, or in[i].force), measures the distances, updates the primal weight and restarts, all on
 * the device; out[i] reports the decision (which: -1 none, SMY_CURRENT, SMY_AVERAGE), the
 * candidate's error and the primal weight. Both batch layouts. Synchronizes once at the end. */
typedef struct smy_pdhg_rule_setup {
  double a_colmax;
  double at_rowmax;
  double c_norm;
  int has_rows;
  double omega;
} smy_pdhg_rule_setup;
typedef struct smy_pdhg_rule_in {
  int due;
  int force;
  long long since; /* iterations since the instance's last restart */
} smy_pdhg_rule_in;
typedef struct smy_pdhg_rule_out {
  int which;
  double kkt_candidate;
  double omega;
} smy_pdhg_rule_out;
int smy_pdhg_rule_init(smy_pdhg* h, const smy_pdhg_rule_setup* init);
int smy_pdhg_rule_decide(smy_pdhg* h, const smy_pdhg_rule_in* in, long long iter1,
                         smy_pdhg_rule_out* out);

/* Product tiles (no reference counterpart; the reference picks threads per row from nnz / rows,
 * src/gpu/spmv.cu:71-77). A tile maps the row blocks of the stacked batch's A (matrix 0) or A'
 * (matrix 1) with values of one precision to workgroups: per_wg (1, 2, 4) consecutive blocks per
 * workgroup, and with xcd = 1 one contiguous range of blocks per XCD. The first run in a
 * precision times every tile on that matrix's product and keeps the fastest; since each row is
 * summed inside its own block the same way under every tile, the choice never changes a result.
 * get returns SMY_STATE before the tuner ran; set fixes one tile for all blocks (the tuner then
 * leaves it).
 *
 * Size bins (a tile per size bin of a batch; no reference counterpart). A stacked batch's
 * instances fall into SMY_TILE_BINS bins by their row blocks in the matrix (about 1,024
 * nonzeros per block): bin 0 under 64 blocks, bin 1 under 512, bin 2 the rest. The tuner also
 * times each tile on each bin's blocks alone, and launches the bins apart, each with its fastest
 * tile, when that beats one launch of all blocks by 5%. get_bin_tile reports bin `bin`'s tile
 * and whether the bins are launched apart (*split); set_bin_tile fixes one bin's tile and turns
 * the split on (the other bins keep theirs, the plain tile unless set). No tile or split changes
 * a result. */
int smy_pdhg_get_tile(smy_pdhg* h, int matrix, int precision, int* per_wg, int* xcd);
/* FP32 steps with FP64 accumulation (no reference counterpart; the reference has no FP32 path).
 * With on = 1, the FP32 step kernels (constant and adaptive) sum each product in FP64 and form
 * the update that uses it in FP64, rounding once to FP32; values, iterates and every stored
 * vector stay FP32. Off (0) by default: FP32 accumulation. FP64 steps are unchanged. Stacked
 * batches only (SMY_INVALID for the shared-matrix batch). */
int smy_pdhg_set_fp64_accumulation(smy_pdhg* h, int on);

int smy_pdhg_set_tile(smy_pdhg* h, int matrix, int precision, int per_wg, int xcd);
int smy_pdhg_get_bin_tile(smy_pdhg* h, int matrix, int precision, int bin, int* per_wg, int* xcd,
                          int* split);
int smy_pdhg_set_bin_tile(smy_pdhg* h, int matrix, int precision, int bin, int per_wg, int xcd);

/* Replaces gpu::csr_spmv (spmv.hpp:16-21): y = A x for a CSR matrix in device memory (rows x ?,
 * ptr[rows + 1], idx[nnz], val[nnz]), x and y device arrays. threads_per_row is accepted for
 * the reference's signature; the gfx950 kernel picks its lane groups per row block. Runs on
 * `stream` (NULL: the default stream) and synchronizes it. */
int smy_csr_spmv(int rows, const int* ptr, const int* idx, const double* val, const double* x,
                 double* y, void* stream, int threads_per_row);
/* The same with FP32 values, x and y (FP32 accumulation). */
int smy_csr_spmv_f32(int rows, const int* ptr, const int* idx, const float* val, const float* x,
                     float* y, void* stream);
/* No counterpart in the reference: smy_csr_spmv / smy_csr_spmv_f32 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). y = A x with values, x and y of `precision` (SMY_FP64 / SMY_FP32), the same
 * kernels and row-block schedule, so the same bits as those two. It allocates nothing: the
 * schedule is built in a fixed per-GPU workspace of the library (about 8 MiB, held until the
 * process ends), which limits rows to 1,048,576 (SMY_TOO_LARGE beyond). Asynchronous on `stream`
 * (NULL: the default stream). Every call on one GPU shares that workspace: calls that overlap on
 * one GPU (on different streams or threads) race; calls on one stream are ordered. */
int smy_csr_spmv_noalloc(int rows, const int* ptr, const int* idx, const void* val, const void* x,
                         void* y, int precision, void* stream);
/* Replaces gpu::spmv_threads_per_row (spmv.hpp:23): smallest power of two in [2, 32] not below
 * the mean row length (the reference's CUSP rule, kept for its known answers). */
int smy_spmv_threads_per_row(long long nnz, int rows);

/* A prepared SpMV plan over a CSR matrix already in device memory (the row-block schedule is
 * built once): used to time the product alone. */
typedef struct smy_spmv_plan smy_spmv_plan;
int smy_spmv_plan_create(int rows, const int* ptr_host, const int* ptr_dev, const int* idx_dev,
                         const void* val_dev, int precision, smy_spmv_plan** out);
int smy_spmv_plan_run(smy_spmv_plan* p, const void* x_dev, void* y_dev, void* stream);
void smy_spmv_plan_destroy(smy_spmv_plan* p);

#ifdef __cplusplus
}
#endif

#endif /* SAMAYA_PDHG_H */
