/* C ABI of the PDLP host solver (C++) that drives the gfx950 device layer of samaya_pdhg.h.
 * Plain C, never unwinds; status codes are samaya's SimplexStatus / Status values as noted.
 *
 * Reference interfaces each entry point replaces (paths relative to the samaya tree):
 *   samaya::solve_pdlp(LpProblem, PdlpOptions, Logger) ... src/lp/pdlp.hpp:35 (options :11-16,
 *                                                          result :19-30)
 *   solve_pdlp_lp (LP by PDLP through Solver::solve) ..... src/core/solver.cpp:392-477, reached
 *                                                          from include/samaya/solver.hpp:75-86
 *                                                          with Params::lp_method = kPdlp and
 *                                                          Params::use_gpu (params.hpp:34-36)
 *   samaya_solve ......................................... include/samaya_c.h:41
 *   samaya_model_num_rows / _cols / _nonzeros, _free ..... include/samaya_c.h:33-37
 *   verify_lp_optimality / _infeasibility / _unbounded_ray include/samaya/verify.hpp:42-67
 *   samaya::gpu_info (GpuInfo) ........................... include/samaya/solver.hpp:93-97
 */
#ifndef SAMAYA_PDLP_H
#define SAMAYA_PDLP_H

#include <stddef.h>
#include <stdint.h>

#include "samaya_pdhg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SimplexStatus (src/lp/simplex.hpp:28-35). */
enum {
  SMY_STATUS_OPTIMAL = 0,
  SMY_STATUS_INFEASIBLE = 1,
  SMY_STATUS_UNBOUNDED = 2,
  SMY_STATUS_ITERATION_LIMIT = 3,
  SMY_STATUS_TIME_LIMIT = 4,
  SMY_STATUS_NUMERICAL_ERROR = 5
};

enum { SMY_PRECISION_FP64 = 0, SMY_PRECISION_MIXED = 1 };
/* Enhanced: infinity-norm Ruiz + Pock-Chambolle preconditioning in place of the reference's
 * 2-norm equilibration variant; every other decision unchanged. */
enum { SMY_ALGORITHM_REFERENCE = 0, SMY_ALGORITHM_ENHANCED = 1 };

/* Synthetic instance families (deterministic, seeded mt19937_64 with the library's own uniform
 * and Box-Muller transforms). */
enum {
  SMY_GEN_SPARSE_UNIFORM = 1,   /* p0 rows, p1 cols, p2 nnz per column at uniform distinct rows,
                                   values N(0,1); half equality rows, half <=; feasible x0 */
  SMY_GEN_SPARSE_POWERLAW = 2,  /* p0 rows, p1 cols, column counts power law on [2, p2] with mean
                                   p3 (max realized), values N(0,1) x 10^U[-2,2] */
  SMY_GEN_BLENDING = 3,         /* multi-period blending LP: p0 periods (0: U{4..48}), p1 units,
                                   p2 streams, p3 qualities (0: U{1..30}); staircase structure */
  SMY_GEN_KNAPSACK = 4          /* multi-dimensional knapsack LP: maximize profit'x (U{1..100}),
                                   p0 rows W x <= 0.4 row sums, p1 columns in [0, 1], p2 weights
                                   U{1..100} per column at distinct uniform rows */
};

typedef struct smy_model smy_model;

/* A model min/max obj'x + offset s.t. row_lower <= A x <= row_upper, col_lower <= x <= col_upper
 * (include/samaya/model.hpp:41-80, LP fields); A is CSC (start[n+1], index[nnz], value[nnz]),
 * infinite bounds are +-INFINITY. The arrays are copied. */
int smy_model_create(int m, int n, int maximize, double offset, const int64_t* start,
                     const int32_t* index, const double* value, const double* obj,
                     const double* col_lower, const double* col_upper, const double* row_lower,
                     const double* row_upper, smy_model** out);
int smy_model_generate(int kind, uint64_t seed, long long p0, long long p1, long long p2,
                       double p3, smy_model** out);
void smy_model_free(smy_model* model);
int smy_model_num_rows(const smy_model* model);
int smy_model_num_cols(const smy_model* model);
long long smy_model_num_nonzeros(const smy_model* model);
/* Copies the model out; any pointer may be NULL. */
int smy_model_export(const smy_model* model, int64_t* start, int32_t* index, double* value,
                     double* obj, double* col_lower, double* col_upper, double* row_lower,
                     double* row_upper, int* maximize, double* offset);

/* compute_scaling (src/linalg/scaling.cpp:130-200): power-of-two column (n) and row (m) factors
 * of the model's A. make_problem (src/lp/lp_solver.cpp:28-51): the computational form's cost,
 * lower and upper (n + m each) and the scaled values of A in the model's CSC order. */
int smy_compute_scaling(const smy_model* model, double* col_scale, double* row_scale);
int smy_make_problem(const smy_model* model, double* cost, double* lower, double* upper,
                     double* a_value);

/* solve_pdlp's preconditioner on the host (ruiz_equilibration, src/lp/pdlp.cpp:109-187, for
 * SMY_ALGORITHM_REFERENCE; for SMY_ALGORITHM_ENHANCED `passes` infinity-norm Ruiz rounds and,
 * with pock_chambolle, one Pock-Chambolle step with alpha = 1) of an LpProblem's A (m x n CSC)
 * and At: col_scale (n), row_scale (m), out = {the bound on ||[A_eq -I]|| of
 * estimate_operator_norm (:36-103), max column 2-norm of A, max row 2-norm of A (:377-394)}.
 * The device computes the same (smy_pdhg_equilibrate). */
int smy_pdlp_equilibrate(const smy_csc* A, const smy_csc* At, int algorithm, int passes,
                         int pock_chambolle, double* col_scale, double* row_scale, double out[3]);

/* PdlpOptions (src/lp/pdlp.hpp:11-16) plus this library's additions. */
typedef struct smy_pdlp_params {
  double tol;                /* 1e-8 (the solver's kPdlpTolerance, src/core/solver.cpp:31) */
  long long max_iterations;  /* -1: automatic */
  double time_limit;         /* seconds, INFINITY: none */
  int use_gpu;               /* must be 1: iterations run on the GPU only */
  int precision;             /* SMY_PRECISION_FP64 or SMY_PRECISION_MIXED */
  double fp32_switch_tol;    /* mixed: FP32 steps until this distance, then FP64 */
  int use_graph;             /* replay the 64 steps between evaluations from a HIP graph */
  int detect;                /* infeasibility / unboundedness detection */
  int log_level;             /* 0 silent ... 3 debug, to stdout */
  int verify;                /* model-level solve: run the verifier */
  int polish_rounds;         /* model-level solve: FP64 continuations with tol/10 if rejected */
  int algorithm;             /* SMY_ALGORITHM_REFERENCE (the reference's decisions) or _ENHANCED */
  int share_matrix;          /* batches of one model: run them as a shared-matrix batch (1) */
  int device_detect;         /* infeasibility / unboundedness tests on the device (1) or host */
  double start_omega;        /* primal weight to start with (<= 0: 1); with a warm start */
  int adaptive_steps;        /* PDLP's adaptive step sizes, decided on the device (0: constant
                                0.99 / L as the reference); batches then use the stacked layout */
  double fp32_stall_level;   /* mixed: a plateau of the FP32 steps (no 20% gain within
                                max(4096, iter / 4) iterations) below this distance also
                                switches to FP64 (1e-4; 0: never) */
  int device_restart;        /* restart decisions, distances and primal weights on the device (1)
                                or on the host from downloaded sums (0) */
  int fp64_accumulation;     /* mixed: FP32 steps sum their products in FP64 (1) or FP32 (0);
                                storage stays FP32; batches then use the stacked layout */
  int device_scaling;        /* one LP (either algorithm): the preconditioner, the scaled matrices,
                                cost and bounds and the operator-norm bound computed on the device
                                (smy_pdhg_equilibrate; 1) or on the host (0); the same factors */
  int column_order;          /* model-level solves: 1 the model's order; 2 by decreasing nonzero
                                count (the products' gathers of the hottest columns share cache
                                lines); 3 the hot columns (count >= 2x the mean) by count, the
                                rest by the first row holding them (a row's new columns side by
                                side); 0 auto: 3 for the enhanced algorithm with skewed column
                                counts (the longest >= 16x the mean), else 1. Answers come back in
                                the model's order */
  long long inject_fault_step; /* test hook: the device handle fails its steps after this many
                                  (smy_pdhg_inject_fault), so the solve takes the reference's
                                  GPU-failure path (src/lp/pdlp.cpp:231-242); -1: off */
  int halpern;               /* restarted Halpern PDHG with reflection (smy_pdhg_set_halpern):
                                constant steps, no averages, restarts on the fixed-point residual
                                to T(z) (0: the reference's restarted PDHG) */
  double reflection;         /* Halpern steps: the reflection in [0, 1] (1) */
  int pid_primal_weight;     /* the primal weight at a restart follows log(dy / dx) through a PI
                                controller (gains 0.99, 0.01) instead of the reference's
                                omega <- sqrt(omega dy / dx) (0); decided on the host */
  int farkas_projection;     /* when the reference's Farkas test fails, test the dual with its
                                entries of the wrong sign for a row's infinite bound set to 0,
                                for f_min and for f_max (smy_pdhg_detect_batch flags 2, 3):
                                1 on, 0 off (the reference's test only), -1 (default) on unless
                                algorithm is the reference's */
} smy_pdlp_params;

void smy_pdlp_params_default(smy_pdlp_params* params);

/* PdlpResult (src/lp/pdlp.hpp:19-30) scalars plus timings. */
typedef struct smy_pdlp_info {
  int status;
  long long iterations;
  long long fp32_iterations;
  double primal_residual;
  double dual_residual;
  double gap;
  double operator_norm;
  double omega;         /* final primal weight */
  double setup_seconds;
  double solve_seconds;
  double kernel_ms;
  int restarts;
  int evaluations;
  char message[256];
  long long rejected_steps;  /* adaptive steps: rejected attempts (included in iterations) */
} smy_pdlp_info;

/* solve_pdlp on an LpProblem in computational form (min cost'v, [A -I] v = 0,
 * lower <= v <= upper; A is m x n CSC; cost/lower/upper have n + m entries). Outputs (any may be
 * NULL): x (n + m), y (m), dual_ray (m), primal_ray (n) in that scaled space; the restart log
 * (iteration, to-average flag, KKT, omega) up to restart_cap entries; the evaluation log
 * (iteration, current rel primal/dual/gap, average rel primal/dual/gap) up to eval_cap. */
int smy_solve_pdlp(int m, int n, const int64_t* a_start, const int32_t* a_index,
                   const double* a_value, const double* cost, const double* lower,
                   const double* upper, const smy_pdlp_params* params, const double* start_x,
                   const double* start_y, smy_pdlp_info* info, double* x, double* y,
                   double* dual_ray, double* primal_ray, long long* restart_iter,
                   int* restart_to_average, double* restart_kkt, double* restart_omega,
                   int restart_cap, long long* eval_iter, double* eval_measures, int eval_cap);

/* The model-level answer (samaya::Result fields for an LP, include/samaya/solver.hpp:38-73). */
typedef struct smy_lp_result {
  int status;   /* SimplexStatus of the PDLP run */
  int verified;
  double objective;
  double max_primal_violation;
  double max_dual_violation;
  long long iterations; /* reported by samaya's CLI as barrier_iterations */
  long long fp32_iterations; /* of which FP32 steps (mixed precision) */
  int rounds;           /* PDLP solves run: 1 + polish rounds used */
  smy_pdlp_info pdlp;   /* the last solve */
  char message[256];
  double dual_bound;    /* optimal answers: a bound on the LP's value from the answer's row duals
                           by weak duality (smy_lp_dual_bound), valid whether or not the primal
                           verifies; lower for minimization, upper for maximization; NaN else */
} smy_lp_result;

/* A bound on the model's optimal value from row duals y (m, d = c - A'y convention), by weak
 * duality (no reference counterpart: the reference's branch and bound takes the dual simplex's
 * objective as a node's bound, src/mip/branch_and_bound.hpp:133-145): y made sign-feasible, then
 * sum_j min over the column's box of d_j x_j plus sum_i min over the row's range of y_i r_i, in
 * double-double less a margin that covers its rounding. Lower bound for minimization, upper for
 * maximization (offset included); -inf / +inf when an unbounded variable has a nonzero reduced
 * cost. */
int smy_lp_dual_bound(const smy_model* model, const double* row_dual, double* bound);

/* LP by PDLP on the GPU (solve_pdlp_lp): power-of-two scaling, computational form, PDLP, mapping
 * back and the independent verifier. Vectors (any may be NULL): col_value (n), row_dual (m,
 * d = c - A'y convention), col_dual (n), row_activity (m), certificate (m if infeasible, n if
 * unbounded). */
int smy_solve_lp(const smy_model* model, const smy_pdlp_params* params, smy_lp_result* result,
                 double* col_value, double* row_dual, double* col_dual, double* row_activity,
                 double* certificate);

/* smy_solve_lp of `count` models in one batch (no reference counterpart): the PDLP iterations of
 * all models share the kernel launches on one device handle, every other decision is each
 * model's own, so results[i] equals smy_solve_lp of models[i] alone. col_values / row_duals:
 * NULL, or `count` pointers (each NULL or of the model's size). times[3] (may be NULL): host
 * setup seconds, solve seconds (setup included) and device milliseconds of the step runs, summed
 * over the rounds. */
int smy_solve_lp_batch(const smy_model* const* models, int count, const smy_pdlp_params* params,
                       smy_lp_result* results, double* const* col_values,
                       double* const* row_duals, double* times);

/* Node relaxations of one model (no reference counterpart): smy_solve_lp of `model` with column
 * bounds col_lower[i] / col_upper[i] (n each) for i < count, in one batch on one shared matrix
 * (the shared-matrix batch unless params->share_matrix is 0). start_col_value (n) and
 * start_row_dual (m, d = c - A'y convention) may be NULL; when given they warm-start every
 * relaxation (typically the root LP's answer). Outputs as smy_solve_lp_batch. */
int smy_solve_lp_bounds_batch(const smy_model* model, int count, const double* const* col_lower,
                              const double* const* col_upper, const smy_pdlp_params* params,
                              const double* start_col_value, const double* start_row_dual,
                              smy_lp_result* results, double* const* col_values,
                              double* const* row_duals, double* times);

/* A branch-and-bound bound vector (deterministic, seeded): a fraction U[frac_min, frac_max] of
 * the n columns, distinct and uniform, fixed at their lower or upper bound (even odds); the rest
 * keep lower / upper. */
int smy_generate_fixings(uint64_t seed, int n, double frac_min, double frac_max,
                         const double* lower, const double* upper, double* out_lower,
                         double* out_upper);

/* samaya_solve (include/samaya_c.h:41) for this library: LP by PDLP on the GPU with default
 * parameters. Returns samaya's Status code (1 optimal, 2 infeasible, 3 unbounded, 5 time limit,
 * 6 iteration limit, 8 numerical error) and the objective (NaN without an answer). */
int smy_solve(const smy_model* model, double* objective);

/* Independent verifier (double-double) on model-space x (n) and y (m) / a certificate. out3 =
 * {max primal violation, max dual violation, objective}. Returns 1 when the check passes. */
int smy_verify_lp(const smy_model* model, const double* x, const double* y, double primal_tol,
                  double dual_tol, double* out3, char* message, size_t message_len);
int smy_verify_infeasibility(const smy_model* model, const double* y, char* message,
                             size_t message_len);
int smy_verify_unbounded_ray(const smy_model* model, const double* ray, char* message,
                             size_t message_len);

/* GpuInfo (include/samaya/solver.hpp:93-97): built_with_gpu_kernels is 1 in this library. */
int smy_gpu_info(int* built_with_gpu_kernels, char* device, size_t device_len);

#ifdef __cplusplus
}
#endif

#endif /* SAMAYA_PDLP_H */
