// The PDLP host solver: every algorithmic decision (equilibration, step sizes, termination,
// restarts, primal weight, infeasibility and unboundedness detection) on the host, every
// iteration on the gfx950 device through the C ABI of include/samaya_pdhg.h.
//
// Mirrors samaya's solve_pdlp (src/lp/pdlp.hpp:11-35, src/lp/pdlp.cpp:228-1193): PdlpOptions and
// PdlpResult keep the reference's fields and meaning, and with the default options ("parity"
// algorithm, FP64) the solver takes the reference's decisions with the reference's constants.
// Additions, each off by default or selected explicitly:
//   * precision kMixed: FP32 values and iterates with FP32 accumulation in the steps, FP64
//     evaluation sums, restarts and termination, then FP64 steps (the polish) to the tolerance;
//   * HIP-graph replay of the 64 steps between evaluations;
//   * polish_to_verify: after an optimal answer that the verifier rejects, continue the FP64
//     iteration with a tighter tolerance (the reference hands that case to crossover);
//   * warm start from a given primal-dual point.
#pragma once

#include <cstdint>
#include <limits>
#include <string>
#include <vector>

#include "lp_data.hpp"

namespace smy::host {

// SimplexStatus (src/lp/simplex.hpp:28-35), same order and values.
enum class Status : int {
  kOptimal = 0,
  kInfeasible,
  kUnbounded,
  kIterationLimit,
  kTimeLimit,
  kNumericalError,
};
const char* to_string(Status s);

enum class Precision : int { kFp64 = 0, kMixed = 1 };

// kReference: the reference's decisions exactly (2-norm equilibration variant, pdlp.cpp:109-187).
// kEnhanced: infinity-norm Ruiz (10 passes) followed by Pock-Chambolle (alpha = 1) diagonal
// preconditioning, the usual PDLP preconditioner; every other decision unchanged.
enum class Algorithm : int { kReference = 0, kEnhanced = 1 };
// Column order of a model-level solve: kAuto puts the columns of an enhanced-algorithm solve in
// decreasing order of their nonzero counts when those are skewed (the longest column at least 16
// times the mean); kDegree always, kNatural never. Answers come back in the model's order.
enum class ColumnOrder : int { kAuto = 0, kNatural = 1, kDegree = 2, kTouch = 3 };

struct PdlpOptions {
  double tol = 1e-8;              // max(rel primal, rel dual, |rel gap|) < tol
  long long max_iterations = -1;  // -1: max(1e4, 1000 (m + n)), capped at 1e6 without time limit
  double time_limit = kInf;       // seconds, setup included
  bool use_gpu = true;            // iterations run on the GPU only: false is an error here
  Precision precision = Precision::kFp64;
  Algorithm algorithm = Algorithm::kReference;
  // Mixed precision: FP32 steps until the better of current / average reaches this (FP64)
  // termination distance, or FP32 stops improving below fp32_stall_level (no 20% gain within
  // max(4096, iter / 4) iterations; 0: never); then FP64 steps to `tol`.
  double fp32_switch_tol = 1e-5;
  double fp32_stall_level = 1e-4;
  // PDLP's adaptive step sizes (decided on the device) instead of the constant 0.99 / L; the
  // averages are then weighted by step size. Uses the stacked layout for every batch.
  bool adaptive_steps = false;
  // Mixed precision: FP32 steps sum their products (and form the update that uses them) in FP64;
  // storage stays FP32. Uses the stacked layout for every batch.
  bool fp64_accumulation = false;
  // Test hook: the device handle fails its steps after this many (-1: off), so the solve takes
  // the GPU-failure path.
  long long inject_fault_step = -1;
  // One LP: the preconditioner (either algorithm's) and the operator-norm bound computed on the
  // device from the unscaled LP (the same factors as the host's).
  bool device_scaling = true;
  ColumnOrder column_order = ColumnOrder::kAuto;  // model-level solves (solve_lp*)
  bool use_graph = true;          // replay each 64-step run from a HIP graph
  bool detect = true;             // infeasibility / unboundedness checks (pdlp.cpp:715-1043)
  bool device_detect = true;      // ... computed on the device (false: host, from downloads)
  bool share_matrix = true;       // batches of one problem: the shared-matrix (SpMM) batch
  bool device_restart = true;     // restart decisions, distances and primal weights on the device
  // Restarted Halpern PDHG with reflection (Lu and Yang 2024, as in cuPDLPx) instead of restarted
  // PDHG with averages: constant steps 0.99 / L, the anchored step
  // z' = lambda ((1 + reflection) T(z) - reflection z) + (1 - lambda) z0, restarts decided on the
  // fixed-point residual and made to T(z). Overrides adaptive_steps; both batch layouts take it.
  bool halpern = false;
  double reflection = 1.0;
  // The primal weight at a restart follows log(dy / dx) (the distances moved since the last
  // restart) through a PI controller with gains 0.99 and 0.01 (as in cuPDLPx) instead of the
  // reference's halfway step omega <- sqrt(omega dy / dx). Decided on the host.
  bool pid_primal_weight = false;
  // When the reference's Farkas test fails on a dual y, test its sign projection: the entries
  // whose sign makes a row's term of f_min (then of f_max) infinite set to 0. A certificate it
  // accepts is a Farkas certificate as the reference's test defines one; the reference tests y
  // only.
  bool farkas_projection = false;
  int log_level = 0;
  // Warm start (scaled computational-form space, before equilibration): v (n + m) and y (m),
  // and the primal weight to start with (<= 0: 1, the reference's start).
  std::vector<double> start_x, start_y;
  double start_omega = 0.0;
};

struct RestartEvent {
  long long iter;
  int to_average;
  double kkt;
  double omega;
};

struct EvalEvent {
  long long iter;
  double current[3];  // rel primal, rel dual, rel gap
  double average[3];
};

struct PdlpResult {
  Status status = Status::kNumericalError;
  std::vector<double> x;           // n + m, scaled computational-form space
  std::vector<double> y;           // m
  std::vector<double> dual_ray;    // m, when infeasible
  std::vector<double> primal_ray;  // n, when unbounded
  long long iterations = 0;
  long long fp32_iterations = 0;   // of which in FP32 (mixed precision)
  long long rejected_steps = 0;    // adaptive steps: attempts rejected (counted in iterations)
  double primal_residual = 0.0;
  double dual_residual = 0.0;
  double gap = 0.0;
  double operator_norm = 0.0;      // L, the bound on ||[A_eq -I]||
  double omega = 1.0;              // the primal weight at the end (a warm start's start_omega)
  double setup_seconds = 0.0;      // equilibration, norm estimate, upload
  double solve_seconds = 0.0;      // everything
  double kernel_ms = 0.0;          // device time of the step runs (HIP events)
  std::string message;
  std::vector<RestartEvent> restarts;
  std::vector<EvalEvent> evals;
};

// The preconditioner of solve_pdlp on the host (the reference's 2-norm variant, or for the
// enhanced algorithm `passes` infinity-norm Ruiz rounds and, with pock_chambolle, one
// Pock-Chambolle step), then out = {bound on ||[A_eq -I]||, a_colmax, at_rowmax}.
void equilibrate_host(const Csc& A, const Csc& At, Algorithm algorithm, int passes,
                      bool pock_chambolle, std::vector<double>& rs, std::vector<double>& cs,
                      double out[3]);

PdlpResult solve_pdlp(const LpProblem& lp, const PdlpOptions& options);

// ---- Batches (no reference counterpart) ----
// Many LPs solved together on one device handle: their steps share the kernel launches, every
// other decision is each LP's own, so every result equals solve_pdlp of that LP alone (same
// iterates, iterations, restarts and status).
struct PdlpTask {
  const LpProblem* lp = nullptr;
  const std::vector<double>* start_x = nullptr;  // warm start (null: options.start_x)
  const std::vector<double>* start_y = nullptr;
  // Bounds of this task (n + m, the problem's space; null: the problem's own). Tasks of one
  // problem with equal row bounds and their own column bounds (node relaxations) run as a
  // shared-matrix batch.
  const std::vector<double>* lower = nullptr;
  const std::vector<double>* upper = nullptr;
};

struct PdlpBatchInfo {
  double setup_seconds = 0.0;  // host setup of every LP and the upload
  double solve_seconds = 0.0;  // until the last LP stopped
  double kernel_ms = 0.0;      // device time of the step runs
};

std::vector<PdlpResult> solve_pdlp_batch(const std::vector<PdlpTask>& tasks,
                                         const PdlpOptions& options, PdlpBatchInfo* info = nullptr);

// ---- Model-level path (solve_pdlp_lp, src/core/solver.cpp:392-477) ----

struct VerifyReport {
  bool ok = true;
  double max_bound_violation = 0.0;
  double max_row_violation = 0.0;
  double max_dual_violation = 0.0;
  double objective = 0.0;
  std::string message;
};

// Independent checks in double-double arithmetic (src/verify/verify.cpp:132-401), tolerances
// 1e-6 primal / dual relative (include/samaya/verify.hpp:17-21).
VerifyReport verify_lp_optimality(const Model& model, const double* x, const double* y,
                                  double primal_tol = 1e-6, double dual_tol = 1e-6);
VerifyReport verify_infeasibility(const Model& model, const double* y, double primal_tol = 1e-6);
// A bound on the LP's optimal value from any row duals y (d = c - A'y convention), by weak
// duality: y is first made sign-feasible (the dual of a row with only an upper bound <= 0 in the
// minimization sense, with only a lower bound >= 0), then
//   sum_j min over [l_j, u_j] of d_j x_j + sum_i min over [rl_i, ru_i] of y_i r_i
// in double-double, less a margin for its rounding. A lower bound for minimization, an upper bound
// for maximization (offset included); -inf (+inf) where a needed bound is infinite. A node
// relaxation's bound for branch and bound, valid whether or not the primal point verifies.
double safe_dual_bound(const Model& model, const double* y);
VerifyReport verify_unbounded_ray(const Model& model, const double* ray, double primal_tol = 1e-6,
                                  double dual_tol = 1e-6);

struct LpSolveOptions {
  PdlpOptions pdlp;
  bool verify = true;
  // Optimal but rejected by the verifier: continue FP64 iterations with tol / 10, up to this many
  // times, before reporting the answer as unverified.
  int polish_rounds = 0;
};

struct LpSolveResult {
  Status status = Status::kNumericalError;
  bool verified = false;
  double objective = 0.0;
  std::vector<double> col_value;  // n
  std::vector<double> row_activity;  // m
  std::vector<double> row_dual;   // m, d = c - A'y convention
  std::vector<double> col_dual;   // n
  std::vector<double> infeasibility_certificate;  // m
  std::vector<double> unbounded_ray;              // n
  double max_primal_violation = 0.0;
  double max_dual_violation = 0.0;
  // safe_dual_bound of the answer's row duals (optimal answers; NaN otherwise).
  double dual_bound = std::numeric_limits<double>::quiet_NaN();
  long long iterations = 0;       // reported by the reference as barrier_iterations
  long long fp32_iterations = 0;  // of which FP32 steps (mixed precision)
  int rounds = 0;                 // solves run (1 + polish rounds used)
  std::string message;
  PdlpResult pdlp;
};

LpSolveResult solve_lp(const Model& model, const LpSolveOptions& options);

// solve_lp of every model, the PDLP solves of each round batched (polish rounds batch the models
// that need them).
std::vector<LpSolveResult> solve_lp_batch(const std::vector<const Model*>& models,
                                          const LpSolveOptions& options,
                                          PdlpBatchInfo* info = nullptr);

// Node relaxations of one model: solve_lp of the model with column bounds col_lower[k] /
// col_upper[k] (n each) for every k, batched on one shared matrix (the shared-matrix batch when
// the options allow it). An optional model-space start (column values, and row duals in the
// d = c - A'y convention) warm-starts every relaxation.
std::vector<LpSolveResult> solve_lp_bounds_batch(const Model& model,
                                                 const std::vector<const double*>& col_lower,
                                                 const std::vector<const double*>& col_upper,
                                                 const LpSolveOptions& options,
                                                 const double* start_col_value = nullptr,
                                                 const double* start_row_dual = nullptr,
                                                 PdlpBatchInfo* info = nullptr);

}  // namespace smy::host
