// Branch and bound for mixed-integer LPs whose node relaxations are bounded in waves by batched
// PDLP on the GPU. samaya's branch and bound solves each node's relaxation by the dual simplex,
// warm-started from its parent's basis, one node at a time (src/mip/branch_and_bound.hpp:133-145);
// here the open nodes with the best bounds are solved together as one shared-matrix batch (one
// matrix, per-node column bounds) and pruned by safe dual bounds and Farkas certificates.
#pragma once

#include <limits>
#include <vector>

#include "pdlp.hpp"

namespace smy::host {

struct MipOptions {
  LpSolveOptions lp;             // the node relaxations' solves
  double gap_tol = 1e-4;         // stop when (incumbent - bound) <= gap_tol (1 + |incumbent|)
  int wave = 256;                // open nodes bounded per batch
  long long max_nodes = 100000;  // relaxations solved at most (then kIterationLimit)
  double time_limit = kInf;      // seconds (then kTimeLimit)
  double integrality_tol = 1e-6;
  double feasibility_tol = 1e-6;  // rows and bounds of a candidate, relative to 1 + |bound|
  int log_level = 0;
};

struct MipResult {
  // kOptimal: the gap closed; kInfeasible: every node pruned without an incumbent; a limit
  // otherwise (with the incumbent and bound reached).
  Status status = Status::kNumericalError;
  // The incumbent's objective, in the model's sense with its offset (NaN without one).
  double objective = std::numeric_limits<double>::quiet_NaN();
  // The best bound in the model's sense (a maximization: an upper bound).
  double bound = std::numeric_limits<double>::quiet_NaN();
  // (incumbent - bound) in the minimization sense, over 1 + |incumbent| (infinite without one).
  double gap = std::numeric_limits<double>::infinity();
  std::vector<double> col_value;  // the incumbent (n), empty without one
  long long nodes = 0;            // relaxations solved
  long long waves = 0;            // batches
  long long pruned_infeasible = 0;
  long long pruned_bound = 0;     // at their solve, or before it (bound no better than the incumbent)
  long long integral = 0;         // relaxations whose answer was integral
  long long unresolved = 0;       // relaxations left without a point, or with an integral one that
                                  // missed a row (their bounds stay in the search's bound)
  long long inexact = 0;          // relaxations stopped short of optimality, branched on their point
  long long fixed_solves = 0;     // batches of continuous parts with the integer columns fixed
  long long lp_iterations = 0;
  double seconds = 0.0;
};

// `integer`: n flags (nonzero: the column must take an integer value).
MipResult solve_mip(const Model& model, const std::vector<char>& integer, const MipOptions& options);

// The column bounds the rows imply (and the activity range as a one-sided row's missing bound),
// written into `model`; false when some column's bounds cross (no point satisfies the rows).
bool propagate_bounds(Model& model, const std::vector<char>& integer);

}  // namespace smy::host
