/* C ABI of the branch and bound whose node relaxations are bounded in waves by batched PDLP on the
 * GPU (samaya_amd/csrc/host/mip.cpp). No entry point of the reference corresponds one to one:
 * samaya's MIP search (src/mip/branch_and_bound.hpp:133-145, reached through Solver::solve with
 * integer columns, include/samaya/solver.hpp:75-86) bounds each node by the dual simplex,
 * warm-started from its parent's basis, one node at a time. Here the open nodes with the best
 * bounds are solved together as one shared-matrix PDLP batch, pruned by safe dual bounds (weak
 * duality over each node's box) and Farkas certificates, and the answers rounded for incumbents.
 * No cuts, propagation, strong branching or conflict analysis. */
#ifndef SAMAYA_MIP_H
#define SAMAYA_MIP_H

#include "samaya_pdlp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct smy_mip_params {
  smy_pdlp_params lp;      /* the node relaxations' solves */
  double gap_tol;          /* stop when incumbent - bound <= gap_tol (1 + |incumbent|) (1e-4) */
  int wave;                /* open nodes bounded per batch (256) */
  long long max_nodes;     /* relaxations solved at most (100000) */
  double time_limit;       /* seconds (infinite) */
  double integrality_tol;  /* 1e-6 */
  double feasibility_tol;  /* a candidate's rows and bounds, relative to 1 + |bound| (1e-6) */
  int log_level;           /* 1: a line per wave and per new incumbent (stdout) */
} smy_mip_params;

typedef struct smy_mip_result {
  int status;               /* SMY_STATUS_OPTIMAL: the gap closed; _INFEASIBLE: no integer point;
                               _ITERATION_LIMIT (nodes) / _TIME_LIMIT; _NUMERICAL_ERROR: relaxations
                               left unresolved hold the bound below the gap */
  double objective;         /* the incumbent's, in the model's sense (NaN without one) */
  double bound;             /* the best bound in the model's sense (maximization: upper) */
  double gap;               /* relative, minimization sense (infinite without an incumbent) */
  long long nodes;          /* relaxations solved (the root included) */
  long long waves;          /* batches */
  long long pruned_infeasible;
  long long pruned_bound;
  long long integral;       /* relaxations with an integral answer */
  long long unresolved;     /* relaxations left without a usable point (their bounds kept) */
  long long inexact;        /* relaxations stopped short of optimality, branched on their point
                               with the parent's bound */
  long long fixed_solves;   /* batches re-solving integral nodes' continuous parts (integers
                               fixed, FP64, tolerance <= 1e-9) whose rounding missed a row */
  long long lp_iterations;  /* PDLP iterations over all relaxations */
  double seconds;
  char message[256];        /* why a solve failed (SMY_ERROR) */
} smy_mip_result;

void smy_mip_params_default(smy_mip_params* params);

/* integer: n flags (nonzero: integer column), or null for the model's own integrality (an MPS
 * file's markers; a model without any is refused). col_value: n doubles for the incumbent, or
 * null. Returns SMY_OK with the result filled, or an error code. */
int smy_mip_solve(const smy_model* model, const unsigned char* integer, const smy_mip_params* params,
                  smy_mip_result* result, double* col_value);

/* The bounds the search starts from: column bounds tightened to those the rows imply
 * (activity-based propagation, up to 10 passes, relaxed by margins that never cut off a point that
 * satisfies the rows; integer columns rounded inward) and each one-sided row's missing bound set to
 * its activity range. Outputs n and m doubles (any may be null); *crossed = 1 when some column's
 * bounds cross (no point satisfies the rows; the outputs are then partial), else 0. integer: n
 * flags or null (the model's own integrality). No reference counterpart (samaya's search
 * propagates at every node, src/mip/branch_and_bound.hpp). */
int smy_mip_propagate(const smy_model* model, const unsigned char* integer, double* col_lower,
                      double* col_upper, double* row_lower, double* row_upper, int* crossed);

#ifdef __cplusplus
}
#endif

#endif /* SAMAYA_MIP_H */
