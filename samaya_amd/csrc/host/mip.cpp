// Best-bound branch and bound over waves of batched PDLP node relaxations (mip.hpp).
//
// Each wave takes up to `wave` open nodes with the smallest bounds (minimization sense) that can
// still beat the incumbent, solves their relaxations as one batch (solve_lp_bounds_batch: the
// model's matrix once, one column-bound vector per node), warm-started from the root relaxation's
// answer and primal weight, and for each node:
//  - prunes it when the relaxation is infeasible (a Farkas certificate);
//  - takes its bound from the answer's safe dual bound (weak duality over the node's box with
//    rounding margins: a valid bound whatever the duals' accuracy), never below its parent's, and
//    prunes it when that cannot beat the incumbent by more than the gap tolerance;
//  - rounds the answer's integer columns to the nearest integers within the node's bounds and
//    accepts the point as an incumbent when it satisfies every row and bound;
//  - branches on the integer column of largest fractionality: x_j <= floor(v), x_j >= ceil(v).
// The reference's search (src/mip/branch_and_bound.cpp) adds propagation, cuts, strong branching,
// pseudocosts, diving heuristics and conflict analysis, none of which is here: this is the
// bounding of nodes through the batched GPU path and the loop that uses it.
#include "mip.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>

namespace smy::host {

namespace {

struct Node {
  double bound;  // a lower bound of the node's optimum, minimization sense
  int depth;
  std::vector<double> lower, upper;  // the node's column bounds (n)
};

using NodePtr = std::unique_ptr<Node>;

// Min-heap by bound (ties: the deeper node first, which finds incumbents sooner).
bool heap_less(const NodePtr& a, const NodePtr& b) {
  if (a->bound != b->bound) return a->bound > b->bound;
  return a->depth < b->depth;
}

double objective_of(const Model& md, const std::vector<double>& x) {
  double f = md.offset;
  for (int j = 0; j < md.n; ++j) f += md.obj[j] * x[j];
  return f;
}

bool within(double v, double lo, double up, double tol) {
  return v >= lo - tol * (1.0 + std::fabs(lo)) && v <= up + tol * (1.0 + std::fabs(up));
}

// Whether x satisfies the node's column bounds and the model's rows within tol.
bool feasible(const Model& md, const std::vector<double>& x, const Node& node, double tol) {
  for (int j = 0; j < md.n; ++j)
    if (!within(x[j], node.lower[j], node.upper[j], tol)) return false;
  std::vector<double> act(static_cast<size_t>(md.m), 0.0);
  for (int j = 0; j < md.n; ++j)
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) act[md.A.index[p]] += md.A.value[p] * x[j];
  for (int i = 0; i < md.m; ++i)
    if (!within(act[i], md.row_lower[i], md.row_upper[i], tol)) return false;
  return true;
}

// Column bounds implied by the rows (activity-based propagation, up to 10 passes): for row i with
// bounds [L, U] and entry a of column j, the other entries' least and greatest activities bound
// a x_j from both sides. A new bound is relaxed by 1e-9 (1 + |bound|) plus 1e-11 of the row's
// magnitude over |a| (far above the rounding of its sums) so that it never cuts off a feasible
// point, and an integer column's is rounded inward (1e-6 slack). Every point that
// satisfies the rows and the old bounds satisfies the new ones, so the MIP keeps its optimum,
// while its relaxations get boxed columns (and two-sided rows, below), whose safe dual bounds are
// finite. Returns false when some column's bounds cross (the MIP is infeasible).
}  // namespace

namespace {

// The rows of A (CSR), for propagation.
struct Rows {
  std::vector<int64_t> start;
  std::vector<int> col;
  std::vector<double> val;
};

Rows rows_of(const Model& md) {
  const int m = md.m, n = md.n;
  Rows R;
  R.start.assign(static_cast<size_t>(m) + 1, 0);
  for (int64_t p = 0; p < md.A.start[n]; ++p) ++R.start[md.A.index[p] + 1];
  for (int i = 0; i < m; ++i) R.start[i + 1] += R.start[i];
  R.col.resize(static_cast<size_t>(md.A.start[n]));
  R.val.resize(R.col.size());
  std::vector<int64_t> next(R.start.begin(), R.start.end() - 1);
  for (int j = 0; j < n; ++j)
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const int64_t q = next[md.A.index[p]]++;
      R.col[q] = j;
      R.val[q] = md.A.value[p];
    }
  return R;
}

// The passes of propagate_bounds over column bounds lo / up (the model's rows and row bounds);
// false when some column's bounds cross.
bool tighten_columns(const Model& md, const Rows& R, const std::vector<char>& integer,
                     std::vector<double>& lo, std::vector<double>& up) {
  const int m = md.m, n = md.n;
  const std::vector<int64_t>& start = R.start;
  const std::vector<int>& col = R.col;
  const std::vector<double>& val = R.val;
  const auto tighten = [&](int j, double b, double row_slack, bool upper) -> bool {
    if (!std::isfinite(b)) return false;
    const double slack = 1e-9 * (1.0 + std::fabs(b)) + row_slack;
    if (upper) {
      double v = b + slack;
      if (integer[j]) v = std::floor(b + slack + 1e-6);
      if (!(up[j] < kInf) || v < up[j] - 1e-9 * (1.0 + std::fabs(up[j]))) {
        up[j] = v;
        return true;
      }
    } else {
      double v = b - slack;
      if (integer[j]) v = std::ceil(b - slack - 1e-6);
      if (!(lo[j] > -kInf) || v > lo[j] + 1e-9 * (1.0 + std::fabs(lo[j]))) {
        lo[j] = v;
        return true;
      }
    }
    return false;
  };
  for (int pass = 0; pass < 10; ++pass) {
    bool changed = false;
    for (int i = 0; i < m; ++i) {
      // Least and greatest activities: finite parts and counts of infinite terms.
      double min_f = 0.0, max_f = 0.0, mag = 0.0;
      int min_inf = 0, max_inf = 0;
      for (int64_t q = start[i]; q < start[i + 1]; ++q) {
        const double a = val[q], l = lo[col[q]], u = up[col[q]];
        const double lmin = a > 0 ? l : u, lmax = a > 0 ? u : l;
        if (std::isfinite(lmin)) {
          min_f += a * lmin;
          mag += std::fabs(a * lmin);
        } else {
          ++min_inf;
        }
        if (std::isfinite(lmax)) {
          max_f += a * lmax;
          mag += std::fabs(a * lmax);
        } else {
          ++max_inf;
        }
      }
      if (std::isfinite(md.row_lower[i])) mag += std::fabs(md.row_lower[i]);
      if (std::isfinite(md.row_upper[i])) mag += std::fabs(md.row_upper[i]);
      for (int64_t q = start[i]; q < start[i + 1]; ++q) {
        const int j = col[q];
        const double a = val[q];
        if (a == 0.0) continue;
        const double lmin = a > 0 ? lo[j] : up[j], lmax = a > 0 ? up[j] : lo[j];
        // The other entries' least / greatest activity (infinite unless every other term is finite).
        const double rest_min = min_inf - (std::isfinite(lmin) ? 0 : 1) > 0 ? -kInf
                                : min_f - (std::isfinite(lmin) ? a * lmin : 0.0);
        const double rest_max = max_inf - (std::isfinite(lmax) ? 0 : 1) > 0 ? kInf
                                : max_f - (std::isfinite(lmax) ? a * lmax : 0.0);
        const double row_slack = 1e-11 * mag / std::fabs(a);
        if (std::isfinite(md.row_upper[i]) && std::isfinite(rest_min))
          changed |= tighten(j, (md.row_upper[i] - rest_min) / a, row_slack, a > 0);
        if (std::isfinite(md.row_lower[i]) && std::isfinite(rest_max))
          changed |= tighten(j, (md.row_lower[i] - rest_max) / a, row_slack, a < 0);
      }
    }
    for (int j = 0; j < n; ++j)
      if (lo[j] > up[j] + 1e-9 * (1.0 + std::fabs(up[j]))) return false;
    if (!changed) break;
  }
  return true;
}

}  // namespace

bool propagate_bounds(Model& md, const std::vector<char>& integer) {
  const Rows R = rows_of(md);
  if (!tighten_columns(md, R, integer, md.col_lower, md.col_upper)) return false;
  const std::vector<int64_t>& start = R.start;
  const std::vector<int>& col = R.col;
  const std::vector<double>& val = R.val;
  const std::vector<double>& lo = md.col_lower;
  const std::vector<double>& up = md.col_upper;
  // A row's missing bound becomes its least / greatest activity over the column bounds (relaxed
  // as above): every point within the column bounds satisfies it, and the rows' terms of the
  // safe dual bound are then finite whatever the sign of their duals.
  for (int i = 0; i < md.m; ++i) {
    double min_a = 0.0, max_a = 0.0, mag = 0.0;
    for (int64_t q = start[i]; q < start[i + 1]; ++q) {
      const double a = val[q], l = lo[col[q]], u = up[col[q]];
      min_a += a > 0 ? a * l : a * u;
      max_a += a > 0 ? a * u : a * l;
      mag += std::fabs(a) * std::max(std::fabs(l), std::fabs(u));
    }
    const double slack = 1e-9 * (1.0 + mag);
    if (!(md.row_lower[i] > -kInf) && std::isfinite(min_a)) md.row_lower[i] = min_a - slack;
    if (!(md.row_upper[i] < kInf) && std::isfinite(max_a)) md.row_upper[i] = max_a + slack;
  }
  return true;
}

MipResult solve_mip(const Model& original, const std::vector<char>& integer, const MipOptions& opt) {
  const auto t0 = std::chrono::steady_clock::now();
  const auto seconds = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  MipResult out;
  // The search runs on the model with the column bounds its rows imply.
  Model md = original;
  if (!propagate_bounds(md, integer)) {
    out.status = Status::kInfeasible;
    return out;
  }
  const int n = md.n;
  const double sense = md.maximize ? -1.0 : 1.0;  // minimization sense: sense * objective
  // Each node's bounds are propagated again before its relaxation (a branching often implies
  // more), on models small enough that the host passes cost little beside the batch.
  const bool node_propagation = md.A.nnz() <= 200000;
  const Rows rows = node_propagation ? rows_of(md) : Rows{};
  double incumbent = kInf;  // minimization sense
  // The smallest bound of the nodes closed without being beaten by the incumbent: pruned within
  // the gap tolerance, or left unresolved (their parent's bound).
  double closed_min = kInf;
  const auto gap_abs = [&](double f) { return std::isfinite(f) ? opt.gap_tol * (1.0 + std::fabs(f)) : 0.0; };
  const auto log = [&](const char* what) {
    if (opt.log_level >= 1) {
      std::printf("mip: %s: nodes %lld, waves %lld, incumbent %.9g, %.1f s\n", what, out.nodes, out.waves,
                  std::isfinite(incumbent) ? sense * incumbent : NAN, seconds());
      std::fflush(stdout);
    }
  };

  // The root relaxation (its answer and primal weight warm-start every node).
  LpSolveOptions lpo = opt.lp;
  const LpSolveResult root = solve_lp(md, lpo);
  out.nodes = 1;
  out.lp_iterations = root.iterations;
  if (root.status != Status::kOptimal) {
    // An infeasible relaxation makes the MIP infeasible; an unbounded one leaves it unbounded or
    // infeasible (reported as the relaxation's status); any other stop is the relaxation's.
    out.status = root.status;
    if (root.status == Status::kInfeasible) ++out.pruned_infeasible;
    out.seconds = seconds();
    return out;
  }
  lpo.pdlp.start_omega = root.pdlp.omega;
  const std::vector<double> start_x = root.col_value, start_y = root.row_dual;

  std::vector<NodePtr> heap;
  std::vector<NodePtr> to_fix;  // integral relaxations whose rounding missed a row
  // The continuous part of the nodes in to_fix, integer columns fixed, solved in one batch to a
  // tighter tolerance in FP64: a row-feasible answer is the node's optimum (an incumbent
  // candidate); otherwise the node stays unresolved with its bound.
  LpSolveOptions fix_opts = opt.lp;
  fix_opts.pdlp.tol = std::min(opt.lp.pdlp.tol, 1e-9);
  fix_opts.pdlp.precision = Precision::kFp64;
  fix_opts.pdlp.halpern = false;
  fix_opts.pdlp.adaptive_steps = false;
  fix_opts.verify = true;  // polish rounds follow a rejected answer
  fix_opts.polish_rounds = std::max(opt.lp.polish_rounds, 2);
  const auto fix_and_solve = [&]() {
    if (to_fix.empty()) return;
    std::vector<const double*> lows, ups;
    for (const NodePtr& node : to_fix) {
      lows.push_back(node->lower.data());
      ups.push_back(node->upper.data());
    }
    const std::vector<LpSolveResult> rs =
        solve_lp_bounds_batch(md, lows, ups, fix_opts, start_x.data(), start_y.data());
    ++out.fixed_solves;
    for (size_t k = 0; k < to_fix.size(); ++k) {
      const Node& node = *to_fix[k];
      const LpSolveResult& r = rs[k];
      out.lp_iterations += r.iterations;
      const bool ok = r.status == Status::kOptimal && r.col_value.size() == static_cast<size_t>(n) &&
                      feasible(md, r.col_value, node, opt.feasibility_tol);
      if (ok) {
        const double f = sense * objective_of(md, r.col_value);
        if (f < incumbent) {
          incumbent = f;
          out.col_value = r.col_value;
          log("new incumbent (integer columns fixed)");
        }
      } else if (r.status != Status::kInfeasible) {
        ++out.unresolved;
        closed_min = std::min(closed_min, node.bound);
      }
    }
    to_fix.clear();
  };
  // A solved relaxation of `node`: prune, update the incumbent, or branch.
  const auto handle = [&](const Node& node, const LpSolveResult& r) {
    if (&r != &root) out.lp_iterations += r.iterations;
    if (r.status == Status::kInfeasible) {
      ++out.pruned_infeasible;
      return;
    }
    if (r.col_value.size() != static_cast<size_t>(n)) {
      ++out.unresolved;
      closed_min = std::min(closed_min, node.bound);
      return;
    }
    // A relaxation stopped short of optimality (a limit, or a ray test misled on a node that a
    // bounded root makes bounded) still has a point: the node branches on it with its parent's
    // bound, so no part of the search is dropped.
    const bool exact = r.status == Status::kOptimal;
    if (!exact) ++out.inexact;
    const double lb =
        std::max(node.bound, exact && std::isfinite(r.dual_bound) ? sense * r.dual_bound : -kInf);
    if (lb >= incumbent - gap_abs(incumbent)) {
      ++out.pruned_bound;
      if (lb < incumbent) closed_min = std::min(closed_min, lb);
      return;
    }
    std::vector<double> x = r.col_value;
    int branch = -1;
    double widest = opt.integrality_tol;
    for (int j = 0; j < n; ++j) {
      if (!integer[j]) continue;
      const double f = x[j] - std::floor(x[j]);
      const double frac = std::min(f, 1.0 - f);
      if (frac > widest) {
        widest = frac;
        branch = j;
      }
      x[j] = std::min(std::max(std::nearbyint(x[j]), node.lower[j]), node.upper[j]);
    }
    const bool ok = feasible(md, x, node, opt.feasibility_tol);
    if (ok) {
      const double f = sense * objective_of(md, x);
      if (f < incumbent) {
        incumbent = f;
        out.col_value = x;
        log("new incumbent");
      }
    }
    if (branch < 0) {
      // The relaxation's answer is integral: its rounding is the node's best point. If that point
      // missed a row by more than the tolerance (the continuous columns keep PDLP's accuracy), the
      // node's continuous part is solved again with its integer columns fixed (fix_and_solve).
      ++out.integral;
      if (!ok) {
        auto fixed = std::make_unique<Node>(Node{lb, node.depth, node.lower, node.upper});
        for (int j = 0; j < n; ++j)
          if (integer[j]) fixed->lower[j] = fixed->upper[j] = x[j];
        to_fix.push_back(std::move(fixed));
      }
      return;
    }
    if (lb >= incumbent - gap_abs(incumbent)) {
      ++out.pruned_bound;
      if (lb < incumbent) closed_min = std::min(closed_min, lb);
      return;
    }
    const double v = r.col_value[branch];
    for (int side = 0; side < 2; ++side) {
      auto child = std::make_unique<Node>(Node{lb, node.depth + 1, node.lower, node.upper});
      if (side == 0)
        child->upper[branch] = std::floor(v);
      else
        child->lower[branch] = std::ceil(v);
      if (child->lower[branch] > child->upper[branch]) continue;
      heap.push_back(std::move(child));
      std::push_heap(heap.begin(), heap.end(), heap_less);
    }
  };
  {
    const Node root_node{-kInf, 0, md.col_lower, md.col_upper};
    handle(root_node, root);
    fix_and_solve();
  }

  out.status = Status::kOptimal;
  while (!heap.empty()) {
    if (seconds() > opt.time_limit) {
      out.status = Status::kTimeLimit;
      break;
    }
    if (out.nodes >= opt.max_nodes) {
      out.status = Status::kIterationLimit;
      break;
    }
    // The wave: the open nodes with the smallest bounds that can still beat the incumbent.
    std::vector<NodePtr> wave;
    while (!heap.empty() && static_cast<int>(wave.size()) < opt.wave &&
           out.nodes + static_cast<long long>(wave.size()) < opt.max_nodes) {
      std::pop_heap(heap.begin(), heap.end(), heap_less);
      NodePtr node = std::move(heap.back());
      heap.pop_back();
      if (node->bound >= incumbent - gap_abs(incumbent)) {
        // Every open node's bound is at least this one's: all are pruned.
        out.pruned_bound += 1 + static_cast<long long>(heap.size());
        if (node->bound < incumbent) closed_min = std::min(closed_min, node->bound);
        heap.clear();
        break;
      }
      if (node_propagation && !tighten_columns(md, rows, integer, node->lower, node->upper)) {
        ++out.pruned_infeasible;  // the branching leaves no point that satisfies the rows
        continue;
      }
      wave.push_back(std::move(node));
    }
    if (wave.empty()) break;
    std::vector<const double*> lows, ups;
    for (const NodePtr& node : wave) {
      lows.push_back(node->lower.data());
      ups.push_back(node->upper.data());
    }
    const std::vector<LpSolveResult> rs =
        solve_lp_bounds_batch(md, lows, ups, lpo, start_x.data(), start_y.data());
    ++out.waves;
    out.nodes += static_cast<long long>(wave.size());
    for (size_t k = 0; k < wave.size(); ++k) handle(*wave[k], rs[k]);
    fix_and_solve();
    if (opt.log_level >= 1) {
      double open_min = kInf;
      for (const NodePtr& node : heap) open_min = std::min(open_min, node->bound);
      std::printf("mip: wave %lld: %zu nodes, %zu open, incumbent %.9g, open bound %.9g, %.1f s\n", out.waves,
                  wave.size(), heap.size(), std::isfinite(incumbent) ? sense * incumbent : NAN,
                  std::isfinite(open_min) ? sense * open_min : NAN, seconds());
      std::fflush(stdout);
    }
  }

  // The bound: the incumbent, the open nodes and the nodes closed below the incumbent.
  double bound = std::min(incumbent, closed_min);
  for (const NodePtr& node : heap) bound = std::min(bound, node->bound);
  if (std::isfinite(incumbent)) {
    out.objective = sense * incumbent;
    out.gap = (incumbent - bound) / (1.0 + std::fabs(incumbent));
  }
  // +inf (minimization sense) when every node was pruned infeasible; -inf when an unresolved
  // node had no bound of its own.
  out.bound = sense * bound;
  if (out.status == Status::kOptimal) {
    if (!std::isfinite(incumbent))
      out.status = out.unresolved > 0 ? Status::kNumericalError : Status::kInfeasible;
    else if (out.gap > opt.gap_tol)
      out.status = Status::kNumericalError;  // unresolved nodes hold the bound below the gap
  }
  out.seconds = seconds();
  log("done");
  return out;
}

}  // namespace smy::host
