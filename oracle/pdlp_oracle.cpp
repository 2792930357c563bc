// ORACLE — test infrastructure only. Nothing in the product links, loads or calls this file.
//
// A CPU, FP64, single-threaded restatement of samaya's PDLP path, written to reproduce the
// reference's arithmetic operation by operation (same loop orders, same constants, same branch
// decisions), so that its results can be compared bit for bit with the reference built from its
// own sources (oracle/_ref) and, within rounding, with the HIP path. Every function cites the
// reference file:line it follows. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load the library built from this file.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off, the flags the reference requires for its
// double-double arithmetic, CMakeLists.txt:56-58).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

namespace oracle {

constexpr double kInf = std::numeric_limits<double>::infinity();

// ---------------------------------------------------------------------------------------------
// Compressed sparse column matrix (include/samaya/sparse_matrix.hpp:36-69).
// ---------------------------------------------------------------------------------------------
struct Csc {
  int rows = 0;
  int cols = 0;
  std::vector<int64_t> start{0};
  std::vector<int> index;
  std::vector<double> value;
  int64_t nnz() const { return static_cast<int64_t>(value.size()); }
};

struct Entry {
  int row;
  int col;
  double value;
};

// Sort by (col, row), sum duplicates, drop exact zeros (src/linalg/sparse_matrix.cpp:15-46).
Csc csc_from_entries(int rows, int cols, std::vector<Entry> e) {
  Csc out;
  out.rows = rows;
  out.cols = cols;
  out.start.assign(static_cast<size_t>(cols) + 1, 0);
  std::sort(e.begin(), e.end(), [](const Entry& a, const Entry& b) {
    return a.col != b.col ? a.col < b.col : a.row < b.row;
  });
  out.index.reserve(e.size());
  out.value.reserve(e.size());
  size_t k = 0;
  for (int j = 0; j < cols; ++j) {
    while (k < e.size() && e[k].col == j) {
      const int i = e[k].row;
      double acc = 0.0;
      for (; k < e.size() && e[k].col == j && e[k].row == i; ++k) acc += e[k].value;
      if (acc != 0.0) {
        out.index.push_back(i);
        out.value.push_back(acc);
      }
    }
    out.start[static_cast<size_t>(j) + 1] = static_cast<int64_t>(out.index.size());
  }
  return out;
}

// Counting-sort transpose; keeps indices sorted (src/linalg/sparse_matrix.cpp:48-67).
Csc csc_transpose(const Csc& a) {
  Csc t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.start.assign(static_cast<size_t>(a.rows) + 1, 0);
  t.index.resize(a.index.size());
  t.value.resize(a.value.size());
  for (int i : a.index) ++t.start[static_cast<size_t>(i) + 1];
  for (size_t i = 0; i < static_cast<size_t>(a.rows); ++i) t.start[i + 1] += t.start[i];
  std::vector<int64_t> fill(t.start.begin(), t.start.end() - 1);
  for (int j = 0; j < a.cols; ++j) {
    for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p) {
      const int64_t q = fill[a.index[p]]++;
      t.index[q] = j;
      t.value[q] = a.value[p];
    }
  }
  return t;
}

// ---------------------------------------------------------------------------------------------
// Power-of-two geometric + equilibration scaling (src/linalg/scaling.cpp:60-200).
// ---------------------------------------------------------------------------------------------
struct Scaling {
  std::vector<double> col;
  std::vector<double> row;
};

double pow2_nearest(double v) {  // scaling.cpp:61-64
  if (!(v > 0.0) || !std::isfinite(v)) return 1.0;
  return std::ldexp(1.0, static_cast<int>(std::lround(std::log2(v))));
}
double clamp_scale(double v) { return std::clamp(v, 0x1p-40, 0x1p40); }  // scaling.cpp:67

double magnitude_ratio(const Csc& a, const Scaling& s) {  // scaling.cpp:114-128
  double lo = kInf, hi = 0.0;
  for (int j = 0; j < a.cols; ++j) {
    for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p) {
      const double v = std::fabs(a.value[p]) * s.row[a.index[p]] * s.col[j];
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
  }
  return hi > 0.0 ? hi / lo : 1.0;
}

Scaling compute_scaling(const Csc& a) {  // scaling.cpp:130-200 (default ScalingOptions)
  const int m = a.rows, n = a.cols;
  Scaling s;
  s.row.assign(static_cast<size_t>(m), 1.0);
  s.col.assign(static_cast<size_t>(n), 1.0);
  if (a.nnz() == 0) return s;
  std::vector<double> rmin(static_cast<size_t>(m)), rmax(static_cast<size_t>(m));
  auto row_extremes = [&] {
    std::fill(rmin.begin(), rmin.end(), kInf);
    std::fill(rmax.begin(), rmax.end(), 0.0);
    for (int j = 0; j < n; ++j) {
      for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p) {
        const double v = std::fabs(a.value[p]) * s.col[j];
        rmin[a.index[p]] = std::min(rmin[a.index[p]], v);
        rmax[a.index[p]] = std::max(rmax[a.index[p]], v);
      }
    }
  };
  double ratio = magnitude_ratio(a, s);
  for (int pass = 0; pass < 10 && ratio > 1.0; ++pass) {
    const Scaling before = s;
    row_extremes();
    for (int i = 0; i < m; ++i)
      if (rmax[i] > 0.0) s.row[i] = clamp_scale(1.0 / std::sqrt(rmin[i] * rmax[i]));
    for (int j = 0; j < n; ++j) {
      double lo = kInf, hi = 0.0;
      for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p) {
        const double v = std::fabs(a.value[p]) * s.row[a.index[p]];
        lo = std::min(lo, v);
        hi = std::max(hi, v);
      }
      if (hi > 0.0) s.col[j] = clamp_scale(1.0 / std::sqrt(lo * hi));
    }
    const double next = magnitude_ratio(a, s);
    if (next > ratio) {
      s = before;
      break;
    }
    const bool small_gain = next > 0.9 * ratio;
    ratio = next;
    if (small_gain) break;
  }
  row_extremes();
  for (int i = 0; i < m; ++i)
    if (rmax[i] > 0.0) s.row[i] = clamp_scale(1.0 / rmax[i]);
  for (int j = 0; j < n; ++j) {
    double hi = 0.0;
    for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p)
      hi = std::max(hi, std::fabs(a.value[p]) * s.row[a.index[p]]);
    if (hi > 0.0) s.col[j] = clamp_scale(1.0 / hi);
  }
  for (double& f : s.row) f = pow2_nearest(f);
  for (double& f : s.col) f = pow2_nearest(f);
  return s;
}

Csc scale_matrix(const Csc& a, const std::vector<double>& row, const std::vector<double>& col) {
  // Scaling::apply (scaling.cpp:78-90) and apply_equilibration (pdlp.cpp:190-209) are the same
  // product value * row[i] * col[j], rebuilt through the triplet constructor.
  std::vector<Entry> e;
  e.reserve(a.value.size());
  for (int j = 0; j < a.cols; ++j)
    for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p)
      e.push_back({a.index[p], j, a.value[p] * row[a.index[p]] * col[j]});
  return csc_from_entries(a.rows, a.cols, std::move(e));
}

// ---------------------------------------------------------------------------------------------
// Model and computational form (include/samaya/model.hpp; src/lp/lp_solver.cpp:28-51).
// ---------------------------------------------------------------------------------------------
struct Model {
  int m = 0, n = 0;
  double sense = 1.0;  // -1 maximize
  double offset = 0.0;
  Csc A;
  std::vector<double> obj, col_lower, col_upper, row_lower, row_upper;
};

struct Lp {
  int m = 0, n = 0;
  Csc A, At;
  std::vector<double> cost, lower, upper;  // n + m
};

Lp make_problem(const Model& md, const Scaling& s) {
  Lp lp;
  lp.m = md.m;
  lp.n = md.n;
  lp.A = scale_matrix(md.A, s.row, s.col);
  lp.At = csc_transpose(lp.A);
  const size_t nt = static_cast<size_t>(md.n + md.m);
  lp.cost.assign(nt, 0.0);
  lp.lower.resize(nt);
  lp.upper.resize(nt);
  for (int j = 0; j < md.n; ++j) {
    lp.cost[j] = md.sense * md.obj[j] * s.col[j];
    lp.lower[j] = md.col_lower[j] / s.col[j];
    lp.upper[j] = md.col_upper[j] / s.col[j];
  }
  for (int i = 0; i < md.m; ++i) {
    lp.lower[md.n + i] = md.row_lower[i] * s.row[i];
    lp.upper[md.n + i] = md.row_upper[i] * s.row[i];
  }
  return lp;
}

// ---------------------------------------------------------------------------------------------
// PDLP setup: power iteration and the 2-norm Ruiz variant (src/lp/pdlp.cpp:36-187).
// ---------------------------------------------------------------------------------------------
double operator_norm_bound(const Csc& A, const Csc& At) {  // pdlp.cpp:36-103
  const int m = A.rows, n = A.cols;
  if (m == 0) return 1.0;
  std::vector<double> u(static_cast<size_t>(m)), v(static_cast<size_t>(n));
  std::mt19937_64 gen(12345);
  std::uniform_real_distribution<double> draw(-1.0, 1.0);
  for (int i = 0; i < m; ++i) u[i] = draw(gen);
  double un = 0.0;
  for (double t : u) un += t * t;
  un = std::sqrt(un);
  for (double& t : u) t /= un;
  double est = 0.0;
  for (int it = 0; it < 20; ++it) {
    std::fill(v.begin(), v.end(), 0.0);
    for (int j = 0; j < m; ++j)
      for (int64_t p = At.start[j]; p < At.start[j + 1]; ++p) v[At.index[p]] += At.value[p] * u[j];
    double vn = 0.0;
    for (double t : v) vn += t * t;
    vn = std::sqrt(vn);
    if (vn == 0.0) return 1.0;
    std::fill(u.begin(), u.end(), 0.0);
    for (int j = 0; j < n; ++j) {
      if (v[j] == 0.0) continue;
      for (int64_t p = A.start[j]; p < A.start[j + 1]; ++p) u[A.index[p]] += A.value[p] * v[j];
    }
    un = 0.0;
    for (double t : u) un += t * t;
    un = std::sqrt(un);
    if (un == 0.0) return 1.0;
    est = un / vn;
    for (double& t : u) t /= un;
  }
  const double la = est * 1.01;
  return std::sqrt(la * la + 1.0);
}

void ruiz_2norm(const Csc& A, const Csc& At, std::vector<double>& rs, std::vector<double>& cs,
                int passes) {  // pdlp.cpp:109-187
  const int m = A.rows, n = A.cols;
  rs.assign(static_cast<size_t>(m), 1.0);
  cs.assign(static_cast<size_t>(n), 1.0);
  if (m == 0 || n == 0) return;
  std::vector<double> rn(static_cast<size_t>(m)), cn(static_cast<size_t>(n));
  for (int it = 0; it < passes; ++it) {
    for (int j = 0; j < n; ++j) {
      double acc = 0.0;
      for (int64_t p = A.start[j]; p < A.start[j + 1]; ++p) {
        const double t = A.value[p] * rs[A.index[p]] * cs[j];
        acc += t * t;
      }
      cn[j] = std::sqrt(acc);
    }
    for (int i = 0; i < m; ++i) {
      double acc = 0.0;
      for (int64_t p = At.start[i]; p < At.start[i + 1]; ++p) {
        const double t = At.value[p] * rs[i] * cs[At.index[p]];
        acc += t * t;
      }
      rn[i] = std::sqrt(acc);
    }
    bool settled = true;
    for (int j = 0; j < n; ++j) {
      if (cn[j] > 0) {
        double f = cs[j] / cn[j];
        f = std::max(1e-8, std::min(1e8, f));
        if (std::fabs(f - cs[j]) / std::max(std::fabs(cs[j]), 1.0) > 0.1) settled = false;
        cs[j] = f;
      }
    }
    for (int i = 0; i < m; ++i) {
      if (rn[i] > 0) {
        double f = rs[i] / rn[i];
        f = std::max(1e-8, std::min(1e8, f));
        if (std::fabs(f - rs[i]) / std::max(std::fabs(rs[i]), 1.0) > 0.1) settled = false;
        rs[i] = f;
      }
    }
    if (settled) break;
  }
}

// ---------------------------------------------------------------------------------------------
// The PDLP solve (src/lp/pdlp.cpp:250-1193), CPU arithmetic branch.
// ---------------------------------------------------------------------------------------------
enum Status { kOptimal = 0, kInfeasible, kUnbounded, kIterationLimit, kTimeLimit, kNumericalError };

struct RestartEvent {
  long long iter;
  int to_average;
  double kkt;
  double omega;
};

struct EvalEvent {  // Both measures at one evaluation (current first, then average).
  long long iter;
  double cur[3];
  double avg[3];
};

struct PdlpOptions {
  double tol = 1e-8;
  long long max_iterations = -1;
  double time_limit = kInf;
};

struct PdlpResult {
  Status status = kNumericalError;
  std::vector<double> x, y, dual_ray, primal_ray;
  long long iterations = 0;
  double primal_residual = 0, dual_residual = 0, gap = 0;
  double L = 0;
  std::vector<double> row_scale, col_scale;
  std::vector<RestartEvent> restarts;
  std::vector<EvalEvent> evals;
};

struct Sums {
  double primal_sq = 0, dual_sq = 0, primal_obj = 0, dual_obj = 0, x_sq = 0, r_sq = 0, y_sq = 0;
};
struct Measure {
  double rel_primal, rel_dual, rel_gap, gap, primal_obj, dual_obj;
};

// lambda: the part of the reduced cost g a variable's bounds can absorb (pdlp.cpp:455-463).
inline double absorbable(double g, double lo, double up) {
  const bool l = lo > -kInf, u = up < kInf;
  return l && u ? g : l ? std::max(g, 0.0) : u ? std::min(g, 0.0) : 0.0;
}

PdlpResult solve_pdlp(const Lp& lp, const PdlpOptions& opt) {
  const auto t0 = std::chrono::steady_clock::now();
  auto elapsed = [&] {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  const int m = lp.m, n = lp.n, nt = n + m;
  PdlpResult res;

  std::vector<double> rs, cs;
  ruiz_2norm(lp.A, lp.At, rs, cs, 10);
  const Csc Aq = scale_matrix(lp.A, rs, cs);
  const Csc Atq = scale_matrix(lp.At, cs, rs);

  std::vector<double> c = lp.cost, lo = lp.lower, up = lp.upper;  // pdlp.cpp:284-299
  for (int j = 0; j < n; ++j) {
    c[j] *= cs[j];
    lo[j] /= cs[j];
    up[j] /= cs[j];
  }
  for (int i = 0; i < m; ++i) {
    c[n + i] *= rs[i];
    lo[n + i] *= rs[i];
    up[n + i] *= rs[i];
  }
  const double L = operator_norm_bound(Aq, Atq);
  res.L = L;
  const double eta = 0.99 / L;  // pdlp.cpp:311-315
  double omega = 1.0;
  double tau = eta / omega, sigma = eta * omega;
  const double theta = 1.0;
  long long k_since = 0;

  std::vector<double> v(nt, 0.0), y(m, 0.0), v_prev(nt, 0.0), v_erg(nt, 0.0), y_erg(m, 0.0);
  std::vector<double> resid(std::max(m, 1)), aty(nt), v_new(nt), v_bar(nt), y_new(m);

  long long max_iter = opt.max_iterations;  // pdlp.cpp:365-373
  if (max_iter <= 0) {
    max_iter = std::max<long long>(10000LL, 10000LL * (m + n) / 10);
    if (opt.time_limit >= kInf) max_iter = std::min<long long>(max_iter, 1000000LL);
  }

  double a_colmax = 0.0, at_rowmax = 0.0;  // pdlp.cpp:377-394
  if (m > 0) {
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int64_t p = lp.A.start[j]; p < lp.A.start[j + 1]; ++p) s += lp.A.value[p] * lp.A.value[p];
      a_colmax = std::max(a_colmax, std::sqrt(s));
    }
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      for (int64_t p = lp.At.start[i]; p < lp.At.start[i + 1]; ++p)
        s += lp.At.value[p] * lp.At.value[p];
      at_rowmax = std::max(at_rowmax, std::sqrt(s));
    }
  }
  double kkt_restart = kInf, kkt_prev = kInf;
  std::vector<double> v_rst(nt, 0.0), y_rst(m, 0.0);

  // Sums over an iterate in the model space (pdlp.cpp:411-488).
  auto host_sums = [&](const std::vector<double>& vv, const std::vector<double>& yy) {
    Sums s;
    std::vector<double> xo(n), ro(m), yo(m);
    for (int j = 0; j < n; ++j) xo[j] = vv[j] * cs[j];
    for (int i = 0; i < m; ++i) ro[i] = vv[n + i] / rs[i];
    for (int i = 0; i < m; ++i) yo[i] = yy[i] * rs[i];
    if (m > 0) {
      std::vector<double> ax(m, 0.0);
      for (int j = 0; j < n; ++j) {
        const double xj = xo[j];
        if (xj == 0.0) continue;
        for (int64_t p = lp.A.start[j]; p < lp.A.start[j + 1]; ++p) ax[lp.A.index[p]] += lp.A.value[p] * xj;
      }
      for (int i = 0; i < m; ++i) {
        const double d = ax[i] - ro[i];
        s.primal_sq += d * d;
      }
    }
    std::vector<double> g_aty(nt, 0.0);
    if (m > 0) {
      for (int i = 0; i < m; ++i) {
        const double yi = yo[i];
        if (yi == 0.0) continue;
        for (int64_t p = lp.At.start[i]; p < lp.At.start[i + 1]; ++p)
          g_aty[lp.At.index[p]] += lp.At.value[p] * yi;
      }
      for (int i = 0; i < m; ++i) g_aty[n + i] = -yo[i];
      for (int k = 0; k < nt; ++k) {
        const double g = lp.cost[k] + g_aty[k];
        const double lam = absorbable(g, lp.lower[k], lp.upper[k]);
        s.dual_sq += (g - lam) * (g - lam);
      }
    }
    for (int j = 0; j < n; ++j) s.primal_obj += lp.cost[j] * xo[j];
    for (int i = 0; i < m; ++i) s.primal_obj += lp.cost[n + i] * ro[i];
    for (int k = 0; k < nt; ++k) {
      const double g = lp.cost[k] + g_aty[k];
      const double lam = absorbable(g, lp.lower[k], lp.upper[k]);
      if (lam > 0.0) s.dual_obj += lam * lp.lower[k];
      if (lam < 0.0) s.dual_obj += lam * lp.upper[k];
    }
    for (int j = 0; j < n; ++j) s.x_sq += xo[j] * xo[j];
    for (int i = 0; i < m; ++i) s.r_sq += ro[i] * ro[i];
    for (int i = 0; i < m; ++i) s.y_sq += yo[i] * yo[i];
    return s;
  };
  double c_norm = 0.0;
  for (int k = 0; k < nt; ++k) c_norm += lp.cost[k] * lp.cost[k];
  c_norm = std::sqrt(c_norm);
  auto finish = [&](const Sums& s) {  // pdlp.cpp:493-519
    const double pres = std::sqrt(s.primal_sq), dres = std::sqrt(s.dual_sq);
    const double gap = s.primal_obj - s.dual_obj;
    const double pscale = 1.0 + a_colmax * std::sqrt(s.x_sq) + std::sqrt(s.r_sq);
    const double dscale = 1.0 + at_rowmax * std::sqrt(s.y_sq) + c_norm;
    const double gscale = 1.0 + std::fabs(s.primal_obj) + std::fabs(s.dual_obj);
    return Measure{m > 0 ? pres / pscale : 0.0, m > 0 ? dres / dscale : 0.0, gap / gscale, gap,
                   s.primal_obj, s.dual_obj};
  };
  auto kkt = [&](const Measure& e) {  // pdlp.cpp:524-527
    return std::sqrt(omega * omega * e.rel_primal * e.rel_primal +
                     e.rel_dual * e.rel_dual / (omega * omega) + e.rel_gap * e.rel_gap);
  };

  bool done = false;
  for (long long iter = 0; iter < max_iter && !done; ++iter) {
    if (opt.time_limit < kInf && elapsed() > opt.time_limit) {
      res.status = kTimeLimit;
      res.iterations = iter;
      break;
    }
    const bool evaluate = (iter + 1) % 64 == 0 || iter + 1 == max_iter;
    // One PDHG step (pdlp.cpp:574-641).
    for (int j = 0; j < nt; ++j) v_bar[j] = v[j] + theta * (v[j] - v_prev[j]);
    resid.assign(static_cast<size_t>(m), 0.0);
    if (m > 0) {
      for (int j = 0; j < n; ++j) {
        const double vj = v_bar[j];
        if (vj == 0.0) continue;
        for (int64_t p = Aq.start[j]; p < Aq.start[j + 1]; ++p) resid[Aq.index[p]] += Aq.value[p] * vj;
      }
      for (int i = 0; i < m; ++i) resid[i] -= v_bar[n + i];
      for (int i = 0; i < m; ++i) y_new[i] = y[i] + sigma * resid[i];
    }
    aty.assign(static_cast<size_t>(nt), 0.0);
    if (m > 0) {
      for (int i = 0; i < m; ++i) {
        const double yi = y_new[i];
        if (yi == 0.0) continue;
        for (int64_t p = Atq.start[i]; p < Atq.start[i + 1]; ++p) aty[Atq.index[p]] += Atq.value[p] * yi;
        aty[n + i] = -yi;
      }
    }
    for (int j = 0; j < nt; ++j) v_new[j] = v[j] - tau * (c[j] + aty[j]);
    for (int j = 0; j < nt; ++j) {  // project_box, pdlp.cpp:213-222
      double t = v_new[j];
      if (lo[j] > -kInf && t < lo[j]) t = lo[j];
      if (up[j] < kInf && t > up[j]) t = up[j];
      v_new[j] = t;
    }
    const double alpha = 1.0 / static_cast<double>(k_since + 1);
    ++k_since;
    for (int j = 0; j < nt; ++j) v_erg[j] = (1.0 - alpha) * v_erg[j] + alpha * v_new[j];
    if (m > 0)
      for (int i = 0; i < m; ++i) y_erg[i] = (1.0 - alpha) * y_erg[i] + alpha * y_new[i];
    if (!evaluate) {
      v_prev.swap(v);
      v.swap(v_new);
      y.swap(y_new);
      continue;
    }

    const Measure cur = finish(host_sums(v_new, y_new));
    const Measure avg = finish(host_sums(v_erg, y_erg));
    res.evals.push_back({iter, {cur.rel_primal, cur.rel_dual, cur.rel_gap},
                         {avg.rel_primal, avg.rel_dual, avg.rel_gap}});
    if (cur.rel_primal < opt.tol && cur.rel_dual < opt.tol && std::fabs(cur.rel_gap) < opt.tol) {
      res.status = kOptimal;
      res.iterations = iter + 1;
      res.primal_residual = cur.rel_primal;
      res.dual_residual = cur.rel_dual;
      res.gap = cur.rel_gap;
      v = v_new;
      y = y_new;
      break;
    }
    if (avg.rel_primal < opt.tol && avg.rel_dual < opt.tol && std::fabs(avg.rel_gap) < opt.tol) {
      res.status = kOptimal;
      res.iterations = iter + 1;
      res.primal_residual = avg.rel_primal;
      res.dual_residual = avg.rel_dual;
      res.gap = avg.rel_gap;
      v = v_erg;
      y = y_erg;
      break;
    }

    // Infeasibility / unboundedness detection (pdlp.cpp:715-1043).
    const bool detect = iter > 1000 && m > 0 && ((iter + 1) / 64) % 8 == 0;
    if (detect) {
      std::vector<double> yf(m, 0.0);
      double ymax = 0.0;
      for (int i = 0; i < m; ++i) {
        yf[i] = y_erg[i] * rs[i];
        ymax = std::max(ymax, std::fabs(yf[i]));
      }
      if (ymax > 1e-6) {
        std::vector<double> w(n + m, 0.0);
        for (int i = 0; i < m; ++i) {
          const double yi = yf[i];
          if (yi == 0.0) continue;
          for (int64_t p = lp.At.start[i]; p < lp.At.start[i + 1]; ++p) w[lp.At.index[p]] += lp.At.value[p] * yi;
        }
        for (int i = 0; i < m; ++i) w[n + i] = -yf[i];
        auto range_add = [](double cf, double l, double u, double& fmin, double& fmax) {
          if (cf > 0) {
            fmin += (l > -kInf) ? cf * l : -INFINITY;
            fmax += (u < kInf) ? cf * u : INFINITY;
          } else if (cf < 0) {
            fmin += (u < kInf) ? cf * u : -INFINITY;
            fmax += (l > -kInf) ? cf * l : INFINITY;
          }
        };
        double fmin = 0.0, fmax = 0.0, wsum = 0.0;
        for (int j = 0; j < n; ++j) wsum += std::fabs(w[j]);
        for (int i = 0; i < m; ++i) wsum += std::fabs(w[n + i]);
        const double drop = 1e-9 * wsum;
        for (int j = 0; j < n; ++j) {
          const double cf = std::fabs(w[j]) <= drop ? 0.0 : w[j];
          range_add(cf, lp.lower[j], lp.upper[j], fmin, fmax);
        }
        for (int i = 0; i < m; ++i) {
          double cf = w[n + i];
          cf = std::fabs(cf) <= drop ? 0.0 : cf;
          range_add(cf, lp.lower[n + i], lp.upper[n + i], fmin, fmax);
        }
        double scale = 0.0;
        for (int j = 0; j < n; ++j) {
          const double t = w[j] / ymax;
          if (std::fabs(t) > 1e-9) {
            if (lp.lower[j] > -kInf) scale = std::max(scale, std::fabs(t * lp.lower[j]));
            if (lp.upper[j] < kInf) scale = std::max(scale, std::fabs(t * lp.upper[j]));
          }
        }
        for (int i = 0; i < m; ++i) {
          const double t = -w[n + i] / ymax;
          if (std::fabs(t) > 1e-9) {
            if (lp.lower[n + i] > -kInf) scale = std::max(scale, std::fabs(t * lp.lower[n + i]));
            if (lp.upper[n + i] < kInf) scale = std::max(scale, std::fabs(t * lp.upper[n + i]));
          }
        }
        if (!std::isfinite(scale)) scale = 0.0;
        const double margin = 1e-8 * (1.0 + scale);
        if ((std::isfinite(fmin) && fmin > margin) || (std::isfinite(fmax) && fmax < -margin)) {
          res.status = kInfeasible;
          res.iterations = iter + 1;
          res.primal_residual = cur.rel_primal;
          res.dual_residual = cur.rel_dual;
          res.gap = cur.rel_gap;
          res.dual_ray.assign(y_erg.begin(), y_erg.end());
          v = v_new;
          y = y_new;
          done = true;
          break;
        }
      }
      // Improving ray from the difference of consecutive iterates (pdlp.cpp:852-1042).
      std::vector<double> d(v_new.size());
      for (size_t k = 0; k < v_new.size(); ++k) d[k] = v_new[k] - v[k];
      double dmax = 0.0;
      for (double t : d) dmax = std::max(dmax, std::fabs(t));
      if (dmax > 1e-6) {
        std::vector<double> dir(d.size());
        for (size_t k = 0; k < d.size(); ++k) dir[k] = d[k] / dmax;
        std::vector<double> rc(m, 0.0);
        for (int j = 0; j < n; ++j) {
          const double xj = dir[j] * cs[j];
          if (xj == 0.0) continue;
          for (int64_t p = lp.A.start[j]; p < lp.A.start[j + 1]; ++p) rc[lp.A.index[p]] += lp.A.value[p] * xj;
        }
        for (int i = 0; i < m; ++i) rc[i] -= dir[n + i] / rs[i];
        double pfeas = 0.0;
        for (double t : rc) pfeas += t * t;
        pfeas = std::sqrt(pfeas);
        double cdir = 0.0;
        for (int j = 0; j < n; ++j) cdir += lp.cost[j] * (dir[j] * cs[j]);
        for (int i = 0; i < m; ++i) cdir += lp.cost[n + i] * (dir[n + i] / rs[i]);
        const double box_tol = 1e-4, eq_tol = 1e-3;
        bool ok = true;
        std::vector<double> rx(n);
        for (int j = 0; j < n && ok; ++j) {
          double xj = dir[j] * cs[j];
          const bool lb = lp.lower[j] > -kInf / 2, ub = lp.upper[j] < kInf / 2;
          const bool fixed = lb && ub && (lp.upper[j] - lp.lower[j] <= 1e-10 * (1.0 + std::fabs(lp.lower[j])));
          if (lb && ub) {
            if (std::fabs(xj) > (fixed ? eq_tol : box_tol)) ok = false;
            xj = 0.0;
          } else if (lb) {
            if (xj < -box_tol) ok = false;
            xj = std::max(xj, 0.0);
          } else if (ub) {
            if (xj > box_tol) ok = false;
            xj = std::min(xj, 0.0);
          }
          rx[j] = xj;
        }
        std::vector<double> av(m, 0.0);
        for (int j = 0; j < n; ++j) {
          const double xj = rx[j];
          if (xj == 0.0) continue;
          for (int64_t p = lp.A.start[j]; p < lp.A.start[j + 1]; ++p) av[lp.A.index[p]] += lp.A.value[p] * xj;
        }
        for (int i = 0; i < m && ok; ++i) {
          const double a = av[i];
          const double L_ = lp.lower[n + i], U_ = lp.upper[n + i];
          const bool lb = L_ > -kInf / 2, ub = U_ < kInf / 2;
          const bool fixed = lb && ub && (U_ - L_ <= 1e-10 * (1.0 + std::fabs(L_)));
          if (lb && ub) {
            if (std::fabs(a) > (fixed ? eq_tol : box_tol)) ok = false;
          } else if (lb) {
            if (a < -box_tol) ok = false;
          } else if (ub) {
            if (a > box_tol) ok = false;
          }
        }
        if (pfeas < 1e-6 && cdir < -1e-8 && ok) {
          res.status = kUnbounded;
          res.iterations = iter + 1;
          res.primal_residual = cur.rel_primal;
          res.dual_residual = cur.rel_dual;
          res.gap = cur.rel_gap;
          res.primal_ray = rx;
          v = v_new;
          y = y_new;
          done = true;
          break;
        }
      }
    }

    // Adaptive restart and primal weight (pdlp.cpp:1075-1132).
    {
      const double kc = kkt(cur), ka = kkt(avg);
      const bool to_avg = ka < kc;
      const double kcand = std::min(kc, ka);
      if (!(kkt_restart < kInf)) kkt_restart = kcand;
      const bool restart = kcand <= 0.2 * kkt_restart ||
                           (kcand <= 0.8 * kkt_restart && kcand > kkt_prev) ||
                           static_cast<double>(k_since) >= 0.36 * static_cast<double>(iter + 1);
      kkt_prev = kcand;
      if (restart) {
        const std::vector<double>& vc = to_avg ? v_erg : v_new;
        const std::vector<double>& yc = to_avg ? y_erg : y_new;
        double dx = 0.0, dy = 0.0;
        for (int k = 0; k < nt; ++k) dx += (vc[k] - v_rst[k]) * (vc[k] - v_rst[k]);
        for (int i = 0; i < m; ++i) dy += (yc[i] - y_rst[i]) * (yc[i] - y_rst[i]);
        dx = std::sqrt(dx);
        dy = std::sqrt(dy);
        if (dx > 1e-10 && dy > 1e-10) {
          omega = std::exp(0.5 * std::log(dy / dx) + 0.5 * std::log(omega));
          tau = eta / omega;
          sigma = eta * omega;
        }
        res.restarts.push_back({iter, to_avg ? 1 : 0, kcand, omega});
        v_rst = vc;
        y_rst = yc;
        v_new = v_rst;
        v = v_rst;
        y_new = y_rst;
        k_since = 0;
        kkt_restart = kcand;
        kkt_prev = kInf;
      }
    }
    v_prev.swap(v);
    v.swap(v_new);
    y.swap(y_new);
  }
  if (res.status == kNumericalError) {
    res.status = kIterationLimit;
    res.iterations = max_iter;
  }
  res.x.assign(nt, 0.0);
  for (int j = 0; j < n; ++j) res.x[j] = v[j] * cs[j];
  for (int i = 0; i < m; ++i) res.x[n + i] = v[n + i] / rs[i];
  res.y.assign(m, 0.0);
  for (int i = 0; i < m; ++i) res.y[i] = y[i] * rs[i];
  res.row_scale = rs;
  res.col_scale = cs;
  return res;
}

// ---------------------------------------------------------------------------------------------
// Double-double (src/core/xreal.hpp:17-137): Knuth TwoSum, Dekker TwoProduct with Veltkamp split.
// ---------------------------------------------------------------------------------------------
struct DD {
  double hi = 0.0, lo = 0.0;
  DD() = default;
  DD(double v) : hi(v), lo(0.0) {}  // NOLINT
  double value() const { return hi + lo; }
  static void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
  }
  static void fast_two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    e = b - (s - a);
  }
  static void split(double a, double& h, double& l) {
    const double kS = 134217729.0, kLim = 6.69692879491417e+299;
    if (a > kLim || a < -kLim) {
      a *= 3.7252902984619140625e-09;
      const double t = kS * a;
      h = t - (t - a);
      l = a - h;
      h *= 268435456.0;
      l *= 268435456.0;
      return;
    }
    const double t = kS * a;
    h = t - (t - a);
    l = a - h;
  }
  static void two_prod(double a, double b, double& p, double& e) {
    p = a * b;
    if (!std::isfinite(p)) {
      e = 0.0;
      return;
    }
    double ah, al, bh, bl;
    split(a, ah, al);
    split(b, bh, bl);
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
  }
  DD& add(const DD& o) {
    if (!std::isfinite(hi) || !std::isfinite(o.hi)) {
      hi = hi + o.hi;
      lo = 0.0;
      return *this;
    }
    double s, e, t, f;
    two_sum(hi, o.hi, s, e);
    two_sum(lo, o.lo, t, f);
    e += t;
    fast_two_sum(s, e, s, e);
    e += f;
    fast_two_sum(s, e, hi, lo);
    if (!std::isfinite(hi)) lo = 0.0;
    return *this;
  }
  DD neg() const {
    DD r;
    r.hi = -hi;
    r.lo = -lo;
    return r;
  }
  DD& mul(double b) {
    if (!std::isfinite(hi) || !std::isfinite(b)) {
      hi = hi * b;
      lo = 0.0;
      return *this;
    }
    double p, e;
    two_prod(hi, b, p, e);
    if (!std::isfinite(p)) {
      hi = p;
      lo = 0.0;
      return *this;
    }
    e += lo * b;
    fast_two_sum(p, e, hi, lo);
    return *this;
  }
  bool positive() const { return DD(0.0).less(*this); }
  bool negative() const { return less(DD(0.0)); }
  bool less(const DD& b) const { return hi < b.hi || (hi == b.hi && lo < b.lo); }
  DD abs() const { return (hi < 0.0 || (hi == 0.0 && lo < 0.0)) ? neg() : *this; }
};
inline DD dd_mul(double a, double b) {
  DD r(a);
  r.mul(b);
  return r;
}

// ---------------------------------------------------------------------------------------------
// Verifier for LP (src/verify/verify.cpp:28-403, LP parts only: no Q, no quadratic rows).
// ---------------------------------------------------------------------------------------------
struct Report {
  bool ok = true;
  double bound_violation = 0, row_violation = 0, dual_violation = 0, objective = 0;
  std::string message;
};

void flag(Report& r, double violation, double tol, const char* what, int idx) {  // verify.cpp:18-27
  if (violation <= tol) return;
  if (r.ok || r.message.empty()) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s %d violated by %.3e (tolerance %.1e)", what, idx, violation, tol);
    r.message = buf;
  }
  r.ok = false;
}

double rel_bound_violation(double v, double l, double u, double scale) {  // verify.cpp:30-35
  double viol = 0.0;
  if (v < l) viol = (l - v) / (1.0 + std::max(std::fabs(l), scale));
  if (v > u) viol = (v - u) / (1.0 + std::max(std::fabs(u), scale));
  return viol;
}

Report verify_primal(const Model& md, const double* x, double ptol) {  // verify.cpp:52-128
  Report r;
  for (int j = 0; j < md.n; ++j) {
    if (!std::isfinite(x[j])) {
      r.ok = false;
      r.message = "solution contains a non-finite value";
      return r;
    }
    const double v = rel_bound_violation(x[j], md.col_lower[j], md.col_upper[j], 0.0);
    r.bound_violation = std::max(r.bound_violation, v);
    flag(r, v, ptol, "column bound", j);
  }
  std::vector<DD> act(md.m);
  std::vector<double> mag(md.m, 0.0);
  for (int j = 0; j < md.n; ++j)
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const DD t = dd_mul(md.A.value[p], x[j]);
      act[md.A.index[p]].add(t);
      mag[md.A.index[p]] = std::max(mag[md.A.index[p]], t.abs().value());
    }
  for (int i = 0; i < md.m; ++i) {
    const double a = act[i].value();
    if (!std::isfinite(a) || !std::isfinite(mag[i])) {
      r.ok = false;
      r.row_violation = kInf;
      if (r.message.empty()) r.message = "row " + std::to_string(i) + " activity is not finite";
      continue;
    }
    const double v = rel_bound_violation(a, md.row_lower[i], md.row_upper[i], mag[i]);
    r.row_violation = std::max(r.row_violation, v);
    flag(r, v, ptol, "row", i);
  }
  DD obj(md.offset);
  for (int j = 0; j < md.n; ++j) obj.add(dd_mul(md.obj[j], x[j]));
  r.objective = obj.value();
  return r;
}

Report verify_lp_optimality(const Model& md, const double* x, const double* y, double ptol,
                            double dtol) {  // verify.cpp:132-255 with Q and quadratic rows empty
  Report r = verify_primal(md, x, ptol);
  const double sense = md.sense;
  auto at_bound = [&](double v, double b, double scale) {
    return std::isfinite(b) && std::fabs(v - b) <= ptol * (1.0 + std::max(std::fabs(b), scale));
  };
  std::vector<DD> act(md.m);
  std::vector<double> mag(md.m, 0.0);
  for (int j = 0; j < md.n; ++j) {
    // d = obj + qx - jty with qx = jty = 0: the double-double sums of two zeros leave obj.
    DD d(md.obj[j]);
    d.add(DD(0.0));
    d.add(DD(0.0).neg());
    double scale = std::fabs(md.obj[j]) + 0.0 + 0.0;
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const DD t = dd_mul(md.A.value[p], y[md.A.index[p]]);
      d.add(t.neg());
      scale += t.abs().value();
      act[md.A.index[p]].add(dd_mul(md.A.value[p], x[j]));
      mag[md.A.index[p]] = std::max(mag[md.A.index[p]], std::fabs(md.A.value[p] * x[j]));
    }
    const double dj = sense * d.value();
    if (!std::isfinite(dj) || !std::isfinite(scale)) {
      r.ok = false;
      r.dual_violation = kInf;
      if (r.message.empty()) r.message = "reduced cost of column " + std::to_string(j) + " is not finite";
      continue;
    }
    const bool atl = at_bound(x[j], md.col_lower[j], 0.0);
    const bool atu = at_bound(x[j], md.col_upper[j], 0.0);
    double viol = 0.0;
    if (!atl) viol = std::max(viol, dj);
    if (!atu) viol = std::max(viol, -dj);
    viol /= 1.0 + scale;
    r.dual_violation = std::max(r.dual_violation, viol);
    flag(r, viol, dtol, "reduced cost of column", j);
  }
  for (int i = 0; i < md.m; ++i) {
    if (!std::isfinite(y[i])) {
      r.ok = false;
      r.message = "dual vector contains a non-finite value";
      return r;
    }
    const double yi = sense * y[i];
    const double a = act[i].value();
    const bool atl = at_bound(a, md.row_lower[i], mag[i]);
    const bool atu = at_bound(a, md.row_upper[i], mag[i]);
    double viol = 0.0;
    if (!atl) viol = std::max(viol, yi);
    if (!atu) viol = std::max(viol, -yi);
    viol /= 1.0 + std::fabs(yi);
    r.dual_violation = std::max(r.dual_violation, viol);
    flag(r, viol, dtol, "dual of row", i);
  }
  return r;
}

Report verify_infeasibility(const Model& md, const double* y, double ptol) {  // verify.cpp:269-341
  Report r;
  double ymax = 0.0;
  for (int i = 0; i < md.m; ++i) {
    if (!std::isfinite(y[i])) {
      r.ok = false;
      r.message = "certificate contains a non-finite value";
      return r;
    }
    ymax = std::max(ymax, std::fabs(y[i]));
  }
  if (ymax == 0.0) {
    r.ok = false;
    r.message = "certificate is zero";
    return r;
  }
  DD lo(0.0), hi(0.0);
  double scale = 0.0;
  auto add_range = [&](const DD& cf, double l, double u) {
    if (cf.positive()) {
      lo.add(l > -kInf ? DD(cf).mul(l) : DD(-INFINITY));
      hi.add(u < kInf ? DD(cf).mul(u) : DD(INFINITY));
    } else if (cf.negative()) {
      lo.add(u < kInf ? DD(cf).mul(u) : DD(-INFINITY));
      hi.add(l > -kInf ? DD(cf).mul(l) : DD(INFINITY));
    }
  };
  for (int j = 0; j < md.n; ++j) {
    DD w(0.0);
    double ws = 0.0;
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const DD t = dd_mul(md.A.value[p], y[md.A.index[p]] / ymax);
      w.add(t);
      ws += t.abs().value();
    }
    if (std::fabs(w.value()) <= 1e-9 * ws) w = DD(0.0);
    add_range(w, md.col_lower[j], md.col_upper[j]);
    if (w.hi != 0.0 || w.lo != 0.0)
      scale = std::max({scale, std::fabs(w.value() * md.col_lower[j]), std::fabs(w.value() * md.col_upper[j])});
  }
  for (int i = 0; i < md.m; ++i) {
    const DD cf = DD(y[i] / ymax).neg();
    add_range(cf, md.row_lower[i], md.row_upper[i]);
    if (cf.hi != 0.0 || cf.lo != 0.0)
      scale = std::max({scale, std::fabs(cf.value() * md.row_lower[i]), std::fabs(cf.value() * md.row_upper[i])});
  }
  if (!std::isfinite(scale)) scale = 0.0;
  const double margin = ptol * (1.0 + scale);
  if (DD(margin).less(lo) || hi.less(DD(-margin))) return r;
  r.ok = false;
  char buf[160];
  std::snprintf(buf, sizeof buf, "certificate range [%.3e, %.3e] does not exclude zero", lo.value(), hi.value());
  r.message = buf;
  return r;
}

Report verify_unbounded_ray(const Model& md, const double* ray, double ptol, double dtol) {  // verify.cpp:345-401
  Report r;
  double vmax = 0.0;
  for (int j = 0; j < md.n; ++j) vmax = std::max(vmax, std::fabs(ray[j]));
  if (!(vmax > 0.0) || !std::isfinite(vmax)) {
    r.ok = false;
    r.message = "ray is zero or not finite";
    return r;
  }
  DD slope(0.0);
  double cmax = 0.0;
  for (int j = 0; j < md.n; ++j) {
    const double v = ray[j] / vmax;
    slope.add(dd_mul(md.sense * md.obj[j], v));
    cmax = std::max(cmax, std::fabs(md.obj[j]));
    double viol = 0.0;
    if (v < 0.0 && md.col_lower[j] > -kInf) viol = -v;
    if (v > 0.0 && md.col_upper[j] < kInf) viol = v;
    r.bound_violation = std::max(r.bound_violation, viol);
    flag(r, viol, ptol, "ray leaves the bounds of column", j);
  }
  std::vector<DD> av(md.m);
  std::vector<double> mag(md.m, 0.0);
  for (int j = 0; j < md.n; ++j)
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const double t = md.A.value[p] * ray[j] / vmax;
      av[md.A.index[p]].add(DD(t));
      mag[md.A.index[p]] = std::max(mag[md.A.index[p]], std::fabs(t));
    }
  for (int i = 0; i < md.m; ++i) {
    const double a = av[i].value();
    double viol = 0.0;
    if (a < 0.0 && md.row_lower[i] > -kInf) viol = -a / (1.0 + mag[i]);
    if (a > 0.0 && md.row_upper[i] < kInf) viol = a / (1.0 + mag[i]);
    r.row_violation = std::max(r.row_violation, viol);
    flag(r, viol, ptol, "ray leaves the bounds of row", i);
  }
  if (!slope.less(DD(-dtol * (1.0 + cmax)))) {
    r.ok = false;
    r.message = "ray does not improve the objective";
  }
  return r;
}

}  // namespace oracle

// =============================================================================================
// C ABI for the Python test harness (ctypes). Plain pointers and sizes only.
// =============================================================================================
extern "C" {

struct oracle_model {  // A model in the reference's Model form (include/samaya/model.hpp:41-80).
  int m, n;
  int maximize;
  double offset;
  const int64_t* start;
  const int* index;
  const double* value;
  const double* obj;
  const double* col_lower;
  const double* col_upper;
  const double* row_lower;
  const double* row_upper;
};

struct oracle_lp {  // LpProblem in computational form (src/lp/simplex.hpp:18-26); A as CSC.
  int m, n;
  const int64_t* start;
  const int* index;
  const double* value;
  const double* cost;
  const double* lower;
  const double* upper;
};

struct oracle_pdlp_out {
  int status;
  long long iterations;
  double primal_residual, dual_residual, gap, L;
  double* x;          // n + m, may be null
  double* y;          // m, may be null
  double* dual_ray;   // m, may be null (written when infeasible)
  double* primal_ray; // n, may be null (written when unbounded)
  double* row_scale;  // m, may be null (Ruiz)
  double* col_scale;  // n, may be null (Ruiz)
  long long* restart_iter;  // restart log, capacity restart_cap
  int* restart_to_average;
  double* restart_kkt;
  double* restart_omega;
  int restart_cap;
  int restart_count;
  long long* eval_iter;  // evaluation log (rel_primal, rel_dual, rel_gap of current then average)
  double* eval_measures;  // 6 per evaluation
  int eval_cap;
  int eval_count;
  double seconds;
};

struct oracle_lp_out {  // The model-space answer of solve_pdlp_lp (src/core/solver.cpp:392-477).
  int status;
  long long iterations;
  double objective;
  int verified;
  double max_primal_violation, max_dual_violation;
  double* col_value;  // n
  double* row_dual;   // m
  double* col_dual;   // n
  double* ray;        // m (infeasible) or n (unbounded), may be null
  char message[256];
  double seconds;
};

}  // extern "C"

namespace {

oracle::Model to_model(const oracle_model* p) {
  oracle::Model md;
  md.m = p->m;
  md.n = p->n;
  md.sense = p->maximize ? -1.0 : 1.0;
  md.offset = p->offset;
  md.A.rows = p->m;
  md.A.cols = p->n;
  md.A.start.assign(p->start, p->start + p->n + 1);
  const int64_t nnz = md.A.start.back();
  md.A.index.assign(p->index, p->index + nnz);
  md.A.value.assign(p->value, p->value + nnz);
  md.obj.assign(p->obj, p->obj + p->n);
  md.col_lower.assign(p->col_lower, p->col_lower + p->n);
  md.col_upper.assign(p->col_upper, p->col_upper + p->n);
  md.row_lower.assign(p->row_lower, p->row_lower + p->m);
  md.row_upper.assign(p->row_upper, p->row_upper + p->m);
  return md;
}

oracle::Lp to_lp(const oracle_lp* p) {
  oracle::Lp lp;
  lp.m = p->m;
  lp.n = p->n;
  lp.A.rows = p->m;
  lp.A.cols = p->n;
  lp.A.start.assign(p->start, p->start + p->n + 1);
  const int64_t nnz = lp.A.start.back();
  lp.A.index.assign(p->index, p->index + nnz);
  lp.A.value.assign(p->value, p->value + nnz);
  lp.At = oracle::csc_transpose(lp.A);
  const size_t nt = static_cast<size_t>(p->n + p->m);
  lp.cost.assign(p->cost, p->cost + nt);
  lp.lower.assign(p->lower, p->lower + nt);
  lp.upper.assign(p->upper, p->upper + nt);
  return lp;
}

void copy_out(const std::vector<double>& v, double* dst) {
  if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(double));
}

void fill_pdlp_out(const oracle::PdlpResult& r, oracle_pdlp_out* out) {
  out->status = r.status;
  out->iterations = r.iterations;
  out->primal_residual = r.primal_residual;
  out->dual_residual = r.dual_residual;
  out->gap = r.gap;
  out->L = r.L;
  copy_out(r.x, out->x);
  copy_out(r.y, out->y);
  copy_out(r.dual_ray, out->dual_ray);
  copy_out(r.primal_ray, out->primal_ray);
  copy_out(r.row_scale, out->row_scale);
  copy_out(r.col_scale, out->col_scale);
  out->restart_count = static_cast<int>(r.restarts.size());
  for (int k = 0; k < out->restart_count && k < out->restart_cap; ++k) {
    if (out->restart_iter) out->restart_iter[k] = r.restarts[k].iter;
    if (out->restart_to_average) out->restart_to_average[k] = r.restarts[k].to_average;
    if (out->restart_kkt) out->restart_kkt[k] = r.restarts[k].kkt;
    if (out->restart_omega) out->restart_omega[k] = r.restarts[k].omega;
  }
  out->eval_count = static_cast<int>(r.evals.size());
  for (int k = 0; k < out->eval_count && k < out->eval_cap; ++k) {
    if (out->eval_iter) out->eval_iter[k] = r.evals[k].iter;
    if (out->eval_measures)
      for (int q = 0; q < 3; ++q) {
        out->eval_measures[6 * k + q] = r.evals[k].cur[q];
        out->eval_measures[6 * k + 3 + q] = r.evals[k].avg[q];
      }
  }
}

}  // namespace

extern "C" {

// solve_pdlp on an LpProblem (src/lp/pdlp.cpp:228-246, CPU branch).
int oracle_solve_pdlp(const oracle_lp* lp, double tol, long long max_iterations,
                      oracle_pdlp_out* out) {
  const auto t0 = std::chrono::steady_clock::now();
  oracle::PdlpOptions o;
  o.tol = tol;
  o.max_iterations = max_iterations;
  const oracle::PdlpResult r = oracle::solve_pdlp(to_lp(lp), o);
  fill_pdlp_out(r, out);
  out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// compute_scaling + make_problem (src/linalg/scaling.cpp:130; src/lp/lp_solver.cpp:28): writes
// the power-of-two scales and the computational-form cost/bounds (n + m) and the scaled values
// of A in the model's CSC order (duplicates and zeros must already be absent).
int oracle_make_problem(const oracle_model* model, double* col_scale, double* row_scale,
                        double* cost, double* lower, double* upper, double* a_value) {
  const oracle::Model md = to_model(model);
  const oracle::Scaling s = oracle::compute_scaling(md.A);
  const oracle::Lp lp = oracle::make_problem(md, s);
  copy_out(s.col, col_scale);
  copy_out(s.row, row_scale);
  copy_out(lp.cost, cost);
  copy_out(lp.lower, lower);
  copy_out(lp.upper, upper);
  copy_out(lp.A.value, a_value);
  return static_cast<int>(lp.A.nnz());
}

// Ruiz variant of solve_pdlp (pdlp.cpp:109-187) on an LpProblem's A.
void oracle_ruiz(const oracle_lp* lp, int passes, double* row_scale, double* col_scale) {
  const oracle::Lp p = to_lp(lp);
  std::vector<double> rs, cs;
  oracle::ruiz_2norm(p.A, p.At, rs, cs, passes);
  copy_out(rs, row_scale);
  copy_out(cs, col_scale);
}

// Power-iteration bound on ||[A -I]|| (pdlp.cpp:36-103) for a CSC matrix.
double oracle_operator_norm(int m, int n, const int64_t* start, const int* index,
                            const double* value) {
  oracle::Csc a;
  a.rows = m;
  a.cols = n;
  a.start.assign(start, start + n + 1);
  a.index.assign(index, index + a.start.back());
  a.value.assign(value, value + a.start.back());
  return oracle::operator_norm_bound(a, oracle::csc_transpose(a));
}

// The whole LP-by-PDLP path of Solver::solve (src/core/solver.cpp:392-477): power-of-two scaling,
// computational form, PDLP, mapping back (unscale, y <- -sense y, d = c - A'y, double-double
// objective) and the independent verifier at its default tolerances.
int oracle_solve_lp(const oracle_model* model, double tol, long long max_iterations,
                    oracle_lp_out* out, oracle_pdlp_out* pdlp_out) {
  const auto t0 = std::chrono::steady_clock::now();
  const oracle::Model md = to_model(model);
  const oracle::Scaling s = oracle::compute_scaling(md.A);
  const oracle::Lp lp = oracle::make_problem(md, s);
  oracle::PdlpOptions o;
  o.tol = tol;
  o.max_iterations = max_iterations;
  const oracle::PdlpResult r = oracle::solve_pdlp(lp, o);
  if (pdlp_out) fill_pdlp_out(r, pdlp_out);
  const int n = md.n, m = md.m;
  std::vector<double> x(r.x.begin(), r.x.begin() + n);
  for (int j = 0; j < n; ++j) x[j] *= s.col[j];
  std::vector<double> y = r.y;
  for (int i = 0; i < m; ++i) y[i] *= s.row[i];
  for (double& t : y) t *= -md.sense;
  std::vector<double> d(n, 0.0);
  for (int j = 0; j < n; ++j) {  // SparseMatrix::multiply_transpose (sparse_matrix.cpp:80-88)
    double acc = 0.0;
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) acc += md.A.value[p] * y[md.A.index[p]];
    d[j] = acc;
  }
  oracle::DD obj(md.offset);
  for (int j = 0; j < n; ++j) {
    d[j] = md.obj[j] - d[j];
    obj.add(oracle::dd_mul(md.obj[j], x[j]));
  }
  out->status = r.status;
  out->iterations = r.iterations;
  out->objective = obj.value();
  copy_out(x, out->col_value);
  copy_out(y, out->row_dual);
  copy_out(d, out->col_dual);
  oracle::Report rep;
  bool answered = true;
  if (r.status == oracle::kOptimal) {
    rep = oracle::verify_lp_optimality(md, x.data(), y.data(), 1e-6, 1e-6);
  } else if (r.status == oracle::kInfeasible) {
    std::vector<double> ray = r.dual_ray;
    for (int i = 0; i < m; ++i) ray[i] *= s.row[i];
    for (double& t : ray) t *= md.sense;
    copy_out(ray, out->ray);
    rep = oracle::verify_infeasibility(md, ray.data(), 1e-6);
  } else if (r.status == oracle::kUnbounded) {
    std::vector<double> ray = r.primal_ray;
    for (int j = 0; j < n; ++j) ray[j] *= s.col[j];
    copy_out(ray, out->ray);
    rep = oracle::verify_unbounded_ray(md, ray.data(), 1e-6, 1e-6);
  } else {
    answered = false;
  }
  out->verified = answered && rep.ok;
  out->max_primal_violation = std::max(rep.bound_violation, rep.row_violation);
  out->max_dual_violation = rep.dual_violation;
  std::snprintf(out->message, sizeof out->message, "%s", rep.message.c_str());
  out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// verify_lp_optimality on model-space x (n) and y (m) (verify.cpp:251).
int oracle_verify_lp(const oracle_model* model, const double* x, const double* y, double ptol,
                     double dtol, double* out3, char* msg, int msglen) {
  const oracle::Model md = to_model(model);
  const oracle::Report r = oracle::verify_lp_optimality(md, x, y, ptol, dtol);
  out3[0] = r.bound_violation > r.row_violation ? r.bound_violation : r.row_violation;
  out3[1] = r.dual_violation;
  out3[2] = r.objective;
  if (msg && msglen > 0) std::snprintf(msg, static_cast<size_t>(msglen), "%s", r.message.c_str());
  return r.ok ? 1 : 0;
}

// The seven sums of PdhgDevice::Sums (src/gpu/pdhg_core.hpp:58-66) for a scaled iterate (v, y),
// computed as solve_pdlp's host evaluation does (pdlp.cpp:411-488), from the unequilibrated A of
// the LpProblem and the equilibration scales rs (m) and cs (n).
void oracle_sums(const oracle_lp* lpp, const double* rs, const double* cs, const double* v,
                 const double* y, double* out7) {
  const oracle::Lp lp = to_lp(lpp);
  const int m = lp.m, n = lp.n, nt = n + m;
  std::vector<double> xo(n), ro(m), yo(m), ax(m, 0.0), g_aty(nt, 0.0);
  for (int j = 0; j < n; ++j) xo[j] = v[j] * cs[j];
  for (int i = 0; i < m; ++i) ro[i] = v[n + i] / rs[i];
  for (int i = 0; i < m; ++i) yo[i] = y[i] * rs[i];
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < n; ++j)
    for (int64_t p = lp.A.start[j]; p < lp.A.start[j + 1]; ++p) ax[lp.A.index[p]] += lp.A.value[p] * xo[j];
  for (int i = 0; i < m; ++i) s[0] += (ax[i] - ro[i]) * (ax[i] - ro[i]);
  for (int i = 0; i < m; ++i)
    for (int64_t p = lp.At.start[i]; p < lp.At.start[i + 1]; ++p) g_aty[lp.At.index[p]] += lp.At.value[p] * yo[i];
  for (int i = 0; i < m; ++i) g_aty[n + i] = -yo[i];
  for (int k = 0; k < nt; ++k) {
    const double g = lp.cost[k] + g_aty[k];
    const double lam = oracle::absorbable(g, lp.lower[k], lp.upper[k]);
    s[1] += (g - lam) * (g - lam);
    if (lam > 0.0) s[3] += lam * lp.lower[k];
    if (lam < 0.0) s[3] += lam * lp.upper[k];
    s[2] += lp.cost[k] * (k < n ? xo[k] : ro[k - n]);
  }
  for (int j = 0; j < n; ++j) s[4] += xo[j] * xo[j];
  for (int i = 0; i < m; ++i) s[5] += ro[i] * ro[i];
  for (int i = 0; i < m; ++i) s[6] += yo[i] * yo[i];
  std::memcpy(out7, s, sizeof s);
}

// `steps` PDHG steps from v = y = 0 on an equilibrated computational-form LP given directly as
// CSC A (m x n) with its cost and bounds, fixed tau, sigma, theta and no restarts
// (pdlp.cpp:574-641 with the host's advance between steps). Writes the last step's v_new, y_new
// and the running averages.
void oracle_pdhg_steps(const oracle_lp* lpp, double tau, double sigma, double theta, int steps,
                       double* v_new_out, double* y_new_out, double* v_erg_out, double* y_erg_out) {
  const oracle::Lp lp = to_lp(lpp);
  const int m = lp.m, n = lp.n, nt = n + m;
  std::vector<double> v(nt, 0.0), vp(nt, 0.0), vb(nt), vn(nt), ve(nt, 0.0), y(m, 0.0), yn(m, 0.0),
      ye(m, 0.0), rsd(m), aty(nt);
  for (int s = 0; s < steps; ++s) {
    for (int j = 0; j < nt; ++j) vb[j] = v[j] + theta * (v[j] - vp[j]);
    std::fill(rsd.begin(), rsd.end(), 0.0);
    for (int j = 0; j < n; ++j) {
      if (vb[j] == 0.0) continue;
      for (int64_t p = lp.A.start[j]; p < lp.A.start[j + 1]; ++p) rsd[lp.A.index[p]] += lp.A.value[p] * vb[j];
    }
    for (int i = 0; i < m; ++i) yn[i] = y[i] + sigma * (rsd[i] - vb[n + i]);
    std::fill(aty.begin(), aty.end(), 0.0);
    for (int i = 0; i < m; ++i) {
      if (yn[i] == 0.0) continue;
      for (int64_t p = lp.At.start[i]; p < lp.At.start[i + 1]; ++p) aty[lp.At.index[p]] += lp.At.value[p] * yn[i];
      aty[n + i] = -yn[i];
    }
    for (int j = 0; j < nt; ++j) {
      double t = v[j] - tau * (lp.cost[j] + aty[j]);
      if (lp.lower[j] > -oracle::kInf && t < lp.lower[j]) t = lp.lower[j];
      if (lp.upper[j] < oracle::kInf && t > lp.upper[j]) t = lp.upper[j];
      vn[j] = t;
    }
    const double a = 1.0 / static_cast<double>(s + 1);
    for (int j = 0; j < nt; ++j) ve[j] = (1.0 - a) * ve[j] + a * vn[j];
    for (int i = 0; i < m; ++i) ye[i] = (1.0 - a) * ye[i] + a * yn[i];
    if (s + 1 < steps) {
      vp.swap(v);
      v.swap(vn);
      y.swap(yn);
    }
  }
  copy_out(vn, v_new_out);
  copy_out(yn, y_new_out);
  copy_out(ve, v_erg_out);
  copy_out(ye, y_erg_out);
}

// Brute-force CSR product y = A x (tests/test_gpu.cpp:97-107), row by row in index order.
void oracle_csr_spmv(int rows, const int* ptr, const int* idx, const double* val, const double* x,
                     double* y) {
  for (int i = 0; i < rows; ++i) {
    double acc = 0.0;
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) acc += val[p] * x[idx[p]];
    y[i] = acc;
  }
}

}  // extern "C"
