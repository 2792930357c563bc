// Independent LP answer checks in double-double arithmetic and the model-level LP-by-PDLP path.
//
// The checks recompute row activities and reduced costs from the model alone and share no code
// with the solver; they follow samaya's verifier (src/verify/verify.cpp:28-401, LP parts) and its
// double-double type (src/core/xreal.hpp): Knuth's TwoSum and Dekker's TwoProduct with a
// Veltkamp split, so the sums do not depend on FMA or on the compiler (this file is built with
// -ffp-contract=off).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "pdlp.hpp"

namespace smy::host {
namespace {

class DoubleDouble {
 public:
  DoubleDouble() = default;
  DoubleDouble(double v) : hi_(v) {}  // NOLINT: widening
  double value() const { return hi_ + lo_; }
  bool is_zero() const { return hi_ == 0.0 && lo_ == 0.0; }
  bool less(const DoubleDouble& b) const { return hi_ < b.hi_ || (hi_ == b.hi_ && lo_ < b.lo_); }
  DoubleDouble negated() const {
    DoubleDouble r;
    r.hi_ = -hi_;
    r.lo_ = -lo_;
    return r;
  }
  DoubleDouble magnitude() const { return (hi_ < 0.0 || (hi_ == 0.0 && lo_ < 0.0)) ? negated() : *this; }

  DoubleDouble& operator+=(const DoubleDouble& o) {
    if (!std::isfinite(hi_) || !std::isfinite(o.hi_)) return nonfinite(hi_ + o.hi_);
    double s, e, t, f;
    sum_exact(hi_, o.hi_, s, e);
    sum_exact(lo_, o.lo_, t, f);
    e += t;
    sum_ordered(s, e, s, e);
    e += f;
    sum_ordered(s, e, hi_, lo_);
    if (!std::isfinite(hi_)) lo_ = 0.0;
    return *this;
  }
  DoubleDouble& operator*=(double b) {
    if (!std::isfinite(hi_) || !std::isfinite(b)) return nonfinite(hi_ * b);
    double p, e;
    product_exact(hi_, b, p, e);
    if (!std::isfinite(p)) return nonfinite(p);
    e += lo_ * b;
    sum_ordered(p, e, hi_, lo_);
    return *this;
  }
  static DoubleDouble product(double a, double b) {
    DoubleDouble r(a);
    r *= b;
    return r;
  }

 private:
  static void sum_exact(double a, double b, double& s, double& e) {
    s = a + b;
    const double bv = s - a;
    e = (a - (s - bv)) + (b - bv);
  }
  static void sum_ordered(double a, double b, double& s, double& e) {
    s = a + b;
    e = b - (s - a);
  }
  static void halves(double a, double& h, double& l) {
    constexpr double kSplit = 134217729.0;  // 2^27 + 1
    constexpr double kBig = 6.69692879491417e+299;  // 2^996
    if (a > kBig || a < -kBig) {
      a *= 0x1p-28;
      const double t = kSplit * a;
      h = (t - (t - a)) * 0x1p28;
      l = (a - (t - (t - a))) * 0x1p28;
      return;
    }
    const double t = kSplit * a;
    h = t - (t - a);
    l = a - h;
  }
  static void product_exact(double a, double b, double& p, double& e) {
    p = a * b;
    if (!std::isfinite(p)) {
      e = 0.0;
      return;
    }
    double ah, al, bh, bl;
    halves(a, ah, al);
    halves(b, bh, bl);
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
  }
  DoubleDouble& nonfinite(double v) {
    hi_ = v;
    lo_ = 0.0;
    return *this;
  }
  double hi_ = 0.0;
  double lo_ = 0.0;
};

using DD = DoubleDouble;

void record(VerifyReport& r, double violation, double tol, const char* what, int index) {
  if (violation <= tol) return;
  if (r.ok || r.message.empty()) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s %d violated by %.3e (tolerance %.1e)", what, index, violation, tol);
    r.message = buf;
  }
  r.ok = false;
}

double bound_violation(double v, double lo, double up, double scale) {
  if (v < lo) return (lo - v) / (1.0 + std::max(std::fabs(lo), scale));
  if (v > up) return (v - up) / (1.0 + std::max(std::fabs(up), scale));
  return 0.0;
}

VerifyReport check_primal(const Model& md, const double* x, double tol) {
  VerifyReport r;
  for (int j = 0; j < md.n; ++j) {
    if (!std::isfinite(x[j])) {
      r.ok = false;
      r.message = "solution contains a non-finite value";
      return r;
    }
    const double v = bound_violation(x[j], md.col_lower[j], md.col_upper[j], 0.0);
    r.max_bound_violation = std::max(r.max_bound_violation, v);
    record(r, v, tol, "column bound", j);
  }
  std::vector<DD> act(static_cast<size_t>(md.m));
  std::vector<double> mag(static_cast<size_t>(md.m), 0.0);
  for (int j = 0; j < md.n; ++j)
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const DD t = DD::product(md.A.value[p], x[j]);
      act[md.A.index[p]] += t;
      mag[md.A.index[p]] = std::max(mag[md.A.index[p]], t.magnitude().value());
    }
  for (int i = 0; i < md.m; ++i) {
    const double a = act[i].value();
    if (!std::isfinite(a) || !std::isfinite(mag[i])) {
      r.ok = false;
      r.max_row_violation = kInf;
      if (r.message.empty()) r.message = "row " + std::to_string(i) + " activity is not finite";
      continue;
    }
    const double v = bound_violation(a, md.row_lower[i], md.row_upper[i], mag[i]);
    r.max_row_violation = std::max(r.max_row_violation, v);
    record(r, v, tol, "row", i);
  }
  DD obj(md.offset);
  for (int j = 0; j < md.n; ++j) obj += DD::product(md.obj[j], x[j]);
  r.objective = obj.value();
  return r;
}

}  // namespace

VerifyReport verify_lp_optimality(const Model& md, const double* x, const double* y,
                                  double primal_tol, double dual_tol) {
  VerifyReport r = check_primal(md, x, primal_tol);
  const double sense = md.maximize ? -1.0 : 1.0;
  const auto at_bound = [&](double v, double b, double scale) {
    return std::isfinite(b) && std::fabs(v - b) <= primal_tol * (1.0 + std::max(std::fabs(b), scale));
  };
  std::vector<DD> act(static_cast<size_t>(md.m));
  std::vector<double> mag(static_cast<size_t>(md.m), 0.0);
  for (int j = 0; j < md.n; ++j) {
    DD d(md.obj[j]);
    double scale = std::fabs(md.obj[j]);
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const DD t = DD::product(md.A.value[p], y[md.A.index[p]]);
      d += t.negated();
      scale += t.magnitude().value();
      act[md.A.index[p]] += DD::product(md.A.value[p], x[j]);
      mag[md.A.index[p]] = std::max(mag[md.A.index[p]], std::fabs(md.A.value[p] * x[j]));
    }
    const double dj = sense * d.value();
    if (!std::isfinite(dj) || !std::isfinite(scale)) {
      r.ok = false;
      r.max_dual_violation = kInf;
      if (r.message.empty()) r.message = "reduced cost of column " + std::to_string(j) + " is not finite";
      continue;
    }
    double v = 0.0;
    if (!at_bound(x[j], md.col_lower[j], 0.0)) v = std::max(v, dj);
    if (!at_bound(x[j], md.col_upper[j], 0.0)) v = std::max(v, -dj);
    v /= 1.0 + scale;
    r.max_dual_violation = std::max(r.max_dual_violation, v);
    record(r, v, dual_tol, "reduced cost of column", j);
  }
  for (int i = 0; i < md.m; ++i) {
    if (!std::isfinite(y[i])) {
      r.ok = false;
      r.message = "dual vector contains a non-finite value";
      return r;
    }
    const double yi = sense * y[i];
    const double a = act[i].value();
    double v = 0.0;
    if (!at_bound(a, md.row_lower[i], mag[i])) v = std::max(v, yi);
    if (!at_bound(a, md.row_upper[i], mag[i])) v = std::max(v, -yi);
    v /= 1.0 + std::fabs(yi);
    r.max_dual_violation = std::max(r.max_dual_violation, v);
    record(r, v, dual_tol, "dual of row", i);
  }
  return r;
}

VerifyReport verify_infeasibility(const Model& md, const double* y, double primal_tol) {
  VerifyReport r;
  double y_max = 0.0;
  for (int i = 0; i < md.m; ++i) {
    if (!std::isfinite(y[i])) {
      r.ok = false;
      r.message = "certificate contains a non-finite value";
      return r;
    }
    y_max = std::max(y_max, std::fabs(y[i]));
  }
  if (y_max == 0.0) {
    r.ok = false;
    r.message = "certificate is zero";
    return r;
  }
  DD lo(0.0), hi(0.0);
  double scale = 0.0;
  const auto range = [&](const DD& c, double l, double u) {
    if (DD(0.0).less(c)) {
      lo += l > -kInf ? DD(c) *= l : DD(-INFINITY);
      hi += u < kInf ? DD(c) *= u : DD(INFINITY);
    } else if (c.less(DD(0.0))) {
      lo += u < kInf ? DD(c) *= u : DD(-INFINITY);
      hi += l > -kInf ? DD(c) *= l : DD(INFINITY);
    }
  };
  for (int j = 0; j < md.n; ++j) {
    DD w(0.0);
    double w_scale = 0.0;
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const DD t = DD::product(md.A.value[p], y[md.A.index[p]] / y_max);
      w += t;
      w_scale += t.magnitude().value();
    }
    if (std::fabs(w.value()) <= 1e-9 * w_scale) w = DD(0.0);
    range(w, md.col_lower[j], md.col_upper[j]);
    if (!w.is_zero())
      scale = std::max({scale, std::fabs(w.value() * md.col_lower[j]), std::fabs(w.value() * md.col_upper[j])});
  }
  for (int i = 0; i < md.m; ++i) {
    const DD c = DD(y[i] / y_max).negated();
    range(c, md.row_lower[i], md.row_upper[i]);
    if (!c.is_zero())
      scale = std::max({scale, std::fabs(c.value() * md.row_lower[i]), std::fabs(c.value() * md.row_upper[i])});
  }
  if (!std::isfinite(scale)) scale = 0.0;
  const double margin = primal_tol * (1.0 + scale);
  if (DD(margin).less(lo) || hi.less(DD(-margin))) return r;
  r.ok = false;
  char buf[160];
  std::snprintf(buf, sizeof buf, "certificate range [%.3e, %.3e] does not exclude zero", lo.value(), hi.value());
  r.message = buf;
  return r;
}

VerifyReport verify_unbounded_ray(const Model& md, const double* ray, double primal_tol,
                                  double dual_tol) {
  VerifyReport r;
  double v_max = 0.0;
  for (int j = 0; j < md.n; ++j) v_max = std::max(v_max, std::fabs(ray[j]));
  if (!(v_max > 0.0) || !std::isfinite(v_max)) {
    r.ok = false;
    r.message = "ray is zero or not finite";
    return r;
  }
  const double sense = md.maximize ? -1.0 : 1.0;
  DD slope(0.0);
  double c_max = 0.0;
  for (int j = 0; j < md.n; ++j) {
    const double v = ray[j] / v_max;
    slope += DD::product(sense * md.obj[j], v);
    c_max = std::max(c_max, std::fabs(md.obj[j]));
    double viol = 0.0;
    if (v < 0.0 && md.col_lower[j] > -kInf) viol = -v;
    if (v > 0.0 && md.col_upper[j] < kInf) viol = v;
    r.max_bound_violation = std::max(r.max_bound_violation, viol);
    record(r, viol, primal_tol, "ray leaves the bounds of column", j);
  }
  std::vector<DD> av(static_cast<size_t>(md.m));
  std::vector<double> mag(static_cast<size_t>(md.m), 0.0);
  for (int j = 0; j < md.n; ++j)
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const double t = md.A.value[p] * ray[j] / v_max;
      av[md.A.index[p]] += DD(t);
      mag[md.A.index[p]] = std::max(mag[md.A.index[p]], std::fabs(t));
    }
  for (int i = 0; i < md.m; ++i) {
    const double a = av[i].value();
    double viol = 0.0;
    if (a < 0.0 && md.row_lower[i] > -kInf) viol = -a / (1.0 + mag[i]);
    if (a > 0.0 && md.row_upper[i] < kInf) viol = a / (1.0 + mag[i]);
    r.max_row_violation = std::max(r.max_row_violation, viol);
    record(r, viol, primal_tol, "ray leaves the bounds of row", i);
  }
  if (!slope.less(DD(-dual_tol * (1.0 + c_max)))) {
    r.ok = false;
    r.message = "ray does not improve the objective";
  }
  return r;
}

double safe_dual_bound(const Model& md, const double* y) {
  // src/verify/verify.cpp:132-247 conventions: d_j = sense (c_j - a_j'y), a row's reduced cost is
  // sense y_i, and weak duality holds for the minimization of sense c'x.
  const double sense = md.maximize ? -1.0 : 1.0;
  std::vector<double> ys(static_cast<size_t>(md.m));
  for (int i = 0; i < md.m; ++i) {
    double yi = sense * y[i];
    if (!std::isfinite(yi)) yi = 0.0;
    if (md.row_lower[i] == -kInf && yi > 0.0) yi = 0.0;  // only an upper bound: dual <= 0
    if (md.row_upper[i] == kInf && yi < 0.0) yi = 0.0;   // only a lower bound: dual >= 0
    ys[i] = yi;
  }
  const auto term = [](DD g, double lo, double up, DD& acc, double& mag) -> bool {
    const double gv = g.value();
    if (gv == 0.0) return true;
    const double b = gv > 0.0 ? lo : up;
    if (!std::isfinite(b)) return false;
    g *= b;
    acc += g;
    mag += std::fabs(g.value());
    return true;
  };
  DD bound;
  double mag = 0.0;
  bool finite = true;
  for (int j = 0; j < md.n && finite; ++j) {
    DD d(sense * md.obj[j]);
    double dmag = std::fabs(md.obj[j]);
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) {
      const DD t = DD::product(md.A.value[p], ys[md.A.index[p]]);
      d += t.negated();
      dmag += std::fabs(t.value());
    }
    mag += dmag * std::max(std::fabs(md.col_lower[j]) < kInf ? std::fabs(md.col_lower[j]) : 0.0,
                           std::fabs(md.col_upper[j]) < kInf ? std::fabs(md.col_upper[j]) : 0.0);
    finite = term(d, md.col_lower[j], md.col_upper[j], bound, mag);
  }
  for (int i = 0; i < md.m && finite; ++i)
    finite = term(DD(ys[i]), md.row_lower[i], md.row_upper[i], bound, mag);
  if (!finite) return md.maximize ? kInf : -kInf;
  // Every double-double operation errs by at most ~2^-104 of its magnitudes; one margin of
  // (nnz + n + m + 8) 2^-100 times their total covers them all.
  const double margin =
      static_cast<double>(md.A.nnz() + md.n + md.m + 8) * 0x1p-100 * (mag + std::fabs(bound.value()));
  const double lower = bound.value() - margin;  // minimization of sense c'x
  return sense * lower + md.offset;
}

namespace {

// Whether ColumnOrder::kAuto reorders the model's columns: the enhanced algorithm, and the longest
// column at least 16 times the mean.
bool skewed_for_auto(const Model& md, const PdlpOptions& o) {
  if (o.algorithm != Algorithm::kEnhanced || md.n == 0) return false;
  int64_t longest = 0;
  for (int j = 0; j < md.n; ++j) longest = std::max(longest, md.A.start[j + 1] - md.A.start[j]);
  const double mean = static_cast<double>(md.A.nnz()) / md.n;
  return static_cast<double>(longest) >= 16.0 * std::max(mean, 1.0);
}

// The columns by decreasing count, ties in the model's order: order[k] is the model column that
// becomes column k.
std::vector<int> degree_order(const Model& md) {
  std::vector<int> order(static_cast<size_t>(md.n));
  for (int j = 0; j < md.n; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return md.A.start[a + 1] - md.A.start[a] > md.A.start[b + 1] - md.A.start[b];
  });
  return order;
}

// ColumnOrder::kTouch: the hot columns (count at least twice the mean) by decreasing count, then
// the others by the first row that holds them (the order in which a pass over the rows first
// reaches them), ties in the model's order. A row's columns seen for the first time then sit side
// by side, so its gathers of them share cache lines.
std::vector<int> touch_order(const Model& md) {
  const double mean = md.n > 0 ? static_cast<double>(md.A.nnz()) / md.n : 0.0;
  const auto count = [&](int j) { return md.A.start[j + 1] - md.A.start[j]; };
  std::vector<int> hot, cold;
  std::vector<int> first(static_cast<size_t>(md.n), md.m);
  for (int j = 0; j < md.n; ++j) {
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) first[j] = std::min(first[j], md.A.index[p]);
    (static_cast<double>(count(j)) >= 2.0 * mean ? hot : cold).push_back(j);
  }
  std::stable_sort(hot.begin(), hot.end(), [&](int a, int b) { return count(a) > count(b); });
  std::stable_sort(cold.begin(), cold.end(), [&](int a, int b) { return first[a] < first[b]; });
  hot.insert(hot.end(), cold.begin(), cold.end());
  return hot;
}

// The column order of a model-level solve (empty: the model's own). kAuto takes the touch order
// on skewed column counts.
std::vector<int> solve_order(const Model& md, const PdlpOptions& o) {
  if (md.n == 0 || o.column_order == ColumnOrder::kNatural) return {};
  if (o.column_order == ColumnOrder::kDegree) return degree_order(md);
  if (o.column_order == ColumnOrder::kTouch || skewed_for_auto(md, o)) return touch_order(md);
  return {};
}

template <typename T>
std::vector<T> gather(const std::vector<T>& v, const std::vector<int>& order) {
  std::vector<T> out(order.size());
  for (size_t k = 0; k < order.size(); ++k) out[k] = v[order[k]];
  return out;
}

// The model with its columns in `order` (rows and their order unchanged).
std::shared_ptr<const Model> permute_columns(const Model& md, const std::vector<int>& order) {
  auto out = std::make_shared<Model>();
  out->m = md.m;
  out->n = md.n;
  out->maximize = md.maximize;
  out->offset = md.offset;
  out->A.rows = md.A.rows;
  out->A.cols = md.A.cols;
  out->A.start.assign(static_cast<size_t>(md.n) + 1, 0);
  out->A.index.resize(md.A.index.size());
  out->A.value.resize(md.A.value.size());
  int64_t z = 0;
  for (int k = 0; k < md.n; ++k) {
    const int j = order[k];
    const int64_t b = md.A.start[j], e = md.A.start[j + 1];
    std::copy(md.A.index.begin() + b, md.A.index.begin() + e, out->A.index.begin() + z);
    std::copy(md.A.value.begin() + b, md.A.value.begin() + e, out->A.value.begin() + z);
    z += e - b;
    out->A.start[k + 1] = z;
  }
  out->obj = gather(md.obj, order);
  out->col_lower = gather(md.col_lower, order);
  out->col_upper = gather(md.col_upper, order);
  out->row_lower = md.row_lower;
  out->row_upper = md.row_upper;
  return out;
}

// Column vectors of a permuted solve back in the model's order.
void unpermute(std::vector<double>& v, const std::vector<int>& order) {
  if (v.size() != order.size()) return;
  std::vector<double> out(v.size());
  for (size_t k = 0; k < order.size(); ++k) out[order[k]] = v[k];
  v = std::move(out);
}

// One model of solve_lp / solve_lp_batch between its PDLP rounds.
struct LpJob {
  const Model* model = nullptr;
  // A column-permuted copy of the model the job solves instead, and the permutation (order[k]:
  // the model column at position k); empty order: the model itself.
  std::shared_ptr<const Model> permuted;
  std::vector<int> order;
  std::vector<double> own_lower, own_upper;  // node bounds in the permuted order
  std::shared_ptr<const Scaling> scaled;
  std::shared_ptr<const LpProblem> lp;
  // Node relaxations: this job's column bounds (model space) and their computational-form
  // bounds (n + m); empty: the model's own.
  const double* col_lower = nullptr;
  const double* col_upper = nullptr;
  std::vector<double> lower, upper;
  LpSolveResult out;
  std::vector<double> start_x, start_y;  // the point a round starts from (warm or polish)
  bool warm = false;
  bool open = true;                      // another round is wanted
};

// The model a job's answer is verified against: the model itself, or a copy with the job's
// column bounds.
std::unique_ptr<Model> job_model(const LpJob& job) {
  if (job.col_lower == nullptr) return nullptr;
  auto md = std::make_unique<Model>(*job.model);
  md->col_lower.assign(job.col_lower, job.col_lower + md->n);
  md->col_upper.assign(job.col_upper, job.col_upper + md->n);
  return md;
}

// Maps one PDLP result back to the model (solve_pdlp_lp, src/core/solver.cpp:392-477: unscale;
// y <- -sense y since PDLP's reduced costs are c + A'y; d = c - A'y; objective in double-double),
// runs the verifier and decides whether a polish round follows.
void absorb(LpJob& job, PdlpResult&& result, int round, const LpSolveOptions& options) {
  const std::unique_ptr<Model> own = job_model(job);
  const Model& model = own ? *own : *job.model;
  LpSolveResult& out = job.out;
  const Scaling& scaled = *job.scaled;
  const int n = model.n, m = model.m;
  const double sense = model.maximize ? -1.0 : 1.0;
  out.pdlp = std::move(result);
  const PdlpResult& r = out.pdlp;
  out.status = r.status;
  out.iterations += r.iterations;
  out.fp32_iterations += r.fp32_iterations;
  out.rounds = round + 1;
  out.message = r.message;
  if (r.x.size() == static_cast<size_t>(n + m)) {
    out.col_value.assign(r.x.begin(), r.x.begin() + n);
    out.row_activity.assign(r.x.begin() + n, r.x.end());
    for (int j = 0; j < n; ++j) out.col_value[j] *= scaled.col[j];
    for (int i = 0; i < m; ++i) out.row_activity[i] /= scaled.row[i];
    out.row_dual = r.y;
    for (int i = 0; i < m; ++i) out.row_dual[i] *= scaled.row[i];
    for (double& yv : out.row_dual) yv *= -sense;
    out.col_dual.assign(static_cast<size_t>(n), 0.0);
    DD obj(model.offset);
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int64_t p = model.A.start[j]; p < model.A.start[j + 1]; ++p)
        s += model.A.value[p] * out.row_dual[model.A.index[p]];
      out.col_dual[j] = model.obj[j] - s;
      obj += DD::product(model.obj[j], out.col_value[j]);
    }
    out.objective = obj.value();
  }
  VerifyReport rep;
  bool answered = true;
  out.dual_bound = std::numeric_limits<double>::quiet_NaN();
  if (r.status == Status::kOptimal) {
    if (options.verify) rep = verify_lp_optimality(model, out.col_value.data(), out.row_dual.data());
    out.dual_bound = safe_dual_bound(model, out.row_dual.data());
  } else if (r.status == Status::kInfeasible && !r.dual_ray.empty()) {
    out.infeasibility_certificate = r.dual_ray;
    for (int i = 0; i < m; ++i) out.infeasibility_certificate[i] *= scaled.row[i];
    for (double& t : out.infeasibility_certificate) t *= sense;
    if (options.verify) {
      rep = verify_infeasibility(model, out.infeasibility_certificate.data());
      if (!rep.ok) {
        // A Farkas certificate must be sign-feasible: the dual of a row bounded on one side only
        // has that side's sign, a free row's is 0. The raw PDLP dual (the reference returns it as
        // is) may carry small entries of the wrong sign, each of which makes the certificate's
        // range unbounded, and small entries on rows that play no part in the proof, each of
        // which widens the range. Its sign-feasible part, in either orientation, with the entries
        // below a fraction of its largest dropped (none, then 1e-6, 1e-3, 1e-1), is checked next;
        // the first that the verifier accepts is kept.
        bool done = false;
        for (const double orient : {1.0, -1.0}) {
          std::vector<double> base = out.infeasibility_certificate;
          double y_max = 0.0;
          for (int i = 0; i < m; ++i) {
            double t = orient * base[i];
            const bool lo = std::isfinite(model.row_lower[i]), up = std::isfinite(model.row_upper[i]);
            if (!lo && !up) t = 0.0;
            else if (!lo) t = std::max(t, 0.0);
            else if (!up) t = std::min(t, 0.0);
            base[i] = t;
            y_max = std::max(y_max, std::fabs(t));
          }
          for (const double drop : {0.0, 1e-6, 1e-3, 1e-1}) {
            std::vector<double> y = base;
            for (double& t : y)
              if (std::fabs(t) < drop * y_max) t = 0.0;
            const VerifyReport cleaned = verify_infeasibility(model, y.data());
            if (cleaned.ok) {
              out.infeasibility_certificate = std::move(y);
              rep = cleaned;
              done = true;
              break;
            }
          }
          if (done) break;
        }
      }
    }
  } else if (r.status == Status::kUnbounded && !r.primal_ray.empty()) {
    out.unbounded_ray = r.primal_ray;
    for (int j = 0; j < n; ++j) out.unbounded_ray[j] *= scaled.col[j];
    if (options.verify) rep = verify_unbounded_ray(model, out.unbounded_ray.data());
  } else {
    answered = false;
  }
  out.verified = answered && options.verify && rep.ok;
  out.max_primal_violation = std::max(rep.max_bound_violation, rep.max_row_violation);
  out.max_dual_violation = rep.max_dual_violation;
  if (answered && !rep.ok)
    out.message = "PDLP: " + std::string(to_string(r.status)) + " failed verification (" + rep.message + ")";
  job.open = !(out.verified || !answered || r.status != Status::kOptimal || round >= options.polish_rounds);
  if (job.open) {
    // Polish: the same solve from the rejected point, in FP64, with a tighter tolerance.
    job.start_x = r.x;
    job.start_y = r.y;
  }
}

template <typename F>
void for_each_parallel(int count, F&& fn) {
  const int threads = std::min<int>(count, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  if (threads <= 1) {
    for (int i = 0; i < count; ++i) fn(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (int i = next++; i < count; i = next++) fn(i);
    });
  for (auto& th : pool) th.join();
}


// The PDLP rounds of a set of jobs: every open job in one batch per round, then mapping,
// verification and the polish decision per job.
void run_jobs(std::vector<LpJob>& jobs, const LpSolveOptions& options, PdlpBatchInfo* info) {
  const int K = static_cast<int>(jobs.size());
  PdlpOptions popt = options.pdlp;
  for (int round = 0;; ++round) {
    std::vector<int> open;
    for (int i = 0; i < K; ++i)
      if (jobs[i].open) open.push_back(i);
    if (open.empty()) break;
    std::vector<PdlpTask> tasks;
    for (int i : open) {
      PdlpTask t{jobs[i].lp.get(), nullptr, nullptr, nullptr, nullptr};
      if (round > 0 || jobs[i].warm) {
        t.start_x = &jobs[i].start_x;
        t.start_y = &jobs[i].start_y;
      }
      if (!jobs[i].lower.empty()) {
        t.lower = &jobs[i].lower;
        t.upper = &jobs[i].upper;
      }
      tasks.push_back(t);
    }
    PdlpBatchInfo round_info;
    std::vector<PdlpResult> results = solve_pdlp_batch(tasks, popt, &round_info);
    if (info != nullptr) {
      info->setup_seconds += round_info.setup_seconds;
      info->solve_seconds += round_info.solve_seconds;
      info->kernel_ms += round_info.kernel_ms;
    }
    for_each_parallel(static_cast<int>(open.size()),
                      [&](int q) { absorb(jobs[open[q]], std::move(results[q]), round, options); });
    // Polish rounds: FP64 PDHG steps with averages and constant step sizes (the reference's
    // iteration) from the rejected point, to a tenth of the tolerance. Halpern steps move slowly
    // below ~1e-6 on some LPs (the 1M-column LP took 508k FP64 steps from 1e-6 to 1e-7).
    popt.tol /= 10.0;
    popt.precision = Precision::kFp64;
    popt.halpern = false;
    popt.adaptive_steps = false;
  }
}

std::vector<LpSolveResult> collect(std::vector<LpJob>& jobs) {
  std::vector<LpSolveResult> out;
  out.reserve(jobs.size());
  for (LpJob& j : jobs) {
    if (!j.order.empty()) {
      unpermute(j.out.col_value, j.order);
      unpermute(j.out.col_dual, j.order);
      unpermute(j.out.unbounded_ray, j.order);
    }
    out.push_back(std::move(j.out));
  }
  return out;
}

}  // namespace

std::vector<LpSolveResult> solve_lp_batch(const std::vector<const Model*>& models,
                                          const LpSolveOptions& options, PdlpBatchInfo* info) {
  const int K = static_cast<int>(models.size());
  std::vector<LpJob> jobs(static_cast<size_t>(K));
  for_each_parallel(K, [&](int i) {
    jobs[i].model = models[i];
    jobs[i].order = solve_order(*models[i], options.pdlp);
    if (!jobs[i].order.empty()) {
      jobs[i].permuted = permute_columns(*models[i], jobs[i].order);
      jobs[i].model = jobs[i].permuted.get();
    }
    auto sc = std::make_shared<Scaling>(compute_scaling(jobs[i].model->A));
    jobs[i].lp = std::make_shared<LpProblem>(make_problem(*jobs[i].model, *sc));
    jobs[i].scaled = std::move(sc);
  });
  run_jobs(jobs, options, info);
  return collect(jobs);
}

std::vector<LpSolveResult> solve_lp_bounds_batch(const Model& model,
                                                 const std::vector<const double*>& col_lower,
                                                 const std::vector<const double*>& col_upper,
                                                 const LpSolveOptions& options,
                                                 const double* start_col_value,
                                                 const double* start_row_dual,
                                                 PdlpBatchInfo* info) {
  const int K = static_cast<int>(col_lower.size());
  const int n = model.n, m = model.m;
  // A column order for the whole batch (the model's, or by decreasing count): the model, every
  // node's bounds and the warm start follow it.
  std::vector<int> order;
  std::shared_ptr<const Model> permuted;
  std::vector<double> start_perm;
  order = solve_order(model, options.pdlp);
  if (!order.empty()) {
    permuted = permute_columns(model, order);
    if (start_col_value != nullptr) {
      start_perm.resize(static_cast<size_t>(n));
      for (int k = 0; k < n; ++k) start_perm[k] = start_col_value[order[k]];
      start_col_value = start_perm.data();
    }
  }
  const Model& md = permuted ? *permuted : model;
  auto scaled = std::make_shared<const Scaling>(compute_scaling(md.A));
  auto lp = std::make_shared<const LpProblem>(make_problem(md, *scaled));
  // Warm start in the computational form (make_problem's space): x / col scale, the row
  // activities A x times the row scale, and y = -sense row dual / row scale.
  std::vector<double> sx, sy;
  if (start_col_value != nullptr) {
    const double sense = model.maximize ? -1.0 : 1.0;
    sx.assign(static_cast<size_t>(n + m), 0.0);
    sy.assign(static_cast<size_t>(m), 0.0);
    std::vector<double> act(static_cast<size_t>(m), 0.0);
    for (int j = 0; j < n; ++j)
      for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p)
        act[md.A.index[p]] += md.A.value[p] * start_col_value[j];
    for (int j = 0; j < n; ++j) sx[j] = start_col_value[j] / scaled->col[j];
    for (int i = 0; i < m; ++i) sx[n + i] = act[i] * scaled->row[i];
    if (start_row_dual != nullptr)
      for (int i = 0; i < m; ++i) sy[i] = -sense * start_row_dual[i] / scaled->row[i];
  }
  std::vector<LpJob> jobs(static_cast<size_t>(K));
  for_each_parallel(K, [&](int k) {
    LpJob& job = jobs[k];
    job.model = &md;
    job.scaled = scaled;
    job.lp = lp;
    job.col_lower = col_lower[k];
    job.col_upper = col_upper[k];
    if (!order.empty()) {
      job.order = order;
      job.own_lower.resize(static_cast<size_t>(n));
      job.own_upper.resize(static_cast<size_t>(n));
      for (int q = 0; q < n; ++q) {
        job.own_lower[q] = col_lower[k][order[q]];
        job.own_upper[q] = col_upper[k][order[q]];
      }
      job.col_lower = job.own_lower.data();
      job.col_upper = job.own_upper.data();
    }
    job.lower = lp->lower;
    job.upper = lp->upper;
    for (int j = 0; j < n; ++j) {
      job.lower[j] = job.col_lower[j] / scaled->col[j];
      job.upper[j] = job.col_upper[j] / scaled->col[j];
    }
    if (!sx.empty()) {
      job.start_x = sx;
      job.start_y = sy;
      job.warm = true;
    }
  });
  run_jobs(jobs, options, info);
  return collect(jobs);
}

LpSolveResult solve_lp(const Model& model, const LpSolveOptions& options) {
  return solve_lp_batch({&model}, options)[0];
}

}  // namespace smy::host
