#include "pdlp.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>

#include "../../../include/samaya_pdhg.h"

namespace smy::host {

const char* to_string(Status s) {
  switch (s) {
    case Status::kOptimal: return "optimal";
    case Status::kInfeasible: return "infeasible";
    case Status::kUnbounded: return "unbounded";
    case Status::kIterationLimit: return "iteration limit";
    case Status::kTimeLimit: return "time limit";
    case Status::kNumericalError: return "numerical error";
  }
  return "unknown";
}

namespace {

// Evaluations every this many iterations (src/lp/pdlp.cpp:25), detection every 8th of them after
// iteration 1000 (:29, :718-719), primal-weight distance floor (:31).
constexpr long long kCheckFrequency = 64;
constexpr long long kDetectionEvery = 8;
constexpr double kWeightDistanceFloor = 1e-10;

class Clock {
 public:
  double seconds() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
  }

 private:
  std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();
};

void logf(const PdlpOptions& o, int level, const char* fmt, ...) {
  if (level > o.log_level) return;
  va_list args;
  va_start(args, fmt);
  std::vfprintf(stdout, fmt, args);
  va_end(args);
  std::fputc('\n', stdout);
  std::fflush(stdout);
}

void check(int rc) {
  if (rc != SMY_OK) throw std::runtime_error(smy_last_error());
}

// Owner of one smy_pdhg handle; every failure becomes an exception (the reference's
// SAMAYA_CUDA_CHECK behaviour, src/gpu/device.hpp:83-91).
class Device {
 public:
  // Null when there is no usable GPU or the matrix is too large (PdhgDevice::create's null).
  static std::unique_ptr<Device> create(const Csc& A, const Csc& At, const std::vector<double>& cost,
                                        const std::vector<double>& lower,
                                        const std::vector<double>& upper, int precision) {
    const smy_csc a{A.rows, A.cols, A.start.data(), A.index.data(), A.value.data()};
    const smy_csc at{At.rows, At.cols, At.start.data(), At.index.data(), At.value.data()};
    smy_pdhg* h = nullptr;
    const int rc = smy_pdhg_create(&a, &at, cost.data(), lower.data(), upper.data(), precision, &h);
    if (rc == SMY_NO_DEVICE || rc == SMY_TOO_LARGE) return nullptr;
    check(rc);
    return std::unique_ptr<Device>(new Device(h));
  }
  ~Device() { smy_pdhg_destroy(h_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;

  smy_pdhg* handle() { return h_; }
  void set_original(const std::vector<double>& cs, const std::vector<double>& rs,
                    const std::vector<double>& cost, const std::vector<double>& lower,
                    const std::vector<double>& upper) {
    check(smy_pdhg_set_original(h_, cs.data(), rs.data(), cost.data(), lower.data(), upper.data()));
  }
  void run(double tau, double sigma, double theta, long long k0, int count, bool graph) {
    if (graph) {
      check(smy_pdhg_run(h_, tau, sigma, theta, k0, count));
    } else {
      for (int j = 0; j < count; ++j) {
        check(smy_pdhg_step(h_, tau, sigma, theta, k0 + j));
        if (j + 1 < count) check(smy_pdhg_advance(h_));
      }
    }
  }
  void evaluate_both(double out[14]) { check(smy_pdhg_evaluate_both(h_, out)); }
  void restart_distance(int which, double& dv, double& dy) {
    check(smy_pdhg_restart_distance(h_, which, &dv, &dy));
  }
  void restart_to(int which) { check(smy_pdhg_restart_to(h_, which)); }
  void advance() { check(smy_pdhg_advance(h_)); }
  void download(std::vector<double>* v, std::vector<double>* v_new, std::vector<double>* y,
                std::vector<double>* y_new, std::vector<double>* v_erg, std::vector<double>* y_erg,
                int nt, int m) {
    for (auto* p : {v, v_new, v_erg})
      if (p) p->resize(nt);
    for (auto* p : {y, y_new, y_erg})
      if (p) p->resize(m);
    const auto d = [](std::vector<double>* p) { return p ? p->data() : nullptr; };
    check(smy_pdhg_download(h_, d(v), d(v_new), d(y), d(y_new), d(v_erg), d(y_erg)));
  }
  void set_precision(int p) { check(smy_pdhg_set_precision(h_, p)); }
  void set_start(const std::vector<double>& v, const std::vector<double>& y) {
    check(smy_pdhg_set_start(h_, v.data(), y.data()));
  }
  double last_run_ms() {
    double ms = 0.0;
    return smy_pdhg_last_run_ms(h_, &ms) == SMY_OK ? ms : 0.0;
  }

 private:
  explicit Device(smy_pdhg* h) : h_(h) {}
  smy_pdhg* h_;
};

// The equilibration of solve_pdlp (src/lp/pdlp.cpp:109-187): up to `passes` simultaneous updates
// of every column and row by the inverse of its current 2-norm, clamped to [1e-8, 1e8]; stops when
// no factor moved by more than 10% (relative to max(|factor|, 1)).
void equilibrate(const Csc& A, const Csc& At, std::vector<double>& rs, std::vector<double>& cs,
                 int passes) {
  const int m = A.rows, n = A.cols;
  rs.assign(static_cast<size_t>(m), 1.0);
  cs.assign(static_cast<size_t>(n), 1.0);
  if (m == 0 || n == 0) return;
  std::vector<double> row_norm(static_cast<size_t>(m)), col_norm(static_cast<size_t>(n));
  const auto update = [](std::vector<double>& f, const std::vector<double>& norm, bool& moved) {
    for (size_t k = 0; k < f.size(); ++k) {
      if (!(norm[k] > 0)) continue;
      const double next = std::max(1e-8, std::min(1e8, f[k] / norm[k]));
      if (std::fabs(next - f[k]) / std::max(std::fabs(f[k]), 1.0) > 0.1) moved = true;
      f[k] = next;
    }
  };
  for (int pass = 0; pass < passes; ++pass) {
    for (int j = 0; j < n; ++j) {
      double ss = 0.0;
      for (int64_t p = A.start[j]; p < A.start[j + 1]; ++p) {
        const double t = A.value[p] * rs[A.index[p]] * cs[j];
        ss += t * t;
      }
      col_norm[j] = std::sqrt(ss);
    }
    for (int i = 0; i < m; ++i) {
      double ss = 0.0;
      for (int64_t p = At.start[i]; p < At.start[i + 1]; ++p) {
        const double t = At.value[p] * rs[i] * cs[At.index[p]];
        ss += t * t;
      }
      row_norm[i] = std::sqrt(ss);
    }
    bool moved = false;
    update(cs, col_norm, moved);
    update(rs, row_norm, moved);
    if (!moved) break;
  }
}

// Diagonal preconditioner of the enhanced algorithm: `passes` rounds of infinity-norm Ruiz
// equilibration (every row and column of D_r A D_c divided by the square root of its largest
// magnitude, rows and columns from the same current matrix), then one Pock-Chambolle step with
// alpha = 1 (rows by the square root of their 1-norm, columns likewise). Factors of empty rows and
// columns stay 1.
void ruiz_pock_chambolle(const Csc& A, const Csc& At, std::vector<double>& rs,
                         std::vector<double>& cs, int passes) {
  const int m = A.rows, n = A.cols;
  rs.assign(static_cast<size_t>(m), 1.0);
  cs.assign(static_cast<size_t>(n), 1.0);
  if (m == 0 || n == 0) return;
  std::vector<double> rn(static_cast<size_t>(m)), cn(static_cast<size_t>(n));
  const auto norms = [&](bool one_norm) {
    for (int j = 0; j < n; ++j) {
      double acc = 0.0;
      for (int64_t p = A.start[j]; p < A.start[j + 1]; ++p) {
        const double t = std::fabs(A.value[p] * rs[A.index[p]] * cs[j]);
        acc = one_norm ? acc + t : std::max(acc, t);
      }
      cn[j] = acc;
    }
    for (int i = 0; i < m; ++i) {
      double acc = 0.0;
      for (int64_t p = At.start[i]; p < At.start[i + 1]; ++p) {
        const double t = std::fabs(At.value[p] * rs[i] * cs[At.index[p]]);
        acc = one_norm ? acc + t : std::max(acc, t);
      }
      rn[i] = acc;
    }
    for (int j = 0; j < n; ++j)
      if (cn[j] > 0.0) cs[j] /= std::sqrt(cn[j]);
    for (int i = 0; i < m; ++i)
      if (rn[i] > 0.0) rs[i] /= std::sqrt(rn[i]);
  };
  for (int pass = 0; pass < passes; ++pass) norms(false);
  norms(true);
}

// Upper bound on ||[A -I]||_2 by 20 power iterations on A'A from a seeded random start
// (src/lp/pdlp.cpp:36-103): sqrt((1.01 sigma)^2 + 1).
double operator_norm_bound(const Csc& A, const Csc& At) {
  const int m = A.rows, n = A.cols;
  if (m == 0) return 1.0;
  std::vector<double> u(static_cast<size_t>(m)), w(static_cast<size_t>(n));
  std::mt19937_64 gen(12345);
  std::uniform_real_distribution<double> uniform(-1.0, 1.0);
  for (double& t : u) t = uniform(gen);
  const auto norm = [](const std::vector<double>& z) {
    double s = 0.0;
    for (double t : z) s += t * t;
    return std::sqrt(s);
  };
  double un = norm(u);
  for (double& t : u) t /= un;
  double sigma = 0.0;
  for (int it = 0; it < 20; ++it) {
    std::fill(w.begin(), w.end(), 0.0);
    for (int i = 0; i < m; ++i)
      for (int64_t p = At.start[i]; p < At.start[i + 1]; ++p) w[At.index[p]] += At.value[p] * u[i];
    const double wn = norm(w);
    if (wn == 0.0) return 1.0;
    std::fill(u.begin(), u.end(), 0.0);
    for (int j = 0; j < n; ++j) {
      if (w[j] == 0.0) continue;
      for (int64_t p = A.start[j]; p < A.start[j + 1]; ++p) u[A.index[p]] += A.value[p] * w[j];
    }
    un = norm(u);
    if (un == 0.0) return 1.0;
    sigma = un / wn;
    for (double& t : u) t /= un;
  }
  const double la = sigma * 1.01;
  return std::sqrt(la * la + 1.0);
}

struct Sums {
  double primal_sq, dual_sq, primal_obj, dual_obj, x_sq, r_sq, y_sq;
};

struct Measure {
  double rel_primal, rel_dual, rel_gap, gap, primal_obj, dual_obj;
};

double worst(const Measure& e) {
  return std::max({e.rel_primal, e.rel_dual, std::fabs(e.rel_gap)});
}

class Solve {
 public:
  Solve(const LpProblem& lp, const PdlpOptions& opt) : lp_(lp), opt_(opt) {}
  PdlpResult run();

 private:
  Measure finish(const Sums& s) const;
  double kkt(const Measure& e) const {
    return std::sqrt(omega_ * omega_ * e.rel_primal * e.rel_primal +
                     e.rel_dual * e.rel_dual / (omega_ * omega_) + e.rel_gap * e.rel_gap);
  }
  bool detect_farkas(long long iter, const std::vector<double>& y_erg);
  bool detect_ray(long long iter, const std::vector<double>& v, const std::vector<double>& v_new);
  void map_back(const std::vector<double>& v, const std::vector<double>& y);

  const LpProblem& lp_;
  const PdlpOptions& opt_;
  PdlpResult res_;
  Clock clock_;
  std::vector<double> rs_, cs_;
  double a_colmax_ = 0.0, at_rowmax_ = 0.0, c_norm_ = 0.0;
  double omega_ = 1.0;
};

Measure Solve::finish(const Sums& s) const {
  // src/lp/pdlp.cpp:493-519.
  const int m = lp_.m;
  const double primal_scale = 1.0 + a_colmax_ * std::sqrt(s.x_sq) + std::sqrt(s.r_sq);
  const double dual_scale = 1.0 + at_rowmax_ * std::sqrt(s.y_sq) + c_norm_;
  const double gap_scale = 1.0 + std::fabs(s.primal_obj) + std::fabs(s.dual_obj);
  const double gap = s.primal_obj - s.dual_obj;
  return Measure{m > 0 ? std::sqrt(s.primal_sq) / primal_scale : 0.0,
                 m > 0 ? std::sqrt(s.dual_sq) / dual_scale : 0.0, gap / gap_scale, gap,
                 s.primal_obj, s.dual_obj};
}

// Farkas certificate test on the average dual (src/lp/pdlp.cpp:724-842).
bool Solve::detect_farkas(long long iter, const std::vector<double>& y_erg) {
  const int m = lp_.m, n = lp_.n;
  std::vector<double> yf(static_cast<size_t>(m));
  double y_max = 0.0;
  for (int i = 0; i < m; ++i) {
    yf[i] = y_erg[i] * rs_[i];
    y_max = std::max(y_max, std::fabs(yf[i]));
  }
  if (!(y_max > 1e-6)) return false;
  std::vector<double> w(static_cast<size_t>(n + m), 0.0);
  for (int i = 0; i < m; ++i) {
    const double yi = yf[i];
    if (yi == 0.0) continue;
    for (int64_t p = lp_.At.start[i]; p < lp_.At.start[i + 1]; ++p) w[lp_.At.index[p]] += lp_.At.value[p] * yi;
  }
  for (int i = 0; i < m; ++i) w[n + i] = -yf[i];
  double w_sum = 0.0;
  for (int k = 0; k < n + m; ++k) w_sum += std::fabs(w[k]);
  const double drop = 1e-9 * w_sum;
  double f_min = 0.0, f_max = 0.0;
  for (int k = 0; k < n + m; ++k) {
    const double c = std::fabs(w[k]) <= drop ? 0.0 : w[k];
    const double lo = lp_.lower[k], up = lp_.upper[k];
    if (c > 0) {
      f_min += lo > -kInf ? c * lo : -INFINITY;
      f_max += up < kInf ? c * up : INFINITY;
    } else if (c < 0) {
      f_min += up < kInf ? c * up : -INFINITY;
      f_max += lo > -kInf ? c * lo : INFINITY;
    }
  }
  double scale = 0.0;
  for (int k = 0; k < n + m; ++k) {
    const double t = (k < n ? w[k] : -w[k]) / y_max;
    if (std::fabs(t) > 1e-9) {
      if (lp_.lower[k] > -kInf) scale = std::max(scale, std::fabs(t * lp_.lower[k]));
      if (lp_.upper[k] < kInf) scale = std::max(scale, std::fabs(t * lp_.upper[k]));
    }
  }
  if (!std::isfinite(scale)) scale = 0.0;
  const double margin = 1e-8 * (1.0 + scale);
  logf(opt_, 2, "pdlp: infeas check iter=%lld y_max=%.2e min_f=%.2e max_f=%.2e margin=%.2e scale=%.2e",
       iter, y_max, f_min, f_max, margin, scale);
  if ((std::isfinite(f_min) && f_min > margin) || (std::isfinite(f_max) && f_max < -margin)) {
    logf(opt_, 1, "pdlp: detected primal infeasibility at iter %lld (Farkas certificate)", iter);
    res_.dual_ray = y_erg;
    return true;
  }
  return false;
}

// Improving-ray test on the difference of consecutive iterates (src/lp/pdlp.cpp:852-1042).
bool Solve::detect_ray(long long iter, const std::vector<double>& v, const std::vector<double>& v_new) {
  const int m = lp_.m, n = lp_.n;
  std::vector<double> d(v_new.size());
  double d_max = 0.0;
  for (size_t k = 0; k < d.size(); ++k) {
    d[k] = v_new[k] - v[k];
    d_max = std::max(d_max, std::fabs(d[k]));
  }
  if (!(d_max > 1e-6)) return false;
  for (double& t : d) t /= d_max;
  std::vector<double> res(static_cast<size_t>(m), 0.0);
  for (int j = 0; j < n; ++j) {
    const double xj = d[j] * cs_[j];
    if (xj == 0.0) continue;
    for (int64_t p = lp_.A.start[j]; p < lp_.A.start[j + 1]; ++p) res[lp_.A.index[p]] += lp_.A.value[p] * xj;
  }
  for (int i = 0; i < m; ++i) res[i] -= d[n + i] / rs_[i];
  double primal_feas = 0.0;
  for (double t : res) primal_feas += t * t;
  primal_feas = std::sqrt(primal_feas);
  double cost_dir = 0.0;
  for (int j = 0; j < n; ++j) cost_dir += lp_.cost[j] * (d[j] * cs_[j]);
  for (int i = 0; i < m; ++i) cost_dir += lp_.cost[n + i] * (d[n + i] / rs_[i]);
  constexpr double kBoxTol = 1e-4, kFixedTol = 1e-3;
  const auto fixed = [](double lo, double up) { return up - lo <= 1e-10 * (1.0 + std::fabs(lo)); };
  bool ok = true;
  std::vector<double> ray(static_cast<size_t>(n), 0.0);
  for (int j = 0; j < n && ok; ++j) {
    double xj = d[j] * cs_[j];
    const double lo = lp_.lower[j], up = lp_.upper[j];
    const bool lb = lo > -kInf / 2, ub = up < kInf / 2;
    if (lb && ub) {
      if (std::fabs(xj) > (fixed(lo, up) ? kFixedTol : kBoxTol)) ok = false;
      xj = 0.0;
    } else if (lb) {
      if (xj < -kBoxTol) ok = false;
      xj = std::max(xj, 0.0);
    } else if (ub) {
      if (xj > kBoxTol) ok = false;
      xj = std::min(xj, 0.0);
    }
    ray[j] = xj;
  }
  std::vector<double> av(static_cast<size_t>(m), 0.0);
  for (int j = 0; j < n; ++j) {
    if (ray[j] == 0.0) continue;
    for (int64_t p = lp_.A.start[j]; p < lp_.A.start[j + 1]; ++p) av[lp_.A.index[p]] += lp_.A.value[p] * ray[j];
  }
  for (int i = 0; i < m && ok; ++i) {
    const double lo = lp_.lower[n + i], up = lp_.upper[n + i];
    const bool lb = lo > -kInf / 2, ub = up < kInf / 2;
    if (lb && ub) {
      if (std::fabs(av[i]) > (fixed(lo, up) ? kFixedTol : kBoxTol)) ok = false;
    } else if (lb) {
      if (av[i] < -kBoxTol) ok = false;
    } else if (ub) {
      if (av[i] > kBoxTol) ok = false;
    }
  }
  if (primal_feas < 1e-6 && cost_dir < -1e-8 && ok) {
    logf(opt_, 1, "pdlp: detected dual infeasibility (unbounded) at iter %lld (improving ray)", iter);
    res_.primal_ray = std::move(ray);
    return true;
  }
  return false;
}

void Solve::map_back(const std::vector<double>& v, const std::vector<double>& y) {
  // src/lp/pdlp.cpp:1167-1189: x = cs v_x, r = v_r / rs, y = rs y'.
  const int m = lp_.m, n = lp_.n;
  res_.x.assign(static_cast<size_t>(n + m), 0.0);
  for (int j = 0; j < n; ++j) res_.x[j] = v[j] * cs_[j];
  for (int i = 0; i < m; ++i) res_.x[n + i] = v[n + i] / rs_[i];
  res_.y.assign(static_cast<size_t>(m), 0.0);
  for (int i = 0; i < m; ++i) res_.y[i] = y[i] * rs_[i];
}

PdlpResult Solve::run() {
  const int m = lp_.m, n = lp_.n, nt = n + m;
  logf(opt_, 2, "pdlp: m=%d n=%d nnz=%lld tol=%.2e", m, n, static_cast<long long>(lp_.A.nnz()), opt_.tol);
  if (!opt_.use_gpu) {
    res_.message = "this PDLP runs its iterations on the GPU only (use_gpu = false)";
    res_.status = Status::kNumericalError;
    return res_;
  }

  // --- Setup (src/lp/pdlp.cpp:260-318) ---
  if (opt_.algorithm == Algorithm::kEnhanced)
    ruiz_pock_chambolle(lp_.A, lp_.At, rs_, cs_, 10);
  else
    equilibrate(lp_.A, lp_.At, rs_, cs_, 10);
  const Csc A_eq = csc_scaled(lp_.A, rs_, cs_);
  const Csc At_eq = csc_scaled(lp_.At, cs_, rs_);
  std::vector<double> cost = lp_.cost, lower = lp_.lower, upper = lp_.upper;
  for (int j = 0; j < n; ++j) {
    cost[j] *= cs_[j];
    lower[j] /= cs_[j];
    upper[j] /= cs_[j];
  }
  for (int i = 0; i < m; ++i) {
    cost[n + i] *= rs_[i];
    lower[n + i] *= rs_[i];
    upper[n + i] *= rs_[i];
  }
  const double L = operator_norm_bound(A_eq, At_eq);
  res_.operator_norm = L;
  const double eta = 0.99 / L;
  omega_ = 1.0;
  double tau = eta / omega_, sigma = eta * omega_;
  const double theta = 1.0;
  logf(opt_, 3, "pdlp: estimated ||[A -I]|| = %.3e", L);

  long long max_iter = opt_.max_iterations;  // src/lp/pdlp.cpp:365-373
  if (max_iter <= 0) {
    max_iter = std::max<long long>(10000LL, 10000LL * (m + n) / 10);
    if (opt_.time_limit >= kInf) max_iter = std::min<long long>(max_iter, 1000000LL);
  }
  if (m > 0) {  // src/lp/pdlp.cpp:377-394
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int64_t p = lp_.A.start[j]; p < lp_.A.start[j + 1]; ++p) s += lp_.A.value[p] * lp_.A.value[p];
      a_colmax_ = std::max(a_colmax_, std::sqrt(s));
    }
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      for (int64_t p = lp_.At.start[i]; p < lp_.At.start[i + 1]; ++p) s += lp_.At.value[p] * lp_.At.value[p];
      at_rowmax_ = std::max(at_rowmax_, std::sqrt(s));
    }
  }
  for (int k = 0; k < nt; ++k) c_norm_ += lp_.cost[k] * lp_.cost[k];
  c_norm_ = std::sqrt(c_norm_);

  const bool mixed = opt_.precision == Precision::kMixed;
  std::unique_ptr<Device> device;
  std::string why;
  try {
    device = Device::create(A_eq, At_eq, cost, lower, upper, mixed ? SMY_FP32 : SMY_FP64);
    if (device) device->set_original(cs_, rs_, lp_.cost, lp_.lower, lp_.upper);
  } catch (const std::exception& e) {
    device.reset();
    why = e.what();
  }
  if (!device) {
    // src/lp/pdlp.cpp:355-358 (the reference then iterates on the CPU; this library has no CPU
    // iterations and reports the failure instead).
    if (!why.empty()) {
      logf(opt_, 1, "pdlp: the GPU could not start (%s)", why.c_str());
      res_.message = "the GPU could not start: " + why;
    } else {
      logf(opt_, 1, "pdlp: no usable GPU");
      res_.message = std::string("no usable GPU: ") + smy_last_error();
    }
    res_.status = Status::kNumericalError;
    res_.solve_seconds = clock_.seconds();
    return res_;
  }
  {
    char name[256];
    smy_gpu_device_name(name, sizeof name);
    logf(opt_, 1, "pdlp: iterations on the GPU (%s)", name);
  }
  if (!opt_.start_x.empty()) {
    // Warm start: into the equilibrated space (x' = x / cs, r' = rs r, y' = y / rs).
    std::vector<double> v0(static_cast<size_t>(nt)), y0(static_cast<size_t>(m));
    for (int j = 0; j < n; ++j) v0[j] = opt_.start_x[j] / cs_[j];
    for (int i = 0; i < m; ++i) v0[n + i] = opt_.start_x[n + i] * rs_[i];
    for (int i = 0; i < m; ++i) y0[i] = opt_.start_y.empty() ? 0.0 : opt_.start_y[i] / rs_[i];
    device->set_start(v0, y0);
  }
  res_.setup_seconds = clock_.seconds();

  double kkt_at_restart = kInf, kkt_previous = kInf;
  long long k_since = 0;
  bool fp32 = mixed;
  bool done = false;
  Measure best{kInf, kInf, kInf, kInf, 0.0, 0.0};
  std::vector<double> v, v_new, y, y_new, v_erg, y_erg;
  long long iter = 0;
  while (iter < max_iter && !done) {
    if (opt_.time_limit < kInf && clock_.seconds() > opt_.time_limit) {
      res_.status = Status::kTimeLimit;
      res_.iterations = iter;
      logf(opt_, 1, "pdlp: time limit after %lld iterations, %.1f s, KKT at the last restart %.2e; "
                    "last evaluation: primal %.2e, dual %.2e, gap %.2e (tolerance %.0e)",
           iter, clock_.seconds(), kkt_at_restart, best.rel_primal, best.rel_dual, best.rel_gap, opt_.tol);
      break;
    }
    // Steps up to the next evaluated iteration, (iter + 1) % 64 == 0 or the last one.
    const long long eval_iter =
        std::min((iter / kCheckFrequency + 1) * kCheckFrequency - 1, max_iter - 1);
    const int count = static_cast<int>(eval_iter - iter + 1);
    device->run(tau, sigma, theta, k_since, count, opt_.use_graph);
    res_.kernel_ms += device->last_run_ms();
    k_since += count;
    if (fp32) res_.fp32_iterations += count;
    iter = eval_iter;

    double s14[14];
    device->evaluate_both(s14);
    const Measure current = finish(Sums{s14[0], s14[1], s14[2], s14[3], s14[4], s14[5], s14[6]});
    const Measure average = finish(Sums{s14[7], s14[8], s14[9], s14[10], s14[11], s14[12], s14[13]});
    res_.evals.push_back({iter,
                          {current.rel_primal, current.rel_dual, current.rel_gap},
                          {average.rel_primal, average.rel_dual, average.rel_gap}});
    best = worst(average) < worst(current) ? average : current;

    // Termination (src/lp/pdlp.cpp:684-713): strict < tol for the current, then the average.
    const auto converged = [&](const Measure& e) {
      return e.rel_primal < opt_.tol && e.rel_dual < opt_.tol && std::fabs(e.rel_gap) < opt_.tol;
    };
    if (converged(current) || converged(average)) {
      const bool cur = converged(current);
      const Measure& e = cur ? current : average;
      logf(opt_, 2, cur ? "pdlp: converged at iter %lld: primal=%.2e dual=%.2e gap=%.2e"
                        : "pdlp: average converged at iter %lld: primal=%.2e dual=%.2e gap=%.2e",
           iter, e.rel_primal, e.rel_dual, e.rel_gap);
      res_.status = Status::kOptimal;
      res_.iterations = iter + 1;
      res_.primal_residual = e.rel_primal;
      res_.dual_residual = e.rel_dual;
      res_.gap = e.rel_gap;
      if (cur) {
        device->download(nullptr, &v_new, nullptr, &y_new, nullptr, nullptr, nt, m);
        v = v_new;
        y = y_new;
      } else {
        device->download(nullptr, nullptr, nullptr, nullptr, &v_erg, &y_erg, nt, m);
        v = v_erg;
        y = y_erg;
      }
      break;
    }

    // Detection (src/lp/pdlp.cpp:715-1043), at every 8th evaluation after iteration 1000.
    const bool detect = opt_.detect && iter > 1000 && m > 0 &&
                        ((iter + 1) / kCheckFrequency) % kDetectionEvery == 0;
    if (detect) {
      device->download(&v, &v_new, nullptr, &y_new, &v_erg, &y_erg, nt, m);
      if (detect_farkas(iter, y_erg)) {
        res_.status = Status::kInfeasible;
        done = true;
      } else if (detect_ray(iter, v, v_new)) {
        res_.status = Status::kUnbounded;
        done = true;
      }
      if (done) {
        res_.iterations = iter + 1;
        res_.primal_residual = current.rel_primal;
        res_.dual_residual = current.rel_dual;
        res_.gap = current.rel_gap;
        v = v_new;
        y = y_new;
        break;
      }
    }

    // Mixed precision: from FP32 to FP64 steps once FP32 got close to the tolerance.
    bool force_restart = false;
    if (fp32 && std::min(worst(current), worst(average)) < opt_.fp32_switch_tol) {
      logf(opt_, 2, "pdlp: FP32 steps reached %.2e at iter %lld; FP64 steps from here",
           std::min(worst(current), worst(average)), iter);
      device->set_precision(SMY_FP64);
      fp32 = false;
      force_restart = true;
    }

    // Adaptive restart and primal weight (src/lp/pdlp.cpp:1075-1132).
    {
      const double kkt_current = kkt(current);
      const double kkt_average = kkt(average);
      const bool to_average = kkt_average < kkt_current;
      const double kkt_candidate = std::min(kkt_current, kkt_average);
      if (!(kkt_at_restart < kInf)) kkt_at_restart = kkt_candidate;
      const bool restart =
          force_restart || kkt_candidate <= 0.2 * kkt_at_restart ||
          (kkt_candidate <= 0.8 * kkt_at_restart && kkt_candidate > kkt_previous) ||
          static_cast<double>(k_since) >= 0.36 * static_cast<double>(iter + 1);
      kkt_previous = kkt_candidate;
      if (restart) {
        const int which = to_average ? SMY_AVERAGE : SMY_CURRENT;
        double dx = 0.0, dy = 0.0;
        device->restart_distance(which, dx, dy);
        dx = std::sqrt(dx);
        dy = std::sqrt(dy);
        if (dx > kWeightDistanceFloor && dy > kWeightDistanceFloor) {
          omega_ = std::exp(0.5 * std::log(dy / dx) + 0.5 * std::log(omega_));
          tau = eta / omega_;
          sigma = eta * omega_;
        }
        logf(opt_, 2, "pdlp: restart at iter %lld to the %s (KKT %.2e), omega %.3e", iter,
             to_average ? "average" : "current iterate", kkt_candidate, omega_);
        res_.restarts.push_back({iter, to_average ? 1 : 0, kkt_candidate, omega_});
        device->restart_to(which);
        k_since = 0;
        kkt_at_restart = kkt_candidate;
        kkt_previous = kInf;
      }
    }
    device->advance();
    ++iter;
  }
  if (res_.status != Status::kOptimal && res_.status != Status::kInfeasible &&
      res_.status != Status::kUnbounded) {
    // A limit stopped the loop: the current point (src/lp/pdlp.cpp:1153-1165).
    device->download(&v, nullptr, &y, nullptr, nullptr, nullptr, nt, m);
    if (res_.status == Status::kNumericalError) {
      res_.status = Status::kIterationLimit;
      res_.iterations = max_iter;
      logf(opt_, 1, "pdlp: iteration limit reached after %lld iterations", max_iter);
    }
  }
  map_back(v, y);
  res_.solve_seconds = clock_.seconds();
  return res_;
}

}  // namespace

PdlpResult solve_pdlp(const LpProblem& lp, const PdlpOptions& options) {
  try {
    return Solve(lp, options).run();
  } catch (const std::exception& e) {
    // The reference re-solves on the CPU after a GPU failure (src/lp/pdlp.cpp:228-246); this
    // library reports it.
    PdlpResult r;
    r.status = Status::kNumericalError;
    r.message = std::string("the GPU failed during the solve: ") + e.what();
    if (options.log_level >= 1) {
      std::printf("pdlp: the GPU failed during the solve (%s)\n", e.what());
      std::fflush(stdout);
    }
    return r;
  }
}

}  // namespace smy::host
