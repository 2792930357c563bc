#include "pdlp.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <exception>
#include <functional>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>

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
// Per-instance numbers of the device-side detection (include/samaya_pdhg.h,
// smy_pdhg_detect_batch): sum |w|, max |yf|, f_min and f_max (finite parts) with their counts of
// infinite terms, scale, d_max, ||A x - r||^2, cost slope, bound failures.
constexpr int kDetectSlotsHost = 11;

class Clock {
 public:
  double seconds() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
  }

 private:
  std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();
};

void check(int rc) {
  if (rc != SMY_OK) throw std::runtime_error(smy_last_error());
}

// Owner of one smy_pdhg handle over a batch of LPs (one for a single solve); every failure
// becomes an exception (the reference's SAMAYA_CUDA_CHECK behaviour, src/gpu/device.hpp:83-91).
class Device {
 public:
  // Null when there is no usable GPU or the matrices are too large (PdhgDevice::create's null).
  static std::unique_ptr<Device> create(const std::vector<const Csc*>& A, const std::vector<const Csc*>& At,
                                        const std::vector<const std::vector<double>*>& cost,
                                        const std::vector<const std::vector<double>*>& lower,
                                        const std::vector<const std::vector<double>*>& upper,
                                        int precision) {
    const size_t K = A.size();
    std::vector<smy_csc> a(K), at(K);
    std::vector<const double*> c(K), l(K), u(K);
    for (size_t i = 0; i < K; ++i) {
      a[i] = smy_csc{A[i]->rows, A[i]->cols, A[i]->start.data(), A[i]->index.data(), A[i]->value.data()};
      at[i] = smy_csc{At[i]->rows, At[i]->cols, At[i]->start.data(), At[i]->index.data(), At[i]->value.data()};
      c[i] = cost[i]->data();
      l[i] = lower[i]->data();
      u[i] = upper[i]->data();
    }
    smy_pdhg* h = nullptr;
    const int rc = smy_pdhg_create_batch(static_cast<int>(K), a.data(), at.data(), c.data(), l.data(),
                                         u.data(), precision, &h);
    if (rc == SMY_NO_DEVICE || rc == SMY_TOO_LARGE) return nullptr;
    check(rc);
    return std::unique_ptr<Device>(new Device(h, static_cast<int>(K)));
  }
  // One matrix and cost for every instance, per-instance bounds (the logical ones equal).
  static std::unique_ptr<Device> create_shared(const Csc* A, const Csc* At, const std::vector<double>* cost,
                                               const std::vector<const std::vector<double>*>& lower,
                                               const std::vector<const std::vector<double>*>& upper,
                                               int precision) {
    const size_t K = lower.size();
    const smy_csc a{A->rows, A->cols, A->start.data(), A->index.data(), A->value.data()};
    const smy_csc at{At->rows, At->cols, At->start.data(), At->index.data(), At->value.data()};
    std::vector<const double*> l(K), u(K);
    for (size_t i = 0; i < K; ++i) {
      l[i] = lower[i]->data();
      u[i] = upper[i]->data();
    }
    smy_pdhg* h = nullptr;
    const int rc = smy_pdhg_create_shared(&a, &at, static_cast<int>(K), cost->data(), l.data(), u.data(),
                                          precision, &h);
    if (rc == SMY_NO_DEVICE || rc == SMY_TOO_LARGE) return nullptr;
    check(rc);
    return std::unique_ptr<Device>(new Device(h, static_cast<int>(K)));
  }
  ~Device() { smy_pdhg_destroy(h_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;

  int size() const { return k_; }
  void set_fp64_accumulation(bool on) { check(smy_pdhg_set_fp64_accumulation(h_, on ? 1 : 0)); }
  void set_halpern(bool on, double reflection) { check(smy_pdhg_set_halpern(h_, on ? 1 : 0, reflection)); }
  // After a run of Halpern steps: ||v+ - v||^2, ||y+ - y||^2 of each instance's last step.
  void fixed_point(std::vector<double>& out) {
    out.resize(2 * static_cast<size_t>(k_));
    check(smy_pdhg_fixed_point_batch(h_, out.data()));
  }
  void inject_fault(long long step) { check(smy_pdhg_inject_fault(h_, step)); }
  // The enhanced preconditioner on the device (a handle of one unscaled LP): the factors, L and
  // the norms of the termination test.
  // out = {L, a_colmax, at_rowmax}.
  void equilibrate(Algorithm algorithm, int passes, std::vector<double>& cs, std::vector<double>& rs,
                   double out[3]) {
    if (algorithm == Algorithm::kEnhanced)
      check(smy_pdhg_equilibrate(h_, passes, 1, cs.data(), rs.data(), out));
    else
      check(smy_pdhg_equilibrate_reference(h_, passes, cs.data(), rs.data(), out));
  }
  void set_original(int i, const std::vector<double>& cs, const std::vector<double>& rs,
                    const std::vector<double>& cost, const std::vector<double>& lower,
                    const std::vector<double>& upper) {
    check(smy_pdhg_set_original_at(h_, i, cs.data(), rs.data(), cost.data(), lower.data(), upper.data()));
  }
  void set_start(int i, const std::vector<double>& v, const std::vector<double>& y) {
    check(smy_pdhg_set_start_at(h_, i, v.data(), y.data()));
  }
  // `count` steps of every instance still stepping; single steps with advances when `graph` is
  // off (the reference's one launch sequence per iteration).
  void run(const std::vector<double>& tau, const std::vector<double>& sigma, double theta,
           const std::vector<long long>& k0, int count, bool graph) {
    if (graph) {
      check(smy_pdhg_run_batch(h_, tau.data(), sigma.data(), theta, k0.data(), count));
    } else {
      std::vector<long long> k = k0;
      for (int j = 0; j < count; ++j) {
        check(smy_pdhg_run_batch(h_, tau.data(), sigma.data(), theta, k.data(), 1));
        for (long long& t : k) ++t;
        if (j + 1 < count) check(smy_pdhg_advance(h_));
      }
    }
  }
  void rule_init(const std::vector<smy_pdhg_rule_setup>& setup) { check(smy_pdhg_rule_init(h_, setup.data())); }
  void rule_decide(const std::vector<smy_pdhg_rule_in>& in, long long iter1, std::vector<smy_pdhg_rule_out>& out) {
    check(smy_pdhg_rule_decide(h_, in.data(), iter1, out.data()));
  }
  // `count` attempts of the adaptive step-size rule; the state is read and updated.
  void run_adaptive(std::vector<smy_pdhg_adaptive>& state, int count) {
    check(smy_pdhg_run_adaptive(h_, state.data(), count));
  }
  void evaluate(std::vector<double>& out) {
    out.resize(14 * static_cast<size_t>(k_));
    check(smy_pdhg_evaluate_batch(h_, out.data()));
  }
  void restart_distance(const std::vector<int>& which, std::vector<double>& out) {
    out.resize(2 * static_cast<size_t>(k_));
    check(smy_pdhg_restart_distance_batch(h_, which.data(), out.data()));
  }
  void restart_to(const std::vector<int>& which) { check(smy_pdhg_restart_to_batch(h_, which.data())); }
  void set_precision(const std::vector<int>& precision) {
    check(smy_pdhg_set_precision_batch(h_, precision.data()));
  }
  void freeze(int i) { check(smy_pdhg_freeze(h_, i)); }
  void detect(const std::vector<int>& flags, std::vector<double>& out) {
    out.resize(static_cast<size_t>(kDetectSlotsHost) * k_);
    check(smy_pdhg_detect_batch(h_, flags.data(), out.data()));
  }
  void download_ray(int i, std::vector<double>& ray, int n) {
    ray.assign(static_cast<size_t>(n), 0.0);
    check(smy_pdhg_download_ray(h_, i, ray.data()));
  }
  void download_farkas(int i, std::vector<double>& y, int m) {
    y.assign(static_cast<size_t>(m), 0.0);
    check(smy_pdhg_download_farkas(h_, i, y.data()));
  }
  void advance() { check(smy_pdhg_advance(h_)); }
  void download(int i, std::vector<double>* v, std::vector<double>* v_new, std::vector<double>* y,
                std::vector<double>* y_new, std::vector<double>* v_erg, std::vector<double>* y_erg,
                int nt, int m) {
    for (auto* p : {v, v_new, v_erg})
      if (p) p->resize(nt);
    for (auto* p : {y, y_new, y_erg})
      if (p) p->resize(m);
    const auto d = [](std::vector<double>* p) { return p ? p->data() : nullptr; };
    check(smy_pdhg_download_at(h_, i, d(v), d(v_new), d(y), d(y_new), d(v_erg), d(y_erg)));
  }
  double last_run_ms() {
    double ms = 0.0;
    return smy_pdhg_last_run_ms(h_, &ms) == SMY_OK ? ms : 0.0;
  }

 private:
  Device(smy_pdhg* h, int k) : h_(h), k_(k) {}
  smy_pdhg* h_;
  int k_;
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
                         std::vector<double>& cs, int passes, bool pock_chambolle = true) {
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
  if (pock_chambolle) norms(true);
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

// The equilibrated matrix of a problem (src/lp/pdlp.cpp:260-318, 377-394): the row and column
// factors, A_eq = D_r A D_c and its transpose, the bound L on ||[A_eq -I]|| and the norms of the
// termination test. Lanes of one problem (node relaxations differing in bounds) share it.
struct Equilibrated {
  std::vector<double> rs, cs;
  Csc A_eq, At_eq;
  double L = 1.0;
  double a_colmax = 0.0, at_rowmax = 0.0;
};

std::shared_ptr<const Equilibrated> equilibrate_problem(const LpProblem& lp, Algorithm algorithm);

// One LP of a solve: its setup, its decisions, its result. A single solve has one lane; a batch
// has one per LP, each taking exactly the decisions it would take alone.
class Lane {
 public:
  Lane(const LpProblem& lp, const PdlpOptions& opt, std::string tag)
      : lp_(lp), opt_(opt), tag_(std::move(tag)) {}

  // Host setup (src/lp/pdlp.cpp:260-318, 365-394) given the matrix's equilibration: the scaled
  // cost and bounds, the step size, the iteration limit and the norms of the termination test.
  void prepare(std::shared_ptr<const Equilibrated> e);
  void log(int level, const char* fmt, ...) const;
  Measure finish(const Sums& s) const;
  double kkt(const Measure& e) const {
    return std::sqrt(omega * omega * e.rel_primal * e.rel_primal +
                     e.rel_dual * e.rel_dual / (omega * omega) + e.rel_gap * e.rel_gap);
  }
  bool detect_farkas(long long iter, const std::vector<double>& y_erg);
  bool farkas_test(long long iter, const std::vector<double>& y_erg, const char* what);
  bool detect_ray(long long iter, const std::vector<double>& v, const std::vector<double>& v_new);
  // The same decisions from the device-side totals (11 numbers, smy_pdhg_detect_batch).
  bool farkas_from(long long iter, const double* o, const char* what = "") const;
  bool ray_from(long long iter, const double* o) const;
  // The Farkas and ray tests' decisions on the totals, without logging.
  static bool farkas_passes(const double* o);
  static bool ray_passes(const double* o);
  // The sign projection of a scaled dual for the f_min (orient 2) or f_max (3) test: entries that
  // make a row's term infinite set to 0 (smy_pdhg_detect_batch's flags 2 and 3).
  std::vector<double> sign_projected(const std::vector<double>& y, int orient) const;
  void map_back(const std::vector<double>& v, const std::vector<double>& y);

  // The bounds of this LP: the task's own when given, else the problem's.
  const std::vector<double>& lo() const { return lower_ != nullptr ? *lower_ : lp_.lower; }
  const std::vector<double>& up() const { return upper_ != nullptr ? *upper_ : lp_.upper; }

  const LpProblem& lp_;
  const PdlpOptions& opt_;
  std::string tag_;  // "pdlp" alone, "pdlp[i]" in a batch
  // Warm start of this LP (null: the options' start_x / start_y).
  const std::vector<double>* start_x = nullptr;
  const std::vector<double>* start_y = nullptr;
  const std::vector<double>* lower_ = nullptr;
  const std::vector<double>* upper_ = nullptr;
  PdlpResult res;
  // Equilibration of the matrix (shared by the lanes of one problem).
  std::shared_ptr<const Equilibrated> eq;
  std::vector<double> cost, lower, upper;  // equilibrated
  std::vector<double> v0, y0;              // warm start, equilibrated (empty: cold start)
  double eta = 0.0;
  double c_norm = 0.0;
  long long max_iter = 0;
  // Iteration state.
  double omega = 1.0, tau = 0.0, sigma = 0.0;
  long long k_since = 0;
  // Adaptive steps: the rule's state (eta above is then the next attempt's step size).
  smy_pdhg_adaptive adapt{0.0, 1.0, 0.0, 0, 0};
  double kkt_at_restart = kInf, kkt_previous = kInf;
  double weight_integral = 0.0;  // pid_primal_weight: the controller's sum of errors
  bool fp32 = false;
  double fp32_best = kInf;       // best distance to termination in FP32 (by 20% steps)
  long long fp32_best_iter = 0;  // ... and where it was reached
  int fp32_below = 0;            // consecutive evaluations below fp32_stall_level
  double fp32_recent[4] = {kInf, kInf, kInf, kInf};  // Halpern steps: the last 4 distances
  int fp32_seen = 0;
  bool fp32_armed = false;       // the stall test counts (fp32_below reached kStallArm once)
  bool active = true;
  // Every structural variable has two finite bounds: the LP is bounded and has no improving ray,
  // so the ray test is not taken (the device skips it when every tested lane is boxed).
  bool boxed = false;
  Measure best{kInf, kInf, kInf, kInf, 0.0, 0.0};
};

void Lane::log(int level, const char* fmt, ...) const {
  if (level > opt_.log_level) return;
  std::fputs(tag_.c_str(), stdout);
  va_list args;
  va_start(args, fmt);
  std::vfprintf(stdout, fmt, args);
  va_end(args);
  std::fputc('\n', stdout);
  std::fflush(stdout);
}

std::shared_ptr<const Equilibrated> equilibrate_problem(const LpProblem& lp, Algorithm algorithm) {
  auto e = std::make_shared<Equilibrated>();
  const int m = lp.m, n = lp.n;
  if (algorithm == Algorithm::kEnhanced)
    ruiz_pock_chambolle(lp.A, lp.At, e->rs, e->cs, 10);
  else
    equilibrate(lp.A, lp.At, e->rs, e->cs, 10);
  e->A_eq = csc_scaled(lp.A, e->rs, e->cs);
  e->At_eq = csc_scaled(lp.At, e->cs, e->rs);
  e->L = operator_norm_bound(e->A_eq, e->At_eq);
  if (m > 0) {  // src/lp/pdlp.cpp:377-394
    for (int j = 0; j < n; ++j) {
      double t = 0.0;
      for (int64_t p = lp.A.start[j]; p < lp.A.start[j + 1]; ++p) t += lp.A.value[p] * lp.A.value[p];
      e->a_colmax = std::max(e->a_colmax, std::sqrt(t));
    }
    for (int i = 0; i < m; ++i) {
      double t = 0.0;
      for (int64_t p = lp.At.start[i]; p < lp.At.start[i + 1]; ++p) t += lp.At.value[p] * lp.At.value[p];
      e->at_rowmax = std::max(e->at_rowmax, std::sqrt(t));
    }
  }
  return e;
}

}  // namespace

void equilibrate_host(const Csc& A, const Csc& At, Algorithm algorithm, int passes,
                      bool pock_chambolle, std::vector<double>& rs, std::vector<double>& cs,
                      double out[3]) {
  if (algorithm == Algorithm::kEnhanced)
    ruiz_pock_chambolle(A, At, rs, cs, passes, pock_chambolle);
  else
    equilibrate(A, At, rs, cs, passes);
  const Csc A_eq = csc_scaled(A, rs, cs), At_eq = csc_scaled(At, cs, rs);
  out[0] = operator_norm_bound(A_eq, At_eq);
  double a_colmax = 0.0, at_rowmax = 0.0;
  if (A.rows > 0) {
    for (int j = 0; j < A.cols; ++j) {
      double t = 0.0;
      for (int64_t p = A.start[j]; p < A.start[j + 1]; ++p) t += A.value[p] * A.value[p];
      a_colmax = std::max(a_colmax, std::sqrt(t));
    }
    for (int i = 0; i < At.cols; ++i) {
      double t = 0.0;
      for (int64_t p = At.start[i]; p < At.start[i + 1]; ++p) t += At.value[p] * At.value[p];
      at_rowmax = std::max(at_rowmax, std::sqrt(t));
    }
  }
  out[1] = a_colmax;
  out[2] = at_rowmax;
}

namespace {

void Lane::prepare(std::shared_ptr<const Equilibrated> e) {
  const int m = lp_.m, n = lp_.n, nt = n + m;
  eq = std::move(e);
  const std::vector<double>& c_s = eq->cs;
  const std::vector<double>& r_s = eq->rs;
  cost = lp_.cost;
  lower = lo();
  upper = up();
  boxed = true;
  for (int j = 0; j < n && boxed; ++j) boxed = lower[j] > -kInf && upper[j] < kInf;
  for (int j = 0; j < n; ++j) {
    cost[j] *= c_s[j];
    lower[j] /= c_s[j];
    upper[j] /= c_s[j];
  }
  for (int i = 0; i < m; ++i) {
    cost[n + i] *= r_s[i];
    lower[n + i] *= r_s[i];
    upper[n + i] *= r_s[i];
  }
  double L = eq->L;
  // Halpern steps converge only while T is nonexpansive (eta ||[A_eq -I]|| <= 1), and 20 power
  // iterations can fall short of the norm when the top singular values cluster, which the
  // enhanced preconditioner makes them do (seen: L 1.391 for a norm of 1.410, and the iterates
  // diverged). Its last step, Pock-Chambolle with alpha = 1, bounds ||A_eq|| by 1 (Schur's test
  // with the row and column weights it divided by), so ||[A_eq -I]|| <= sqrt(2): that bound is
  // used instead.
  if (opt_.halpern && opt_.algorithm == Algorithm::kEnhanced && lp_.m > 0) L = std::sqrt(2.0);
  res.operator_norm = L;
  eta = 0.99 / L;
  omega = opt_.start_omega > 0.0 ? opt_.start_omega : 1.0;
  tau = eta / omega;
  sigma = eta * omega;
  adapt = smy_pdhg_adaptive{eta, omega, 0.0, 0, 0};
  log(3, ": estimated ||[A -I]|| = %.3e", L);

  max_iter = opt_.max_iterations;  // src/lp/pdlp.cpp:365-373
  if (max_iter <= 0) {
    max_iter = std::max<long long>(10000LL, 10000LL * (m + n) / 10);
    if (opt_.time_limit >= kInf) max_iter = std::min<long long>(max_iter, 1000000LL);
  }
  for (int k = 0; k < nt; ++k) c_norm += lp_.cost[k] * lp_.cost[k];
  c_norm = std::sqrt(c_norm);

  const std::vector<double>& sx = start_x != nullptr ? *start_x : opt_.start_x;
  const std::vector<double>& sy = start_y != nullptr ? *start_y : opt_.start_y;
  if (!sx.empty()) {
    // Warm start: into the equilibrated space (x' = x / cs, r' = rs r, y' = y / rs).
    v0.assign(static_cast<size_t>(nt), 0.0);
    y0.assign(static_cast<size_t>(m), 0.0);
    for (int j = 0; j < n; ++j) v0[j] = sx[j] / c_s[j];
    for (int i = 0; i < m; ++i) v0[n + i] = sx[n + i] * r_s[i];
    for (int i = 0; i < m; ++i) y0[i] = sy.empty() ? 0.0 : sy[i] / r_s[i];
  }
}

Measure Lane::finish(const Sums& s) const {
  // src/lp/pdlp.cpp:493-519.
  const int m = lp_.m;
  const double primal_scale = 1.0 + eq->a_colmax * std::sqrt(s.x_sq) + std::sqrt(s.r_sq);
  const double dual_scale = 1.0 + eq->at_rowmax * std::sqrt(s.y_sq) + c_norm;
  const double gap_scale = 1.0 + std::fabs(s.primal_obj) + std::fabs(s.dual_obj);
  const double gap = s.primal_obj - s.dual_obj;
  return Measure{m > 0 ? std::sqrt(s.primal_sq) / primal_scale : 0.0,
                 m > 0 ? std::sqrt(s.dual_sq) / dual_scale : 0.0, gap / gap_scale, gap,
                 s.primal_obj, s.dual_obj};
}

// Farkas certificate test on the average dual (src/lp/pdlp.cpp:724-842).
std::vector<double> Lane::sign_projected(const std::vector<double>& y, int orient) const {
  const int n = lp_.n;
  std::vector<double> out(y);
  for (int i = 0; i < lp_.m; ++i) {
    const double s = orient == 2 ? y[i] : -y[i];  // oriented as for f_min
    if ((s < 0.0 && !(lo()[n + i] > -kInf)) || (s > 0.0 && !(up()[n + i] < kInf))) out[i] = 0.0;
  }
  return out;
}

bool Lane::detect_farkas(long long iter, const std::vector<double>& y_erg) {
  // The reference's test on y_erg, then (farkas_projection) on its sign projections.
  if (farkas_test(iter, y_erg, "")) return true;
  if (!opt_.farkas_projection) return false;
  for (int orient : {2, 3})
    if (farkas_test(iter, sign_projected(y_erg, orient), ", sign-projected dual")) return true;
  return false;
}

bool Lane::farkas_test(long long iter, const std::vector<double>& y_erg, const char* what) {
  const int m = lp_.m, n = lp_.n;
  std::vector<double> yf(static_cast<size_t>(m));
  double y_max = 0.0;
  for (int i = 0; i < m; ++i) {
    yf[i] = y_erg[i] * eq->rs[i];
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
    const double lo = this->lo()[k], up = this->up()[k];
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
      if (lo()[k] > -kInf) scale = std::max(scale, std::fabs(t * lo()[k]));
      if (up()[k] < kInf) scale = std::max(scale, std::fabs(t * up()[k]));
    }
  }
  if (!std::isfinite(scale)) scale = 0.0;
  const double margin = 1e-8 * (1.0 + scale);
  log(2, ": infeas check iter=%lld y_max=%.2e min_f=%.2e max_f=%.2e margin=%.2e scale=%.2e",
      iter, y_max, f_min, f_max, margin, scale);
  if ((std::isfinite(f_min) && f_min > margin) || (std::isfinite(f_max) && f_max < -margin)) {
    log(1, ": detected primal infeasibility at iter %lld (Farkas certificate%s)", iter, what);
    res.dual_ray = y_erg;
    return true;
  }
  return false;
}

// Improving-ray test on the difference of consecutive iterates (src/lp/pdlp.cpp:852-1042).
bool Lane::detect_ray(long long iter, const std::vector<double>& v, const std::vector<double>& v_new) {
  const int m = lp_.m, n = lp_.n;
  std::vector<double> d(v_new.size());
  double d_max = 0.0;
  for (size_t k = 0; k < d.size(); ++k) {
    d[k] = v_new[k] - v[k];
    d_max = std::max(d_max, std::fabs(d[k]));
  }
  if (!(d_max > 1e-6)) return false;
  for (double& t : d) t /= d_max;
  std::vector<double> r(static_cast<size_t>(m), 0.0);
  for (int j = 0; j < n; ++j) {
    const double xj = d[j] * eq->cs[j];
    if (xj == 0.0) continue;
    for (int64_t p = lp_.A.start[j]; p < lp_.A.start[j + 1]; ++p) r[lp_.A.index[p]] += lp_.A.value[p] * xj;
  }
  for (int i = 0; i < m; ++i) r[i] -= d[n + i] / eq->rs[i];
  double primal_feas = 0.0;
  for (double t : r) primal_feas += t * t;
  primal_feas = std::sqrt(primal_feas);
  double cost_dir = 0.0;
  for (int j = 0; j < n; ++j) cost_dir += lp_.cost[j] * (d[j] * eq->cs[j]);
  for (int i = 0; i < m; ++i) cost_dir += lp_.cost[n + i] * (d[n + i] / eq->rs[i]);
  constexpr double kBoxTol = 1e-4, kFixedTol = 1e-3;
  const auto fixed = [](double lo, double up) { return up - lo <= 1e-10 * (1.0 + std::fabs(lo)); };
  bool ok = true;
  std::vector<double> ray(static_cast<size_t>(n), 0.0);
  for (int j = 0; j < n && ok; ++j) {
    double xj = d[j] * eq->cs[j];
    const double lo = this->lo()[j], up = this->up()[j];
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
    const double lo = this->lo()[n + i], up = this->up()[n + i];
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
    log(1, ": detected dual infeasibility (unbounded) at iter %lld (improving ray)", iter);
    res.primal_ray = std::move(ray);
    return true;
  }
  return false;
}

bool Lane::ray_passes(const double* o) {
  // src/lp/pdlp.cpp:852-1042 from the device totals.
  return o[7] > 1e-6 && std::sqrt(o[8]) < 1e-6 && o[9] < -1e-8 && o[10] == 0.0;
}

bool Lane::farkas_passes(const double* o) {
  // src/lp/pdlp.cpp:724-842 from the device totals.
  if (!(o[1] > 1e-6)) return false;
  const double f_min = o[4] > 0.0 ? -INFINITY : o[2];
  const double f_max = o[5] > 0.0 ? INFINITY : o[3];
  const double scale = std::isfinite(o[6]) ? o[6] : 0.0;
  const double margin = 1e-8 * (1.0 + scale);
  return (std::isfinite(f_min) && f_min > margin) || (std::isfinite(f_max) && f_max < -margin);
}

bool Lane::farkas_from(long long iter, const double* o, const char* what) const {
  if (!(o[1] > 1e-6)) return false;
  const double scale = std::isfinite(o[6]) ? o[6] : 0.0;
  log(2, ": infeas check iter=%lld y_max=%.2e min_f=%.2e max_f=%.2e margin=%.2e scale=%.2e%s", iter,
      o[1], o[4] > 0.0 ? -INFINITY : o[2], o[5] > 0.0 ? INFINITY : o[3], 1e-8 * (1.0 + scale), o[6], what);
  if (farkas_passes(o)) {
    log(1, ": detected primal infeasibility at iter %lld (Farkas certificate%s)", iter, what);
    return true;
  }
  return false;
}

bool Lane::ray_from(long long iter, const double* o) const {
  if (ray_passes(o)) {
    log(1, ": detected dual infeasibility (unbounded) at iter %lld (improving ray)", iter);
    return true;
  }
  return false;
}

void Lane::map_back(const std::vector<double>& v, const std::vector<double>& y) {
  // src/lp/pdlp.cpp:1167-1189: x = cs v_x, r = v_r / rs, y = rs y'.
  // The enhanced algorithm first puts the point back into its box: an average kept in FP32 drifts
  // off a fixed bound by rounding over many steps (by up to ~3e-6 relative on the node
  // relaxations), though every averaged iterate lies in the box. The reference's algorithm
  // returns the point as it is.
  const int m = lp_.m, n = lp_.n;
  const bool clip = opt_.algorithm == Algorithm::kEnhanced;
  const auto at = [&](int k) { return clip ? std::min(std::max(v[k], lower[k]), upper[k]) : v[k]; };
  res.x.assign(static_cast<size_t>(n + m), 0.0);
  for (int j = 0; j < n; ++j) res.x[j] = at(j) * eq->cs[j];
  for (int i = 0; i < m; ++i) res.x[n + i] = at(n + i) / eq->rs[i];
  res.y.assign(static_cast<size_t>(m), 0.0);
  for (int i = 0; i < m; ++i) res.y[i] = y[i] * eq->rs[i];
}

// Runs fn(i) for i in [0, count) on up to `threads` host threads.
template <typename F>
void parallel_for(int count, int threads, F&& fn) {
  threads = std::max(1, std::min(threads, count));
  if (threads == 1) {
    for (int i = 0; i < count; ++i) fn(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  std::vector<std::exception_ptr> errors(static_cast<size_t>(threads));
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t] {
      try {
        for (int i = next++; i < count; i = next++) fn(i);
      } catch (...) {
        errors[t] = std::current_exception();
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errors)
    if (e) std::rethrow_exception(e);
}

int host_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return static_cast<int>(std::min(16u, std::max(1u, hw)));
}

// The PDLP loop of src/lp/pdlp.cpp:522-1150 over a set of lanes sharing one device handle. Every
// lane steps, evaluates, restarts, switches precision and stops on its own schedule; the lanes
// only share the kernel launches.
std::vector<PdlpResult> run_lanes(std::vector<Lane>& lanes, const PdlpOptions& opt, const Clock& clock,
                                  PdlpBatchInfo* info) {
  const int K = static_cast<int>(lanes.size());
  for (Lane& L : lanes)
    L.log(2, ": m=%d n=%d nnz=%lld tol=%.2e", L.lp_.m, L.lp_.n, static_cast<long long>(L.lp_.A.nnz()), opt.tol);
  const auto fail_all = [&](Status status, const std::string& message) {
    std::vector<PdlpResult> out(static_cast<size_t>(K));
    for (PdlpResult& r : out) {
      r.status = status;
      r.message = message;
      r.solve_seconds = clock.seconds();
    }
    return out;
  };
  if (!opt.use_gpu) {
    return fail_all(Status::kNumericalError,
                    "this PDLP runs its iterations on the GPU only (use_gpu = false)");
  }
  // --- Setup (src/lp/pdlp.cpp:260-318), lanes in parallel on the host; lanes of one problem
  // share its equilibration ---
  std::vector<int> first(static_cast<size_t>(K));
  std::vector<int> owners;
  for (int i = 0; i < K; ++i) {
    first[i] = i;
    for (int o : owners)
      if (&lanes[o].lp_ == &lanes[i].lp_) first[i] = o;
    if (first[i] == i) owners.push_back(i);
  }
  // The preconditioner of a single LP (either algorithm's) runs on the device: the handle is
  // created from the unscaled LP and scales itself (the same factors as the host's).
  const bool device_eq = opt.device_scaling && K == 1;
  std::vector<std::shared_ptr<const Equilibrated>> eqs(static_cast<size_t>(K));
  const int threads = K > 1 ? host_threads() : 1;
  if (!device_eq) {
    parallel_for(static_cast<int>(owners.size()), threads, [&](int q) {
      eqs[owners[q]] = equilibrate_problem(lanes[owners[q]].lp_, opt.algorithm);
    });
    parallel_for(K, threads, [&](int i) { lanes[i].prepare(eqs[first[i]]); });
  }

  // One problem for all lanes, with one set of row bounds: the shared-matrix batch (iterates
  // stored instance-minor, SpMM products); otherwise the stacked batch.
  const bool halpern = opt.halpern;
  const bool adaptive = opt.adaptive_steps && !halpern;
  bool share = opt.share_matrix && !adaptive && !opt.fp64_accumulation && K > 1 &&
               owners.size() == 1 && !device_eq;
  for (int i = 1; share && i < K; ++i) {
    const int n = lanes[0].lp_.n;
    share = std::equal(lanes[i].lower.begin() + n, lanes[i].lower.end(), lanes[0].lower.begin() + n) &&
            std::equal(lanes[i].upper.begin() + n, lanes[i].upper.end(), lanes[0].upper.begin() + n);
  }

  const bool mixed = opt.precision == Precision::kMixed;
  std::unique_ptr<Device> device;
  std::string why;
  try {
    if (device_eq) {
      Lane& L0 = lanes[0];
      device = Device::create({&L0.lp_.A}, {&L0.lp_.At}, {&L0.lp_.cost}, {&L0.lo()}, {&L0.up()},
                              mixed ? SMY_FP32 : SMY_FP64);
      if (device) {
        auto e = std::make_shared<Equilibrated>();
        e->cs.assign(static_cast<size_t>(L0.lp_.n), 1.0);
        e->rs.assign(static_cast<size_t>(L0.lp_.m), 1.0);
        double out[3];
        device->equilibrate(opt.algorithm, 10, e->cs, e->rs, out);
        e->L = out[0];
        e->a_colmax = out[1];
        e->at_rowmax = out[2];
        L0.prepare(e);
      }
    } else {
      std::vector<const Csc*> A, At;
      std::vector<const std::vector<double>*> c, l, u;
      for (Lane& L : lanes) {
        A.push_back(&L.eq->A_eq);
        At.push_back(&L.eq->At_eq);
        c.push_back(&L.cost);
        l.push_back(&L.lower);
        u.push_back(&L.upper);
      }
      device = share ? Device::create_shared(A[0], At[0], c[0], l, u, mixed ? SMY_FP32 : SMY_FP64)
                     : Device::create(A, At, c, l, u, mixed ? SMY_FP32 : SMY_FP64);
    }
    if (device) {
      if (share) lanes[0].log(2, ": %d LPs share one matrix: shared-matrix batch", K);
      if (opt.fp64_accumulation) device->set_fp64_accumulation(true);
      if (halpern) device->set_halpern(true, opt.reflection);
      if (opt.inject_fault_step >= 0) device->inject_fault(opt.inject_fault_step);
      for (int i = 0; i < K; ++i) {
        Lane& L = lanes[i];
        device->set_original(i, L.eq->cs, L.eq->rs, L.lp_.cost, L.lo(), L.up());
      }
    }
  } catch (const std::exception& e) {
    device.reset();
    why = e.what();
  }
  if (!device) {
    // src/lp/pdlp.cpp:355-358 (the reference then iterates on the CPU; this library has no CPU
    // iterations and reports the failure instead).
    for (Lane& L : lanes) {
      if (!why.empty())
        L.log(1, ": the GPU could not start (%s)", why.c_str());
      else
        L.log(1, ": no usable GPU");
    }
    return fail_all(Status::kNumericalError, !why.empty() ? "the GPU could not start: " + why
                                                          : std::string("no usable GPU: ") + smy_last_error());
  }
  {
    char name[256];
    smy_gpu_device_name(name, sizeof name);
    for (Lane& L : lanes) L.log(1, ": iterations on the GPU (%s)", name);
  }
  for (int i = 0; i < K; ++i)
    if (!lanes[i].v0.empty()) device->set_start(i, lanes[i].v0, lanes[i].y0);
  const double setup_seconds = clock.seconds();
  for (Lane& L : lanes) {
    L.res.setup_seconds = setup_seconds;
    L.fp32 = mixed;
  }
  if (info != nullptr) info->setup_seconds = setup_seconds;

  const double theta = 1.0;
  std::vector<double> tau(static_cast<size_t>(K)), sigma(static_cast<size_t>(K)), sums, dist, dout, fixed;
  std::vector<long long> k0(static_cast<size_t>(K));
  std::vector<int> which(static_cast<size_t>(K)), precision(static_cast<size_t>(K));
  std::vector<smy_pdhg_adaptive> adapt(static_cast<size_t>(K));
  // The restart rule on the device: decisions, distances, primal weights and restarts there; the
  // host keeps termination, detection and precision switches.
  // Halpern steps restart on their fixed-point residual, decided on the host.
  const bool device_rule = opt.device_restart && !halpern && !opt.pid_primal_weight;
  std::vector<smy_pdhg_rule_in> rule_in(static_cast<size_t>(K));
  std::vector<smy_pdhg_rule_out> rule_out(static_cast<size_t>(K));
  if (device_rule) {
    std::vector<smy_pdhg_rule_setup> setup(static_cast<size_t>(K));
    for (int i = 0; i < K; ++i) {
      const Lane& L = lanes[i];
      setup[i] = smy_pdhg_rule_setup{L.eq->a_colmax, L.eq->at_rowmax, L.c_norm, L.lp_.m > 0 ? 1 : 0, L.omega};
    }
    device->rule_init(setup);
  }
  std::vector<double> v, v_new, y, y_new, v_erg, y_erg;
  int active = K;
  const auto stop = [&](int i) {
    device->freeze(i);
    lanes[i].active = false;
    lanes[i].res.solve_seconds = clock.seconds();
    --active;
  };
  // The lane's answer at its current point (a limit stopped it: src/lp/pdlp.cpp:1153-1165); with
  // Halpern steps T(z) of the last run, which is within the bounds where z need not be.
  const auto stop_at_current = [&](int i) {
    Lane& L = lanes[i];
    const int nt = L.lp_.n + L.lp_.m;
    if (halpern) {
      device->download(i, nullptr, nullptr, nullptr, nullptr, &v_erg, &y_erg, nt, L.lp_.m);
      L.map_back(v_erg, y_erg);
    } else {
      device->download(i, &v, nullptr, &y, nullptr, nullptr, nullptr, nt, L.lp_.m);
      L.map_back(v, y);
    }
    stop(i);
  };
  long long iter = 0;
  double next_progress = clock.seconds() + 10.0;
  while (active > 0) {
    if (K > 1 && opt.log_level >= 1 && clock.seconds() >= next_progress) {
      // A long batch reports itself every 10 seconds.
      std::printf("pdlp batch: iteration %lld, %d of %d LPs still running, %.1f s\n", iter, active, K,
                  clock.seconds());
      std::fflush(stdout);
      next_progress = clock.seconds() + 10.0;
    }
    if (opt.time_limit < kInf && clock.seconds() > opt.time_limit) {
      for (int i = 0; i < K; ++i) {
        Lane& L = lanes[i];
        if (!L.active) continue;
        L.res.status = Status::kTimeLimit;
        L.res.iterations = iter;
        L.log(1, ": time limit after %lld iterations, %.1f s, KKT at the last restart %.2e; "
                 "last evaluation: primal %.2e, dual %.2e, gap %.2e (tolerance %.0e)",
              iter, clock.seconds(), L.kkt_at_restart, L.best.rel_primal, L.best.rel_dual,
              L.best.rel_gap, opt.tol);
        stop_at_current(i);
      }
      break;
    }
    // Steps up to the next evaluated iteration of the earliest lane: (iter + 1) % 64 == 0 or the
    // lane's last one.
    long long eval_iter = LLONG_MAX;
    for (const Lane& L : lanes)
      if (L.active)
        eval_iter = std::min(eval_iter, std::min((iter / kCheckFrequency + 1) * kCheckFrequency - 1,
                                                 L.max_iter - 1));
    const int count = static_cast<int>(eval_iter - iter + 1);
    if (adaptive) {
      for (int i = 0; i < K; ++i) adapt[i] = lanes[i].adapt;
      device->run_adaptive(adapt, count);
      for (int i = 0; i < K; ++i) {
        Lane& L = lanes[i];
        if (!L.active) continue;
        L.adapt = adapt[i];
        L.res.rejected_steps = adapt[i].rejected;
      }
    } else {
      for (int i = 0; i < K; ++i) {
        tau[i] = lanes[i].tau;
        sigma[i] = lanes[i].sigma;
        k0[i] = lanes[i].k_since;
      }
      device->run(tau, sigma, theta, k0, count, opt.use_graph);
    }
    const double ms = device->last_run_ms();
    if (info != nullptr) info->kernel_ms += ms;
    for (Lane& L : lanes) {
      if (!L.active) continue;
      L.res.kernel_ms += ms;
      L.k_since += count;
      if (L.fp32) L.res.fp32_iterations += count;
    }
    iter = eval_iter;

    device->evaluate(sums);
    fixed.clear();
    std::fill(which.begin(), which.end(), -1);
    std::fill(precision.begin(), precision.end(), -1);
    std::vector<char> force(static_cast<size_t>(K), 0), due(static_cast<size_t>(K), 0);
    std::vector<double> kkt_candidate(static_cast<size_t>(K), 0.0);
    std::vector<Measure> cur_m(static_cast<size_t>(K)), avg_m(static_cast<size_t>(K));
    // Measures and termination of the lanes at an evaluation point (every 64th iteration and
    // their last).
    for (int i = 0; i < K; ++i) {
      Lane& L = lanes[i];
      if (!L.active || ((iter + 1) % kCheckFrequency != 0 && iter != L.max_iter - 1)) continue;
      const double* s14 = &sums[14 * static_cast<size_t>(i)];
      const Measure current = L.finish(Sums{s14[0], s14[1], s14[2], s14[3], s14[4], s14[5], s14[6]});
      const Measure average = L.finish(Sums{s14[7], s14[8], s14[9], s14[10], s14[11], s14[12], s14[13]});
      cur_m[i] = current;
      avg_m[i] = average;
      L.res.evals.push_back({iter,
                             {current.rel_primal, current.rel_dual, current.rel_gap},
                             {average.rel_primal, average.rel_dual, average.rel_gap}});
      L.best = worst(average) < worst(current) ? average : current;
      const int nt = L.lp_.n + L.lp_.m, m = L.lp_.m;

      // Termination (src/lp/pdlp.cpp:684-713): strict < tol for the current, then the average.
      // Halpern steps: T(z) only (in the averages' slot).
      const auto converged = [&](const Measure& e) {
        return e.rel_primal < opt.tol && e.rel_dual < opt.tol && std::fabs(e.rel_gap) < opt.tol;
      };
      if ((!halpern && converged(current)) || converged(average)) {
        const bool cur = !halpern && converged(current);
        const Measure& e = cur ? current : average;
        L.log(2, cur ? ": converged at iter %lld: primal=%.2e dual=%.2e gap=%.2e"
                     : ": average converged at iter %lld: primal=%.2e dual=%.2e gap=%.2e",
              iter, e.rel_primal, e.rel_dual, e.rel_gap);
        L.res.status = Status::kOptimal;
        L.res.iterations = iter + 1;
        L.res.primal_residual = e.rel_primal;
        L.res.dual_residual = e.rel_dual;
        L.res.gap = e.rel_gap;
        if (cur) {
          device->download(i, nullptr, &v_new, nullptr, &y_new, nullptr, nullptr, nt, m);
          L.map_back(v_new, y_new);
        } else {
          device->download(i, nullptr, nullptr, nullptr, nullptr, &v_erg, &y_erg, nt, m);
          L.map_back(v_erg, y_erg);
        }
        stop(i);
        continue;
      }
      due[i] = 1;
    }

    // Detection (src/lp/pdlp.cpp:715-1043), at every 8th evaluation after iteration 1000: the
    // Farkas test, then the ray test. On the device for all lanes at once (only per-instance
    // totals come back), or on the host from downloaded iterates (the reference's arithmetic).
    std::vector<int> dflag(static_cast<size_t>(K), 0);
    bool any_detect = false;
    for (int i = 0; i < K; ++i)
      if (due[i] && opt.detect && iter > 1000 && lanes[i].lp_.m > 0 &&
          ((iter + 1) / kCheckFrequency) % kDetectionEvery == 0) {
        dflag[i] = 1;
        any_detect = true;
      }
    if (any_detect && opt.device_detect) device->detect(dflag, dout);
    // farkas_projection: the lanes whose Farkas test failed test other candidates, in one device
    // pass per candidate for all of them (smy_pdhg_detect_batch flags): the sign projection of
    // their dual for f_min (2) or f_max (3) when that sum had infinite terms; the dual's change
    // since the restart point as it is (5), then projected (6, 7) when that had infinite terms.
    // The first that passes is the certificate (downloaded as tested).
    std::vector<int> found(static_cast<size_t>(K), 0);
    std::vector<std::vector<double>> certificate(static_cast<size_t>(K));
    if (any_detect && opt.device_detect && opt.farkas_projection) {
      std::vector<double> diff(dout.size(), 0.0);  // per lane: the numbers of candidate 5
      const auto pass = [&](const std::function<int(int)>& pick) {
        std::vector<int> pflag(static_cast<size_t>(K), 0);
        bool any = false;
        for (int i = 0; i < K; ++i) {
          const double* r = &dout[static_cast<size_t>(i) * kDetectSlotsHost];
          if (dflag[i] && found[i] == 0 && !Lane::farkas_passes(r) && (lanes[i].boxed || !Lane::ray_passes(r))) {
            pflag[i] = pick(i);
            any = any || pflag[i] != 0;
          }
        }
        if (!any) return;
        std::vector<double> pout;
        device->detect(pflag, pout);
        for (int i = 0; i < K; ++i) {
          if (pflag[i] == 0) continue;
          const double* o = &pout[static_cast<size_t>(i) * kDetectSlotsHost];
          if (pflag[i] == 5)
            std::copy(o, o + kDetectSlotsHost, diff.begin() + static_cast<ptrdiff_t>(i) * kDetectSlotsHost);
          if (Lane::farkas_passes(o)) {
            found[i] = pflag[i];
            device->download_farkas(i, certificate[i], lanes[i].lp_.m);
          }
        }
      };
      const auto raw = [&](int i) { return &dout[static_cast<size_t>(i) * kDetectSlotsHost]; };
      const auto dif = [&](int i) { return &diff[static_cast<size_t>(i) * kDetectSlotsHost]; };
      pass([&](int i) { return raw(i)[1] > 1e-6 && raw(i)[4] > 0.0 ? 2 : 0; });
      pass([&](int i) { return raw(i)[1] > 1e-6 && raw(i)[5] > 0.0 ? 3 : 0; });
      pass([&](int) { return 5; });
      pass([&](int i) { return dif(i)[1] > 1e-6 && dif(i)[4] > 0.0 ? 6 : 0; });
      pass([&](int i) { return dif(i)[1] > 1e-6 && dif(i)[5] > 0.0 ? 7 : 0; });
    }
    for (int i = 0; any_detect && i < K; ++i) {
      if (!dflag[i]) continue;
      Lane& L = lanes[i];
      const int nt = L.lp_.n + L.lp_.m, m = L.lp_.m;
      bool done = false;
      if (opt.device_detect) {
        const double* o = &dout[static_cast<size_t>(i) * kDetectSlotsHost];
        if (L.farkas_from(iter, o)) {
          device->download(i, nullptr, nullptr, nullptr, nullptr, nullptr, &y_erg, nt, m);
          L.res.dual_ray = y_erg;
          L.res.status = Status::kInfeasible;
          done = true;
        } else if (found[i] != 0) {
          L.log(1, ": detected primal infeasibility at iter %lld (Farkas certificate, %s)", iter,
                found[i] < 4 ? "sign-projected dual" : found[i] == 5 ? "the dual's change since the restart"
                                                                     : "sign-projected change of the dual");
          L.res.dual_ray = std::move(certificate[i]);
          L.res.status = Status::kInfeasible;
          done = true;
        } else if (!L.boxed && L.ray_from(iter, o)) {
          device->download_ray(i, L.res.primal_ray, L.lp_.n);
          L.res.status = Status::kUnbounded;
          done = true;
        }
        if (done) device->download(i, nullptr, &v_new, nullptr, &y_new, nullptr, nullptr, nt, m);
      } else {
        device->download(i, &v, &v_new, nullptr, &y_new, &v_erg, &y_erg, nt, m);
        if (L.detect_farkas(iter, y_erg)) {
          L.res.status = Status::kInfeasible;
          done = true;
        } else if (!L.boxed && L.detect_ray(iter, v, v_new)) {
          L.res.status = Status::kUnbounded;
          done = true;
        }
      }
      if (done) {
        L.res.iterations = iter + 1;
        L.res.primal_residual = cur_m[i].rel_primal;
        L.res.dual_residual = cur_m[i].rel_dual;
        L.res.gap = cur_m[i].rel_gap;
        L.map_back(v_new, y_new);
        stop(i);
        due[i] = 0;
      }
    }

    for (int i = 0; i < K; ++i) {
      if (!due[i]) continue;
      Lane& L = lanes[i];
      const Measure& current = cur_m[i];
      const Measure& average = avg_m[i];
      // Mixed precision: from FP32 to FP64 steps once FP32 got close to the tolerance, or once
      // FP32 stopped improving near its rounding floor (which lies above the switch point): no
      // 20% gain on its best distance to termination within max(4096, iter / 4) iterations,
      // counted from the point where the distance has stayed below fp32_stall_level for
      // kStallArm evaluations in a row (a plateau above it belongs to the algorithm, which FP64
      // steps would only make slower; one evaluation below it can be the relative gap passing
      // through zero, seen at iteration 447 of the 1M-column LP in count order).
      if (L.fp32) {
        constexpr int kStallArm = 4;
        double w = std::min(worst(current), worst(average));
        if (halpern) {
          // T(z)'s distance at FP32's rounding floor scatters over an order of magnitude from one
          // evaluation to the next (measured on a 3-row LP: 5e-6 to 1e-4), which would keep the
          // stall rule from arming: the rule sees the least of the last 4 evaluations.
          L.fp32_recent[L.fp32_seen++ % 4] = w;
          w = std::min(std::min(L.fp32_recent[0], L.fp32_recent[1]),
                       std::min(L.fp32_recent[2], L.fp32_recent[3]));
        }
        const double level = opt.fp32_stall_level;
        L.fp32_below = w < level ? L.fp32_below + 1 : 0;
        const bool arm = !L.fp32_armed && L.fp32_below >= kStallArm;
        if (arm) L.fp32_armed = true;
        if (w < 0.8 * L.fp32_best || arm) {
          L.fp32_best = w;
          L.fp32_best_iter = iter;
        }
        const bool reached = w < opt.fp32_switch_tol;
        const bool stalled = L.fp32_armed &&
                             iter - L.fp32_best_iter >= std::max<long long>(4096, iter / 4);
        if (reached || stalled) {
          L.log(2, reached ? ": FP32 steps reached %.2e at iter %lld; FP64 steps from here"
                           : ": FP32 steps stalled at %.2e at iter %lld; FP64 steps from here",
                w, iter);
          precision[i] = SMY_FP64;
          L.fp32 = false;
          force[i] = 1;
        }
      }

      if (halpern) {
        // Restarted Halpern steps: the same three tests on the fixed-point residual
        // r = ||z - T(z)|| (primal part weighted by omega, dual part by 1 / omega) of the run's
        // last step instead of the KKT error; a restart moves to T(z), the new anchor.
        if (fixed.empty()) device->fixed_point(fixed);
        const double r = std::sqrt(L.omega * fixed[2 * i] + fixed[2 * i + 1] / L.omega);
        if (!(L.kkt_at_restart < kInf)) L.kkt_at_restart = r;
        const bool restart =
            force[i] || r <= 0.2 * L.kkt_at_restart ||
            (r <= 0.8 * L.kkt_at_restart && r > L.kkt_previous) ||
            static_cast<double>(L.k_since) >= 0.36 * static_cast<double>(iter + 1);
        L.kkt_previous = r;
        if (restart) {
          which[i] = SMY_AVERAGE;
          kkt_candidate[i] = r;
        }
        continue;
      }
      // Adaptive restart (src/lp/pdlp.cpp:1075-1132): the decision; the primal weight follows
      // from the distance below. With the device rule, both happen on the device (below).
      if (device_rule) continue;
      const double kkt_current = L.kkt(current);
      const double kkt_average = L.kkt(average);
      const bool to_average = kkt_average < kkt_current;
      const double cand = std::min(kkt_current, kkt_average);
      if (!(L.kkt_at_restart < kInf)) L.kkt_at_restart = cand;
      const bool restart =
          force[i] || cand <= 0.2 * L.kkt_at_restart ||
          (cand <= 0.8 * L.kkt_at_restart && cand > L.kkt_previous) ||
          static_cast<double>(L.k_since) >= 0.36 * static_cast<double>(iter + 1);
      L.kkt_previous = cand;
      if (restart) {
        which[i] = to_average ? SMY_AVERAGE : SMY_CURRENT;
        kkt_candidate[i] = cand;
      }
    }
    if (std::any_of(precision.begin(), precision.end(), [](int q) { return q >= 0; }))
      device->set_precision(precision);
    if (device_rule && std::any_of(due.begin(), due.end(), [](char d) { return d != 0; })) {
      for (int i = 0; i < K; ++i)
        rule_in[i] = smy_pdhg_rule_in{due[i] ? 1 : 0, force[i] ? 1 : 0, lanes[i].k_since};
      device->rule_decide(rule_in, iter + 1, rule_out);
      for (int i = 0; i < K; ++i) {
        if (!due[i] || rule_out[i].which < 0) continue;
        Lane& L = lanes[i];
        L.omega = rule_out[i].omega;
        L.tau = L.eta / L.omega;
        L.sigma = L.eta * L.omega;
        L.adapt.omega = L.omega;
        L.adapt.weight = 0.0;
        const bool to_average = rule_out[i].which == SMY_AVERAGE;
        L.log(2, ": restart at iter %lld to the %s (KKT %.2e), omega %.3e", iter,
              to_average ? "average" : "current iterate", rule_out[i].kkt_candidate, L.omega);
        L.res.restarts.push_back({iter, to_average ? 1 : 0, rule_out[i].kkt_candidate, L.omega});
        L.k_since = 0;
      }
    }
    if (std::any_of(which.begin(), which.end(), [](int w) { return w >= 0; })) {
      device->restart_distance(which, dist);
      for (int i = 0; i < K; ++i) {
        if (which[i] < 0) continue;
        Lane& L = lanes[i];
        const double dx = std::sqrt(dist[2 * i]);
        const double dy = std::sqrt(dist[2 * i + 1]);
        if (dx > kWeightDistanceFloor && dy > kWeightDistanceFloor) {
          if (opt.pid_primal_weight) {
            // log omega follows log(dy / dx) through a PI controller (gains 0.99 and 0.01).
            const double err = std::log(dy / dx) - std::log(L.omega);
            L.weight_integral += err;
            L.omega = std::exp(std::log(L.omega) + 0.99 * err + 0.01 * L.weight_integral);
          } else {
            L.omega = std::exp(0.5 * std::log(dy / dx) + 0.5 * std::log(L.omega));
          }
          L.tau = L.eta / L.omega;
          L.sigma = L.eta * L.omega;
        }
        L.adapt.omega = L.omega;
        L.adapt.weight = 0.0;  // the average starts again from the next accepted iterate
        const bool to_average = which[i] == SMY_AVERAGE;
        L.log(2, ": restart at iter %lld to the %s (KKT %.2e), omega %.3e", iter,
              to_average ? "average" : "current iterate", kkt_candidate[i], L.omega);
        L.res.restarts.push_back({iter, to_average ? 1 : 0, kkt_candidate[i], L.omega});
        L.k_since = 0;
        L.kkt_at_restart = kkt_candidate[i];
        L.kkt_previous = kInf;
      }
      device->restart_to(which);
    }
    device->advance();
    ++iter;
    for (int i = 0; i < K; ++i) {
      Lane& L = lanes[i];
      if (!L.active || iter < L.max_iter) continue;
      L.res.status = Status::kIterationLimit;
      L.res.iterations = L.max_iter;
      L.log(1, ": iteration limit reached after %lld iterations", L.max_iter);
      stop_at_current(i);
    }
  }
  std::vector<PdlpResult> out;
  out.reserve(static_cast<size_t>(K));
  for (Lane& L : lanes) {
    L.res.omega = L.omega;
    out.push_back(std::move(L.res));
  }
  if (info != nullptr) info->solve_seconds = clock.seconds();
  return out;
}

std::vector<PdlpResult> solve_lanes(const std::vector<PdlpTask>& tasks, const PdlpOptions& options,
                                    PdlpBatchInfo* info) {
  const Clock clock;
  const int K = static_cast<int>(tasks.size());
  try {
    std::vector<Lane> lanes;
    lanes.reserve(static_cast<size_t>(K));
    for (int i = 0; i < K; ++i) {
      lanes.emplace_back(*tasks[i].lp, options, K == 1 ? std::string("pdlp") : "pdlp[" + std::to_string(i) + "]");
      lanes.back().start_x = tasks[i].start_x;
      lanes.back().start_y = tasks[i].start_y;
      lanes.back().lower_ = tasks[i].lower;
      lanes.back().upper_ = tasks[i].upper;
    }
    return run_lanes(lanes, options, clock, info);
  } catch (const std::exception& e) {
    // The reference re-solves on the CPU after a GPU failure (src/lp/pdlp.cpp:228-246); this
    // library reports it.
    std::vector<PdlpResult> out(static_cast<size_t>(K));
    for (PdlpResult& r : out) {
      r.status = Status::kNumericalError;
      r.message = std::string("the GPU failed during the solve: ") + e.what();
      r.solve_seconds = clock.seconds();
    }
    if (options.log_level >= 1) {
      std::printf("pdlp: the GPU failed during the solve (%s)\n", e.what());
      std::fflush(stdout);
    }
    return out;
  }
}

}  // namespace

PdlpResult solve_pdlp(const LpProblem& lp, const PdlpOptions& options) {
  return solve_lanes({PdlpTask{&lp, nullptr, nullptr}}, options, nullptr)[0];
}

std::vector<PdlpResult> solve_pdlp_batch(const std::vector<PdlpTask>& tasks,
                                         const PdlpOptions& options, PdlpBatchInfo* info) {
  return solve_lanes(tasks, options, info);
}

}  // namespace smy::host
