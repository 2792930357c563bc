// The restart rule of PDLP on the device (the reference takes it on the host from downloaded sums:
// src/lp/pdlp.cpp:1075-1132, with the measures of :493-533).
//
// At an evaluation point, from each instance's 14 evaluation totals (current and average
// iterate), rule_decide_kernel forms the relative residuals and gap, the weighted KKT error of
// both candidates, and the reference's decision: restart to the better candidate when its error
// fell to 0.2 of the error at the last restart, or to 0.8 of it and stopped falling, or when the
// restart has lasted 36% of all iterations (or when the host forces one: a change of precision).
// The decision goes into the per-precision flags that the distance and restart kernels read, so
// distances, the primal-weight update (rule_weight_kernel) and the restart itself follow on the
// device without a round trip through the host. The host reads the outcome afterwards.
#pragma once

#include "hip_common.hpp"

namespace smy {

struct RuleState {
  // Normalizers of the relative residuals (src/lp/pdlp.cpp:493-519).
  double a_colmax, at_rowmax, c_norm;
  int has_rows;
  int which;  // the last decision: -1 none, 0 current, 1 average
  double omega;
  double kkt_at_restart, kkt_previous;
  double kkt_candidate;  // the last decision's candidate error
  long long since;       // iterations since the last restart
};

struct RuleInput {
  int due;    // take the decision for this instance now
  int force;  // restart regardless (the host changed its precision)
  long long since;  // iterations since the last restart (kept by the host, which counts steps)
};

// rel primal, rel dual, rel gap of 7 totals {||A x - r||^2, dual residual^2, primal objective,
// dual objective, ||x||^2, ||r||^2, ||y||^2}.
__device__ inline void rule_measure(const RuleState& s, const double* t, double out[3]) {
  const double primal_scale = 1.0 + s.a_colmax * sqrt(t[4]) + sqrt(t[5]);
  const double dual_scale = 1.0 + s.at_rowmax * sqrt(t[6]) + s.c_norm;
  const double gap_scale = 1.0 + fabs(t[2]) + fabs(t[3]);
  out[0] = s.has_rows ? sqrt(t[0]) / primal_scale : 0.0;
  out[1] = s.has_rows ? sqrt(t[1]) / dual_scale : 0.0;
  out[2] = (t[2] - t[3]) / gap_scale;
}

__device__ inline double rule_kkt(double omega, const double m[3]) {
  return sqrt(omega * omega * m[0] * m[0] + m[1] * m[1] / (omega * omega) + m[2] * m[2]);
}

// One thread per instance. sums: instance i's totals start at i * 28 and come in 2 groups of 14
// (the A part, then the A' part), added in that order. flags: 2 rows of ld_flags (>= K; the
// entries past K are padding), by precision.
static __global__ void rule_decide_kernel(int K, int ld_flags, const double* __restrict__ sums,
                                          long long iter1, const RuleInput* __restrict__ in,
                                          const int* __restrict__ prec, RuleState* __restrict__ st,
                                          int* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ld_flags) return;
  flags[i] = -1;
  flags[ld_flags + i] = -1;
  if (i >= K) return;
  RuleState s = st[i];
  s.which = -1;
  if (!in[i].due || prec[i] < 0) {
    st[i] = s;
    return;
  }
  double t[14];
  for (int q = 0; q < 14; ++q) t[q] = sums[static_cast<size_t>(i) * 28 + q] + sums[static_cast<size_t>(i) * 28 + 14 + q];
  double cur[3], avg[3];
  rule_measure(s, t, cur);
  rule_measure(s, t + 7, avg);
  const double kc = rule_kkt(s.omega, cur);
  const double ka = rule_kkt(s.omega, avg);
  const bool to_average = ka < kc;
  const double cand = fmin(kc, ka);
  if (!(s.kkt_at_restart < INFINITY)) s.kkt_at_restart = cand;
  s.since = in[i].since;
  const bool restart = in[i].force || cand <= 0.2 * s.kkt_at_restart ||
                       (cand <= 0.8 * s.kkt_at_restart && cand > s.kkt_previous) ||
                       static_cast<double>(s.since) >= 0.36 * static_cast<double>(iter1);
  s.kkt_previous = cand;
  s.kkt_candidate = cand;
  if (restart) {
    s.which = to_average ? 1 : 0;
    flags[prec[i] * ld_flags + i] = s.which;
  }
  st[i] = s;
}

// After the distances of the restarting instances (dist: instance i's at i * 4, `segments` groups
// of 2 totals, ||dv||^2 then ||dy||^2, added in order): the primal weight moves halfway, in log
// scale, to the ratio of the distances when both exceed 1e-10 (src/lp/pdlp.cpp:1095-1113); the
// restart resets the error reference and the iteration count.
static __global__ void rule_weight_kernel(int K, int segments, const double* __restrict__ dist,
                                          RuleState* __restrict__ st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  RuleState s = st[i];
  if (s.which < 0) return;
  double d[2];
  for (int q = 0; q < 2; ++q) {
    double v = dist[static_cast<size_t>(i) * 4 + q];
    for (int g = 1; g < segments; ++g) v = v + dist[static_cast<size_t>(i) * 4 + 2 * g + q];
    d[q] = v;
  }
  const double dx = sqrt(d[0]);
  const double dy = sqrt(d[1]);
  if (dx > 1e-10 && dy > 1e-10) s.omega = exp(0.5 * log(dy / dx) + 0.5 * log(s.omega));
  s.since = 0;
  s.kkt_at_restart = s.kkt_candidate;
  s.kkt_previous = INFINITY;
  st[i] = s;
}

}  // namespace smy
