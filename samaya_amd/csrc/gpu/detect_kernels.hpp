// Device-side infeasibility and unboundedness tests of PDLP (src/lp/pdlp.cpp:715-1043) for both
// batch layouts, so that no iterate leaves the GPU for them.
//
// Farkas test on the average dual y_erg (pdlp.cpp:724-842), in the problem space before
// equilibration: yf = rs y_erg, w_x = A' yf, w_r = -yf. Phase 1 gives per instance
// sum |w| and max |yf|; phase 2, with drop = 1e-9 sum |w| and y_max, gives the finite parts of
// f_min = sum over c = w (|w| > drop) of c x (the bound that minimizes), f_max likewise, the counts
// of infinite contributions, and scale = max |t bound| over |t| > 1e-9, t = +-w / y_max.
// Ray test on d = v_new - v (pdlp.cpp:852-1042): phase 1 gives d_max; phase 2, with d normalized
// by it, gives ||A x - r||^2 (x = cs d_x, r = d_r / rs), the cost slope, the projected ray of the
// structural variables with its bound failures, and the bound failures of A ray on the rows.
// The host takes the decisions from these numbers exactly as the reference does.
//
// Slots of the per-instance results (block partials, then per-instance totals):
//   0 sum |w|   1 max |yf|   2 f_min (finite part)   3 f_max (finite part)   4 f_min -inf count
//   5 f_max +inf count   6 scale   7 d_max   8 ||A x - r||^2   9 cost slope   10 bound failures
// Slots 1, 6 and 7 are maxima, the others sums.
#pragma once

#include "csr_kernels.hpp"

namespace smy {

constexpr int kDetectSlots = 11;
constexpr unsigned kDetectMaxSlots = (1u << 1) | (1u << 6) | (1u << 7);

// One instance's data in either layout: stacked (global indices, one value per variable) or
// shared (one matrix for all instances; per-instance values at index * kp + instance, cost and
// row bounds shared).
template <typename T, bool Shared>
struct DetectView {
  int n;   // structural variables: v_{n+i} is row i's logical variable
  int kp;  // instances per variable row (shared layout)
  const T* __restrict__ v;
  const T* __restrict__ v_new;
  const T* __restrict__ y_erg;
  const T* __restrict__ y_rst;           // the restart point's duals (flags 5-7)
  const double* __restrict__ rs;
  const double* __restrict__ cs;
  const double* __restrict__ cost;       // per variable (shared by the instances when Shared)
  const double* __restrict__ lower;      // stacked: per variable; shared: structural n x kp
  const double* __restrict__ upper;
  const double* __restrict__ row_lower;  // shared layout: the logical bounds (m)
  const double* __restrict__ row_upper;
  double* __restrict__ scratch;          // structural: w, then ray / cs
  double* __restrict__ ray;              // structural: the projected ray
  const double* __restrict__ prm;        // per instance: drop, y_max, d_max
  const int* __restrict__ flags;         // per instance: > 0 when the tests run
  double* __restrict__ partial;          // slot base of the op
  size_t stride;                         // between slots
  int inst_stride;                       // between instances within a slot (shared layout)

  __device__ size_t at(int k, int inst) const {
    return Shared ? static_cast<size_t>(k) * kp + inst : static_cast<size_t>(k);
  }
  __device__ double lo(int k, int inst) const {
    if constexpr (Shared) return k < n ? lower[static_cast<size_t>(k) * kp + inst] : row_lower[k - n];
    return lower[k];
  }
  __device__ double up(int k, int inst) const {
    if constexpr (Shared) return k < n ? upper[static_cast<size_t>(k) * kp + inst] : row_upper[k - n];
    return upper[k];
  }
  __device__ double d(int k, int inst) const {
    return static_cast<double>(v_new[at(k, inst)]) - static_cast<double>(v[at(k, inst)]);
  }
  // Products of instances that do not take part are not computed at all.
  __device__ bool skip(int inst) const { return flags[inst] <= 0; }
};

// Before phase 1 when some flag is above 1: the dual the Farkas test takes, into a copy (ydst, the
// layout of y_erg) that the phases then read for every flagged instance. Flag bit 2 (4): the dual's
// change since the restart point, y_erg - y_rst (on an infeasible LP the iterates drift along the
// certificate's direction and the difference sheds the point they started from), else y_erg.
// Flag & 3 = 1: that vector as it is; 2 or 3: its sign projection for the f_min (2) or f_max (3)
// test, the entry of a row whose term in that sum would be infinite (f_min: -yf_i > 0 against an
// infinite lower bound of the row's logical variable, or -yf_i < 0 against an infinite upper
// bound; f_max: the opposite signs) set to 0. The reference tests y_erg only (pdlp.cpp:724-842);
// every y that passes the test is a certificate.
template <typename T, bool Shared>
struct FarkasProjectOp : DetectView<T, Shared> {
  static constexpr int kSums = 0;
  static constexpr int kMaxFrom = 0;
  T* __restrict__ ydst;
  __device__ int flag(int inst) const { return this->flags[inst] > 0 ? this->flags[inst] : -1; }
  __device__ void row_a(int i, int k, int f, double*, int inst) const {
    const size_t o = this->at(i, inst);
    double y = static_cast<double>(this->y_erg[o]);
    if (f & 4) y -= static_cast<double>(this->y_rst[o]);
    const int orient = f & 3;
    bool cut = false;
    if (orient >= 2) {
      const double s = orient == 2 ? y : -y;  // as for f_min
      cut = (s < 0.0 && !(this->lo(k, inst) > -INFINITY)) || (s > 0.0 && !(this->up(k, inst) < INFINITY));
    }
    ydst[o] = cut ? T(0) : static_cast<T>(y);
  }
  __device__ void col(int, int, double*, int) const {}
};

// Phase 1, rows of A': w_j = (A' yf)_j = (A_eq' y_erg)_j / cs_j, kept in scratch; slot 0 += |w_j|.
template <typename T, bool Shared>
struct FarkasColsOp : DetectView<T, Shared> {
  using Acc = double;
  using Pre = NoPre;
  static constexpr int kVec = 1;
  static constexpr int kSums = 2;
  static constexpr int kMaxFrom = 1;
  __device__ Pre pre(int, int) const { return {}; }
  __device__ Acc gather(int, int c, int inst) const {
    return static_cast<double>(this->y_erg[this->at(c, inst)]);
  }
  __device__ void row(int j, int inst, const Acc* acc, const Pre&, double* s) const {
    if (this->flags[inst] <= 0) return;
    const double w = acc[0] / this->cs[j];
    this->scratch[this->at(j, inst)] = w;
    s[0] += fabs(w);
  }
};

// Phase 1, rows of A (logical variables): slot 0 += |yf_i|, slot 1 = max |yf_i|.
template <typename T, bool Shared>
struct FarkasRowsOp : DetectView<T, Shared> {
  static constexpr int kSums = 2;
  static constexpr int kMaxFrom = 1;
  __device__ int flag(int inst) const { return this->flags[inst] > 0 ? 0 : -1; }
  __device__ void row_a(int i, int, int, double* s, int inst) const {
    const double yf = static_cast<double>(this->y_erg[this->at(i, inst)]) * this->rs[i];
    s[0] += fabs(yf);
    s[1] = fmax(s[1], fabs(yf));
  }
  __device__ void col(int, int, double*, int) const {}
};

// Phase 2, both schedules: the bound sums of the Farkas test (slots 2..6).
template <typename T, bool Shared>
struct FarkasBoundsOp : DetectView<T, Shared> {
  static constexpr int kSums = 5;
  static constexpr int kMaxFrom = 4;
  __device__ int flag(int inst) const { return this->flags[inst] > 0 ? 0 : -1; }
  __device__ void add(int k, double w, int inst, double* s) const {
    const double drop = this->prm[3 * inst], y_max = this->prm[3 * inst + 1];
    const double lo = this->lo(k, inst), up = this->up(k, inst);
    const bool lf = lo > -INFINITY, uf = up < INFINITY;
    const double c = fabs(w) <= drop ? 0.0 : w;
    if (c > 0) {
      if (lf) s[0] += c * lo; else s[2] += 1.0;
      if (uf) s[1] += c * up; else s[3] += 1.0;
    } else if (c < 0) {
      if (uf) s[0] += c * up; else s[2] += 1.0;
      if (lf) s[1] += c * lo; else s[3] += 1.0;
    }
    const double t = (k < this->n ? w : -w) / y_max;
    if (fabs(t) > 1e-9) {
      if (lf) s[4] = fmax(s[4], fabs(t * lo));
      if (uf) s[4] = fmax(s[4], fabs(t * up));
    }
  }
  __device__ void row_a(int i, int k, int, double* s, int inst) const {
    const double yf = static_cast<double>(this->y_erg[this->at(i, inst)]) * this->rs[i];
    add(k, -yf, inst, s);
  }
  __device__ void col(int j, int, double* s, int inst) const { add(j, this->scratch[this->at(j, inst)], inst, s); }
};

// Phase 1, both schedules: slot 7 = max |v_new - v|.
template <typename T, bool Shared>
struct RayMaxOp : DetectView<T, Shared> {
  static constexpr int kSums = 1;
  static constexpr int kMaxFrom = 0;
  __device__ int flag(int inst) const { return this->flags[inst] > 0 ? 0 : -1; }
  __device__ void row_a(int, int k, int, double* s, int inst) const { s[0] = fmax(s[0], fabs(this->d(k, inst))); }
  __device__ void col(int j, int, double* s, int inst) const { s[0] = fmax(s[0], fabs(this->d(j, inst))); }
};

// Phase 2, rows of A: res_i = (A x)_i - r_i with x = cs d_x / d_max, r_i = d_{n+i} / (rs_i d_max):
// slot 8 += res_i^2, slot 9 += cost_{n+i} r_i.
template <typename T, bool Shared>
struct RayRowsOp : DetectView<T, Shared> {
  using Acc = double;
  using Pre = NoPre;
  static constexpr int kVec = 1;
  static constexpr int kSums = 2;
  static constexpr int kMaxFrom = 2;
  __device__ Pre pre(int, int) const { return {}; }
  __device__ Acc gather(int, int c, int inst) const { return this->d(c, inst); }
  __device__ void row(int i, int inst, const Acc* acc, const Pre&, double* s) const {
    if (this->flags[inst] <= 0) return;
    const double dmax = this->prm[3 * inst + 2];
    if (!(dmax > 1e-6)) return;
    const int k = this->n + i;
    const double dr = this->d(k, inst) / dmax;
    const double r = dr / this->rs[i];
    const double res = acc[0] / dmax / this->rs[i] - r;
    s[0] += res * res;
    s[1] += this->cost[k] * r;
  }
};

__device__ inline bool fixed_bounds(double lo, double up) { return up - lo <= 1e-10 * (1.0 + fabs(lo)); }

// Phase 2, rows of A' (structural): x_j = cs_j d_j / d_max, slot 9 += cost_j x_j, the projected
// ray (kept in `ray`, and ray / cs in scratch for the product below), slot 10 += bound failures.
template <typename T, bool Shared>
struct RayColsOp : DetectView<T, Shared> {
  static constexpr int kSums = 2;
  static constexpr int kMaxFrom = 2;
  __device__ int flag(int inst) const { return this->flags[inst] > 0 ? 0 : -1; }
  __device__ void row_a(int, int, int, double*, int) const {}
  __device__ void col(int j, int, double* s, int inst) const {
    const double dmax = this->prm[3 * inst + 2];
    double r = 0.0;
    if (dmax > 1e-6) {
      double x = this->d(j, inst) / dmax * this->cs[j];
      s[0] += this->cost[j] * x;
      const double lo = this->lo(j, inst), up = this->up(j, inst);
      const bool lb = lo > -INFINITY, ub = up < INFINITY;
      constexpr double kBoxTol = 1e-4, kFixedTol = 1e-3;
      if (lb && ub) {
        if (fabs(x) > (fixed_bounds(lo, up) ? kFixedTol : kBoxTol)) s[1] += 1.0;
        x = 0.0;
      } else if (lb) {
        if (x < -kBoxTol) s[1] += 1.0;
        x = fmax(x, 0.0);
      } else if (ub) {
        if (x > kBoxTol) s[1] += 1.0;
        x = fmin(x, 0.0);
      }
      r = x;
    }
    this->ray[this->at(j, inst)] = r;
    this->scratch[this->at(j, inst)] = r / this->cs[j];
  }
};

// Phase 2, rows of A: (A ray)_i = (A_eq (ray / cs))_i / rs_i against row i's bounds; slot 10 +=
// failures.
template <typename T, bool Shared>
struct RayCheckOp : DetectView<T, Shared> {
  using Acc = double;
  using Pre = NoPre;
  static constexpr int kVec = 1;
  static constexpr int kSums = 1;
  static constexpr int kMaxFrom = 1;
  __device__ Pre pre(int, int) const { return {}; }
  __device__ Acc gather(int, int c, int inst) const { return this->scratch[this->at(c, inst)]; }
  __device__ void row(int i, int inst, const Acc* acc, const Pre&, double* s) const {
    if (this->flags[inst] <= 0) return;
    if (!(this->prm[3 * inst + 2] > 1e-6)) return;
    const double av = acc[0] / this->rs[i];
    const int k = this->n + i;
    const double lo = this->lo(k, inst), up = this->up(k, inst);
    const bool lb = lo > -INFINITY, ub = up < INFINITY;
    constexpr double kBoxTol = 1e-4, kFixedTol = 1e-3;
    if (lb && ub) {
      if (fabs(av) > (fixed_bounds(lo, up) ? kFixedTol : kBoxTol)) s[0] += 1.0;
    } else if (lb) {
      if (av < -kBoxTol) s[0] += 1.0;
    } else if (ub) {
      if (av > kBoxTol) s[0] += 1.0;
    }
  }
};

// Per-instance totals of the detection slots: workgroup (slot, instance) combines the instance's
// two block ranges in a fixed order (sum, or max for the slots of kDetectMaxSlots) and writes
// out[instance * kDetectSlots + slot].
static __global__ __launch_bounds__(kBlock) void finish_detect_kernel(size_t stride,
                                                                       const double* __restrict__ partial,
                                                                       const int2* __restrict__ ranges,
                                                                       double* __restrict__ out) {
  __shared__ double wsum[kBlock / kWave];
  const int tid = threadIdx.x;
  const int q = blockIdx.x;
  const int inst = blockIdx.y;
  const bool is_max = (kDetectMaxSlots >> q) & 1u;
  const double* row = partial + static_cast<size_t>(q) * stride;
  double acc = 0.0;
  for (int part = 0; part < 2; ++part) {
    const int2 r = ranges[2 * inst + part];
    for (int b = r.x + tid; b < r.y; b += kBlock) acc = is_max ? fmax(acc, row[b]) : acc + row[b];
  }
  acc = is_max ? group_max<kWave>(acc) : group_sum<kWave>(acc);
  if ((tid & (kWave - 1)) == 0) wsum[tid / kWave] = acc;
  __syncthreads();
  if (tid == 0)
    out[static_cast<size_t>(inst) * kDetectSlots + q] =
        is_max ? fmax(fmax(wsum[0], wsum[1]), fmax(wsum[2], wsum[3]))
               : ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

}  // namespace smy
