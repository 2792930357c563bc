// Diagonal preconditioning on the device (the enhanced algorithm's infinity-norm Ruiz passes and
// Pock-Chambolle step, the norms of the termination test, and the power iteration that bounds
// ||[A -I]||), over the two CSR copies a handle holds: A by rows and A' by rows.
//
// The per-row and per-column reductions take one thread per row of a copy and walk its entries
// in index order: a maximum is exact in any order, and a sum formed that way is the host's
// (csrc/host/pdlp.cpp, ruiz_pock_chambolle and equilibrate_problem), so the factors equal the
// host's bit for bit. Every product is formed as the host forms it. The power iteration's
// products do the same, and its norms are taken on the host from the downloaded vectors, so the
// bound on the operator norm is the host's too.
#pragma once

#include "hip_common.hpp"

namespace smy {

// Per-row reductions of a CSR copy take one wavefront per row, for rows of any length (the
// 1M-column LP has columns of 5,000 entries, which one thread would walk for milliseconds): the
// 64 lanes form the terms of 64 consecutive entries at once, then every lane adds them in entry
// order from the lanes' registers (v_readlane), so the total is the sequential one, bit for bit.
// The power iteration's products are formed with the host's scaled values: A_eq' entries
// (a * cs_j) * rs_i, A_eq entries (a * rs_i) * cs_j.
__device__ inline double lane_value(double v, int lane) {
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  return __hiloint2double(hi, lo);
}

struct ScaledAbs {  // |(a * rs_i) * cs_j|
  const int* __restrict__ idx;
  const double* __restrict__ val;
  const double* __restrict__ rs;
  const double* __restrict__ cs;
  bool seg_is_row;
  __device__ double operator()(int s, int p) const {
    const int i = seg_is_row ? s : idx[p];
    const int j = seg_is_row ? idx[p] : s;
    return fabs(val[p] * rs[i] * cs[j]);
  }
};

struct ScaledSquare {  // ((a * rs_i) * cs_j)^2: the reference's 2-norm equilibration terms
  const int* __restrict__ idx;
  const double* __restrict__ val;
  const double* __restrict__ rs;
  const double* __restrict__ cs;
  bool seg_is_row;
  __device__ double operator()(int s, int p) const {
    const int i = seg_is_row ? s : idx[p];
    const int j = seg_is_row ? idx[p] : s;
    const double t = val[p] * rs[i] * cs[j];
    return t * t;
  }
};

struct Square {  // a^2
  const double* __restrict__ val;
  __device__ double operator()(int, int p) const { return val[p] * val[p]; }
};

struct PowerW {  // ((a * cs_j) * rs_i) * u_i over column j
  const int* __restrict__ idx;
  const double* __restrict__ val;
  const double* __restrict__ rs;
  const double* __restrict__ cs;
  const double* __restrict__ u;
  __device__ double operator()(int j, int p) const {
    const int i = idx[p];
    return val[p] * cs[j] * rs[i] * u[i];
  }
};

struct PowerU {  // ((a * rs_i) * cs_j) * w_j over row i; w_j = 0 adds nothing (the host skips it)
  const int* __restrict__ idx;
  const double* __restrict__ val;
  const double* __restrict__ rs;
  const double* __restrict__ cs;
  const double* __restrict__ w;
  __device__ double operator()(int i, int p) const {
    const int j = idx[p];
    const double wj = w[j];
    return wj == 0.0 ? 0.0 : val[p] * rs[i] * cs[j] * wj;
  }
};

// out[s] = sum (Max: max) of term(s, p) over row s's entries, in entry order.
template <typename Term, bool Max>
__global__ __launch_bounds__(kBlock) void segment_wave_kernel(int segs, const int* __restrict__ ptr,
                                                              Term term, double* __restrict__ out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int s = blockIdx.x * (kBlock / kWave) + static_cast<int>(threadIdx.x) / kWave;
  if (s >= segs) return;  // uniform over the wavefront
  const int b = ptr[s], e = ptr[s + 1];
  double acc = 0.0;
  for (int c = b; c < e; c += kWave) {
    const int p = c + lane;
    const double t = p < e ? term(s, p) : 0.0;
    const int k = min(kWave, e - c);
    for (int q = 0; q < k; ++q) {
      const double v = lane_value(t, q);
      if constexpr (Max)
        acc = fmax(acc, v);
      else
        acc += v;
    }
  }
  if (lane == 0) out[s] = acc;
}

// In place, one wavefront per row of the copy: A by rows (a * cs_j) * rs_i, A' by rows
// (a * rs_i) * cs_j.
template <bool SegIsRow>
__global__ __launch_bounds__(kBlock) void scale_values_wave_kernel(int segs, const int* __restrict__ ptr,
                                                                   const int* __restrict__ idx,
                                                                   double* __restrict__ val,
                                                                   const double* __restrict__ rs,
                                                                   const double* __restrict__ cs) {
  const int lane = threadIdx.x & (kWave - 1);
  const int s = blockIdx.x * (kBlock / kWave) + static_cast<int>(threadIdx.x) / kWave;
  if (s >= segs) return;
  const int e = ptr[s + 1];
  for (int p = ptr[s] + lane; p < e; p += kWave) {
    if constexpr (SegIsRow)
      val[p] = val[p] * cs[idx[p]] * rs[s];
    else
      val[p] = val[p] * rs[idx[p]] * cs[s];
  }
}

inline int wave_blocks(long long segs) {
  constexpr int per = kBlock / kWave;
  return static_cast<int>((segs + per - 1) / per);
}

// f[k] /= sqrt(norm[k]) where norm[k] > 0 (the factors of empty rows and columns stay).
__global__ __launch_bounds__(kBlock) void divide_by_sqrt_kernel(int count, double* __restrict__ f,
                                                                const double* __restrict__ norm) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < count && norm[k] > 0.0) f[k] /= sqrt(norm[k]);
}

// One update of the reference's 2-norm equilibration (src/lp/pdlp.cpp:131-178): where norm[k] > 0,
// f[k] = clamp(f[k] / norm[k], 1e-8, 1e8); *moved = 1 if some factor changed by more than 10%
// relative to max(|f[k]|, 1) (every thread that sees one stores the same value).
__global__ __launch_bounds__(kBlock) void ruiz2_update_kernel(int count, double* __restrict__ f,
                                                              const double* __restrict__ norm,
                                                              int* __restrict__ moved) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= count) return;
  const double nk = norm[k];
  if (!(nk > 0.0)) return;
  const double fk = f[k];
  const double next = fmax(1e-8, fmin(1e8, fk / nk));
  if (fabs(next - fk) / fmax(fabs(fk), 1.0) > 0.1) *moved = 1;
  f[k] = next;
}

// The scaled cost and bounds: structural j: cost * cs, bounds / cs; logical n + i: all * rs.
__global__ __launch_bounds__(kBlock) void scale_vectors_kernel(int n, int m, const double* __restrict__ cs,
                                                               const double* __restrict__ rs,
                                                               double* __restrict__ cost,
                                                               double* __restrict__ lower,
                                                               double* __restrict__ upper) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < n) {
    cost[k] *= cs[k];
    lower[k] /= cs[k];
    upper[k] /= cs[k];
  } else if (k < n + m) {
    const double r = rs[k - n];
    cost[k] *= r;
    lower[k] *= r;
    upper[k] *= r;
  }
}

constexpr int kReduceThreads = 1024;

// One workgroup: out[0] = sum (Max = false) or max (Max = true) of in[0..count), in a fixed order.
template <bool Max>
__global__ __launch_bounds__(kReduceThreads) void reduce_finish_kernel(long long count,
                                                                       const double* __restrict__ in,
                                                                       double* __restrict__ out) {
  __shared__ double part[kReduceThreads];
  double acc = 0.0;
  for (long long k = threadIdx.x; k < count; k += kReduceThreads) acc = Max ? fmax(acc, in[k]) : acc + in[k];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kReduceThreads / 2; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off)
      part[threadIdx.x] = Max ? fmax(part[threadIdx.x], part[threadIdx.x + off])
                              : part[threadIdx.x] + part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

// x[k] = sqrt(x[k]).
__global__ __launch_bounds__(kBlock) void sqrt_kernel(int count, double* __restrict__ x) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < count) x[k] = sqrt(x[k]);
}

}  // namespace smy
