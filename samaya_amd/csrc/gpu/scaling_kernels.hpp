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

enum SegmentNorm { kSegMax = 0, kSegSum = 1, kSegSumSq = 2 };

// out[s] over the entries of row s of a CSR copy: max or sum of |(a * rs_i) * cs_j|, or the sum
// of a^2 (the unscaled 2-norm's square). SegIsRow: row s of the copy is row i of A (indices are
// columns j); otherwise it is column j of A (indices are rows i).
template <int Kind, bool SegIsRow>
__global__ __launch_bounds__(kBlock) void segment_norm_kernel(int segs, const int* __restrict__ ptr,
                                                              const int* __restrict__ idx,
                                                              const double* __restrict__ val,
                                                              const double* __restrict__ rs,
                                                              const double* __restrict__ cs,
                                                              double* __restrict__ out) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= segs) return;
  double acc = 0.0;
  const int end = ptr[s + 1];
  for (int p = ptr[s]; p < end; ++p) {
    if constexpr (Kind == kSegSumSq) {
      const double a = val[p];
      acc += a * a;
    } else {
      const int i = SegIsRow ? s : idx[p];
      const int j = SegIsRow ? idx[p] : s;
      const double t = fabs(val[p] * rs[i] * cs[j]);
      acc = Kind == kSegMax ? fmax(acc, t) : acc + t;
    }
  }
  out[s] = acc;
}

// f[k] /= sqrt(norm[k]) where norm[k] > 0 (the factors of empty rows and columns stay).
__global__ __launch_bounds__(kBlock) void divide_by_sqrt_kernel(int count, double* __restrict__ f,
                                                                const double* __restrict__ norm) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < count && norm[k] > 0.0) f[k] /= sqrt(norm[k]);
}

// In place: A by rows gets (a * cs_j) * rs_i (the host scales A' that way and uploads it as A by
// rows), A' by rows gets (a * rs_i) * cs_j.
template <bool SegIsRow>
__global__ __launch_bounds__(kBlock) void scale_values_kernel(int segs, const int* __restrict__ ptr,
                                                              const int* __restrict__ idx,
                                                              double* __restrict__ val,
                                                              const double* __restrict__ rs,
                                                              const double* __restrict__ cs) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= segs) return;
  const int end = ptr[s + 1];
  for (int p = ptr[s]; p < end; ++p) {
    if constexpr (SegIsRow)
      val[p] = val[p] * cs[idx[p]] * rs[s];
    else
      val[p] = val[p] * rs[idx[p]] * cs[s];
  }
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

// The power iteration's products on the unscaled values and the factors, one thread per row
// summing in the order the host's scatter adds (csrc/host/pdlp.cpp, operator_norm_bound), with the
// host's scaled values: w = A_eq' u with A_eq' entries (a * cs_j) * rs_i, and u = A_eq w with
// A_eq entries (a * rs_i) * cs_j, skipping w_j = 0 as the host does.
__global__ __launch_bounds__(kBlock) void power_w_kernel(int n, const int* __restrict__ ptr,
                                                         const int* __restrict__ idx,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ rs,
                                                         const double* __restrict__ cs,
                                                         const double* __restrict__ u,
                                                         double* __restrict__ w) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  double acc = 0.0;
  const double c = cs[j];
  for (int p = ptr[j]; p < ptr[j + 1]; ++p) {
    const int i = idx[p];
    acc += val[p] * c * rs[i] * u[i];
  }
  w[j] = acc;
}

__global__ __launch_bounds__(kBlock) void power_u_kernel(int m, const int* __restrict__ ptr,
                                                         const int* __restrict__ idx,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ rs,
                                                         const double* __restrict__ cs,
                                                         const double* __restrict__ w,
                                                         double* __restrict__ u) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m) return;
  double acc = 0.0;
  const double r = rs[i];
  for (int p = ptr[i]; p < ptr[i + 1]; ++p) {
    const int j = idx[p];
    const double wj = w[j];
    if (wj == 0.0) continue;
    acc += val[p] * r * cs[j] * wj;
  }
  u[i] = acc;
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
