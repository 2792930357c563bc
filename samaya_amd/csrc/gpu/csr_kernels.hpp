// Device kernels for the PDHG hot path on gfx950 (CDNA4, wavefront 64).
//
// Every product with A or A' runs through one kernel shape, csr_rows_kernel: a 256-thread
// workgroup takes one block of consecutive rows from the row-block schedule (row_blocks.hpp),
// streams the block's values and column indices once with coalesced non-temporal loads (they
// are read once per product and must not evict the gathered vector from L2 / Infinity Cache),
// multiplies by the gathered vector entries, stages the products in LDS, and sums each row with a
// group of g lanes (g = 64 ... 1 depending on how many rows the block holds) in a fixed butterfly
// order. A row too long for LDS is summed by the whole workgroup with a strided loop. The row
// totals then go through an epilogue functor (Op::row) that applies the PDHG update of that row's
// variable, or accumulates the evaluation sums, so no product ever makes a round trip through
// HBM. Sums are formed in an order fixed by the schedule: repeated runs are bit-identical.
#pragma once

#include "hip_common.hpp"
#include "row_blocks.hpp"

namespace smy {

template <typename V>
struct CsrRows {
  int rows = 0;
  const int* __restrict__ ptr = nullptr;
  const int* __restrict__ idx = nullptr;
  const V* __restrict__ val = nullptr;
  const int* __restrict__ blocks = nullptr;  // nblocks + 1 row starts
};

// Scalars of one PDHG step, read by the step kernels from device memory so that the same kernel
// nodes can be replayed from a HIP graph with new values. alpha = 1 / (k0 + j + 1) for the j-th
// step of a run whose first step is the k0-th since the last restart.
struct StepScalars {
  double tau;
  double sigma;
  double theta;
  long long k0;
};

template <typename T>
__device__ inline T clamp_to(T v, T lo, T up) {
  return fmin(fmax(v, lo), up);
}

// Sums of K per-thread values over the workgroup (wavefront butterfly, then the four wavefront
// totals in index order); thread q < K writes total q to partial[q * gridDim.x + blockIdx.x].
template <int K>
__device__ inline void block_partials(const double (&local)[K], double* __restrict__ partial) {
  __shared__ double wsum[K][kBlock / kWave];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const double t = group_sum<kWave>(local[q]);
    if ((tid & (kWave - 1)) == 0) wsum[q][tid / kWave] = t;
  }
  __syncthreads();
  if (tid < K) {
    partial[tid * gridDim.x + blockIdx.x] =
        ((wsum[tid][0] + wsum[tid][1]) + wsum[tid][2]) + wsum[tid][3];
  }
}

template <typename V, typename Op>
__global__ __launch_bounds__(kBlock) void csr_rows_kernel(CsrRows<V> A, Op op) {
  using Acc = typename Op::Acc;
  constexpr int NV = Op::kVec;
  __shared__ Acc prod[NV][kBlockNnz];
  __shared__ Acc rowsum[NV][kBlockRows];
  const int tid = threadIdx.x;
  const int r0 = A.blocks[blockIdx.x];
  const int r1 = A.blocks[blockIdx.x + 1];
  const int p0 = A.ptr[r0];
  const int p1 = A.ptr[r1];
  const int nrows = r1 - r0;
  const int nnz = p1 - p0;
  if (nnz > kBlockNnz) {
    // One row longer than the LDS stage: the whole workgroup strides over it.
    Acc s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = Acc(0);
    for (int p = p0 + tid; p < p1; p += kBlock) {
      const Acc a = static_cast<Acc>(__builtin_nontemporal_load(A.val + p));
      const int c = __builtin_nontemporal_load(A.idx + p);
#pragma unroll
      for (int v = 0; v < NV; ++v) s[v] += a * op.gather(v, c);
    }
    __shared__ Acc wsum[NV][kBlock / kWave];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const Acc t = group_sum<kWave>(s[v]);
      if ((tid & (kWave - 1)) == 0) wsum[v][tid / kWave] = t;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v)
        rowsum[v][0] = ((wsum[v][0] + wsum[v][1]) + wsum[v][2]) + wsum[v][3];
    }
  } else {
    for (int k = tid; k < nnz; k += kBlock) {
      const Acc a = static_cast<Acc>(__builtin_nontemporal_load(A.val + p0 + k));
      const int c = __builtin_nontemporal_load(A.idx + p0 + k);
#pragma unroll
      for (int v = 0; v < NV; ++v) prod[v][k] = a * op.gather(v, c);
    }
    __syncthreads();
    // Lanes per row: the largest power of two <= 64 that still gives every row a group.
    int g = kWave;
    while (g > 1 && g * nrows > kBlock) g >>= 1;
    const int grp = tid / g;
    const int lane = tid & (g - 1);
    Acc s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = Acc(0);
    if (grp < nrows) {
      const int b = A.ptr[r0 + grp] - p0;
      const int e = A.ptr[r0 + grp + 1] - p0;
      for (int q = b + lane; q < e; q += g) {
#pragma unroll
        for (int v = 0; v < NV; ++v) s[v] += prod[v][q];
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = group_sum_dyn(s[v], g);
    if (grp < nrows && lane == 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) rowsum[v][grp] = s[v];
    }
  }
  __syncthreads();
  double sums[Op::kSums > 0 ? Op::kSums : 1];
#pragma unroll
  for (int q = 0; q < (Op::kSums > 0 ? Op::kSums : 1); ++q) sums[q] = 0.0;
  if (tid < nrows) {
    Acc acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = rowsum[v][tid];
    op.row(r0 + tid, acc, sums);
  }
  if constexpr (Op::kSums > 0) block_partials<Op::kSums>(sums, op.partial);
}

// ------------------------------------------------------------------------------------------
// Epilogues.
// ------------------------------------------------------------------------------------------

// Plain product y = A x (the standalone SpMV entry point).
template <typename V, typename X>
struct SpmvOp {
  using Acc = X;
  static constexpr int kVec = 1;
  static constexpr int kSums = 0;
  const X* __restrict__ x;
  X* __restrict__ y;
  __device__ Acc gather(int, int c) const { return x[c]; }
  __device__ void row(int r, const Acc* acc, double*) const { y[r] = acc[0]; }
};

// Rows of A: the dual update of row i and the primal update of its logical variable n + i.
//   y_new_i = y_i + sigma (A vbar_x - vbar_{n+i})_i,  y_erg_i += alpha (y_new_i - y_erg_i)
//   v_new_{n+i} = clamp(v_{n+i} - tau (cost_{n+i} - y_new_i)),  its average and next vbar.
template <typename T>
struct DualRowsOp {
  using Acc = T;
  static constexpr int kVec = 1;
  static constexpr int kSums = 0;
  int n;
  int j;  // step index within the run (alpha)
  const StepScalars* __restrict__ sc;
  const T* __restrict__ vbar;
  const T* __restrict__ y;
  T* __restrict__ y_new;
  T* __restrict__ y_erg;
  const T* __restrict__ v;
  T* __restrict__ v_new;
  T* __restrict__ v_erg;
  T* __restrict__ vbar_next;
  const T* __restrict__ cost;
  const T* __restrict__ lower;
  const T* __restrict__ upper;
  __device__ Acc gather(int, int c) const { return vbar[c]; }
  __device__ void row(int i, const Acc* acc, double*) const {
    const T sigma = static_cast<T>(sc->sigma);
    const T tau = static_cast<T>(sc->tau);
    const T theta = static_cast<T>(sc->theta);
    const T alpha = static_cast<T>(1.0 / static_cast<double>(sc->k0 + j + 1));
    const int k = n + i;
    const T yn = y[i] + sigma * (acc[0] - vbar[k]);
    y_new[i] = yn;
    y_erg[i] = (T(1) - alpha) * y_erg[i] + alpha * yn;
    const T vk = v[k];
    const T grad = cost[k] + (-yn);
    const T vn = clamp_to(vk - tau * grad, lower[k], upper[k]);
    v_new[k] = vn;
    v_erg[k] = (T(1) - alpha) * v_erg[k] + alpha * vn;
    vbar_next[k] = vn + theta * (vn - vk);
  }
};

// Rows of A' (structural variables j < n): v_new_j = clamp(v_j - tau (cost_j + (A' y_new)_j)),
// its running average and the next extrapolation vbar_j = v_new_j + theta (v_new_j - v_j).
template <typename T>
struct PrimalColsOp {
  using Acc = T;
  static constexpr int kVec = 1;
  static constexpr int kSums = 0;
  int j;
  const StepScalars* __restrict__ sc;
  const T* __restrict__ y_new;
  const T* __restrict__ v;
  T* __restrict__ v_new;
  T* __restrict__ v_erg;
  T* __restrict__ vbar_next;
  const T* __restrict__ cost;
  const T* __restrict__ lower;
  const T* __restrict__ upper;
  __device__ Acc gather(int, int c) const { return y_new[c]; }
  __device__ void row(int col, const Acc* acc, double*) const {
    const T tau = static_cast<T>(sc->tau);
    const T theta = static_cast<T>(sc->theta);
    const T alpha = static_cast<T>(1.0 / static_cast<double>(sc->k0 + j + 1));
    const T vk = v[col];
    const T grad = cost[col] + acc[0];
    const T vn = clamp_to(vk - tau * grad, lower[col], upper[col]);
    v_new[col] = vn;
    v_erg[col] = (T(1) - alpha) * v_erg[col] + alpha * vn;
    vbar_next[col] = vn + theta * (vn - vk);
  }
};

// lambda: the part of a reduced cost g that the bounds of its variable can absorb (any sign when
// both bounds are finite, >= 0 with only a lower bound, <= 0 with only an upper one, 0 if free).
__device__ inline double absorbable(double g, double lo, double up) {
  const bool l = lo > -INFINITY;
  const bool u = up < INFINITY;
  return l && u ? g : l ? fmax(g, 0.0) : u ? fmin(g, 0.0) : 0.0;
}

// Evaluation sums (7 per iterate, in the model space before equilibration): 0 ||A x - r||^2,
// 1 dual residual^2, 2 primal objective, 3 dual objective, 4 ||x||^2, 5 ||r||^2, 6 ||y||^2.
// x = cs v_x, r = v_r / rs, y_o = rs y; A x = (A_eq v_x) / rs and A' y_o = (A_eq' y) / cs.
// Rows of A carry the logical variables' terms.
template <typename T, int NV>
struct MeasureRowsOp {
  using Acc = double;
  static constexpr int kVec = NV;
  static constexpr int kSums = 7 * NV;
  int n;
  const T* __restrict__ vv[NV];
  const T* __restrict__ yy[NV];
  const double* __restrict__ rs;
  const double* __restrict__ cost;
  const double* __restrict__ lower;
  const double* __restrict__ upper;
  double* __restrict__ partial;
  __device__ Acc gather(int v, int c) const { return static_cast<double>(vv[v][c]); }
  __device__ void row(int i, const Acc* acc, double* s) const {
    const int k = n + i;
    const double r = rs[i];
    const double c = cost[k], lo = lower[k], up = upper[k];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double value = static_cast<double>(vv[v][k]) / r;
      const double y_o = static_cast<double>(yy[v][i]) * r;
      const double g = c - y_o;
      const double e = acc[v] / r - value;
      s[7 * v + 0] += e * e;
      s[7 * v + 5] += value * value;
      s[7 * v + 6] += y_o * y_o;
      s[7 * v + 2] += c * value;
      const double lam = absorbable(g, lo, up);
      s[7 * v + 1] += (g - lam) * (g - lam);
      if (lam > 0.0) s[7 * v + 3] += lam * lo;
      if (lam < 0.0) s[7 * v + 3] += lam * up;
    }
  }
};

// Rows of A' carry the structural variables' terms.
template <typename T, int NV>
struct MeasureColsOp {
  using Acc = double;
  static constexpr int kVec = NV;
  static constexpr int kSums = 7 * NV;
  int m;
  const T* __restrict__ vv[NV];
  const T* __restrict__ yy[NV];
  const double* __restrict__ cs;
  const double* __restrict__ cost;
  const double* __restrict__ lower;
  const double* __restrict__ upper;
  double* __restrict__ partial;
  __device__ Acc gather(int v, int c) const { return static_cast<double>(yy[v][c]); }
  __device__ void row(int col, const Acc* acc, double* s) const {
    const double sc = cs[col];
    const double c = cost[col], lo = lower[col], up = upper[col];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double value = static_cast<double>(vv[v][col]) * sc;
      const double g = c + (m > 0 ? acc[v] / sc : 0.0);
      s[7 * v + 4] += value * value;
      s[7 * v + 2] += c * value;
      const double lam = absorbable(g, lo, up);
      s[7 * v + 1] += (g - lam) * (g - lam);
      if (lam > 0.0) s[7 * v + 3] += lam * lo;
      if (lam < 0.0) s[7 * v + 3] += lam * up;
    }
  }
};

// ------------------------------------------------------------------------------------------
// Elementwise and reduction kernels outside the products.
// ------------------------------------------------------------------------------------------

__global__ void set_scalars_kernel(StepScalars* sc, double tau, double sigma, double theta,
                                   long long k0) {
  if (threadIdx.x == 0) {
    sc->tau = tau;
    sc->sigma = sigma;
    sc->theta = theta;
    sc->k0 = k0;
  }
}

// vbar = v + theta (v - v_prev): only when theta changed since vbar was formed.
template <typename T>
__global__ void extrapolate_kernel(int nt, const T* __restrict__ v, const T* __restrict__ v_prev,
                                   T theta, T* __restrict__ vbar) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nt) vbar[k] = v[k] + theta * (v[k] - v_prev[k]);
}

// Squared distances ||v - v0||^2 (nt entries) and ||y - y0||^2 (m entries), block partials.
template <typename T>
__global__ __launch_bounds__(kBlock) void distance_kernel(int nt, int m, const T* __restrict__ v,
                                                          const T* __restrict__ v0,
                                                          const T* __restrict__ y,
                                                          const T* __restrict__ y0,
                                                          double* __restrict__ partial) {
  double local[2] = {0.0, 0.0};
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nt; k += gridDim.x * blockDim.x) {
    const double d = static_cast<double>(v[k]) - static_cast<double>(v0[k]);
    local[0] += d * d;
    if (k < m) {
      const double e = static_cast<double>(y[k]) - static_cast<double>(y0[k]);
      local[1] += e * e;
    }
  }
  block_partials<2>(local, partial);
}

// out[q] = sum over blocks of partial[q * nblocks + b], in a fixed order (one workgroup).
template <int K>
__global__ __launch_bounds__(kBlock) void finish_sums_kernel(int nblocks,
                                                             const double* __restrict__ partial,
                                                             double* __restrict__ out) {
  __shared__ double wsum[K][kBlock / kWave];
  const int tid = threadIdx.x;
  for (int q = 0; q < K; ++q) {
    double acc = 0.0;
    for (int b = tid; b < nblocks; b += kBlock) acc += partial[q * nblocks + b];
    acc = group_sum<kWave>(acc);
    if ((tid & (kWave - 1)) == 0) wsum[q][tid / kWave] = acc;
  }
  __syncthreads();
  if (tid < K) out[tid] = ((wsum[tid][0] + wsum[tid][1]) + wsum[tid][2]) + wsum[tid][3];
}

// A restart at an iterate (cv, cy): it becomes the restart point, the current and next iterate,
// and the extrapolation point (so the next step starts without momentum).
template <typename T>
__global__ void restart_to_kernel(int nt, int m, const T* __restrict__ cv,
                                   const T* __restrict__ cy, T* __restrict__ v_rst,
                                   T* __restrict__ v, T* __restrict__ v_new,
                                   T* __restrict__ vbar, T* __restrict__ y_rst,
                                   T* __restrict__ y, T* __restrict__ y_new) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nt) {
    const T c = cv[k];
    v_rst[k] = c;
    v[k] = c;
    v_new[k] = c;
    vbar[k] = c;
  }
  if (k < m) {
    const T c = cy[k];
    y_rst[k] = c;
    y[k] = c;
    y_new[k] = c;
  }
}

template <typename T, typename U>
__global__ void convert_kernel(long long count, const T* __restrict__ in, U* __restrict__ out) {
  const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) out[k] = static_cast<U>(in[k]);
}

template <typename T>
__global__ void copy_kernel(long long count, const T* __restrict__ in, T* __restrict__ out) {
  const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) out[k] = in[k];
}

}  // namespace smy
