// Device kernels for the PDHG hot path on gfx950 (CDNA4, wavefront 64).
//
// Every product with A or A' runs through one kernel shape, csr_rows_kernel: a persistent grid
// of 256-thread workgroups (as many as fit on the chip) walks the row-block schedule
// (row_blocks.hpp), block b, b + grid, b + 2 grid, ... For each block the workgroup
//   1. streams the block's values and column indices once, 4 per thread, with coalesced
//      non-temporal loads (read once per product, they must not evict the gathered vector from
//      L2 / Infinity Cache), together with the block's row offsets and, for each row, the inputs
//      of its epilogue (Op::pre);
//   2. gathers the vector entries, multiplies, and stages the products in LDS;
//   3. sums each row with a group of g lanes (g = 64 ... 1, from the block's row count) in a
//      fixed butterfly order, or, for a row longer than kLongRow that has a block of its own,
//      with a strided loop over the whole workgroup;
//   4. applies the epilogue functor (Op::row) to each row total: the PDHG update of that row's
//      variable, or the evaluation sums, so no product makes a round trip through HBM.
// The loads of step 1 for the next block are issued before steps 2-4 of the current one, and the
// schedule entry of the block after that is loaded one block further ahead, so each workgroup
// keeps two blocks' worth of memory traffic in flight instead of one dependent chain. All sums
// are formed in an order fixed by the schedule and the grid: repeated runs are bit-identical.
#pragma once

#include "hip_common.hpp"
#include "row_blocks.hpp"

namespace smy {

constexpr int kPer = kBlockNnz / kBlock;  // nonzeros each thread streams per block

template <typename V>
struct CsrRows {
  int rows = 0;
  int nblocks = 0;
  const int* __restrict__ ptr = nullptr;
  const int* __restrict__ idx = nullptr;
  const V* __restrict__ val = nullptr;
  const int2* __restrict__ meta = nullptr;  // nblocks + 1 entries (first row, first nonzero)
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

// What one thread holds of a block between issuing its loads and using them.
template <typename V, typename Pre>
struct BlockStage {
  int r0, p0, nrows, nnz;
  bool long_row;
  V a[kPer];
  int c[kPer];
  int rp;   // ptr[r0 + tid] - p0 for tid < nrows
  Pre pre;  // the epilogue inputs of row r0 + tid
};

template <typename V, typename Op>
__device__ inline void stage_block(const CsrRows<V>& A, const Op& op, int2 m0, int2 m1,
                                   BlockStage<V, typename Op::Pre>& s) {
  const int tid = threadIdx.x;
  s.r0 = m0.x;
  s.p0 = m0.y;
  s.nrows = m1.x - m0.x;
  s.nnz = m1.y - m0.y;
  s.long_row = s.nrows == 1 && s.nnz > kLongRow;
  if (!s.long_row) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int k = tid + u * kBlock;
      if (k < s.nnz) {
        s.a[u] = __builtin_nontemporal_load(A.val + s.p0 + k);
        s.c[u] = __builtin_nontemporal_load(A.idx + s.p0 + k);
      } else {
        s.a[u] = V(0);
        s.c[u] = 0;
      }
    }
  }
  if (tid < s.nrows) {
    s.rp = A.ptr[s.r0 + tid] - s.p0;
    s.pre = op.pre(s.r0 + tid);
  }
}

template <typename V, typename Op>
__global__ __launch_bounds__(kBlock) void csr_rows_kernel(CsrRows<V> A, Op op) {
  using Acc = typename Op::Acc;
  using Pre = typename Op::Pre;
  constexpr int NV = Op::kVec;
  constexpr int NS = Op::kSums > 0 ? Op::kSums : 1;
  __shared__ Acc prod[NV][kBlockNnz];
  __shared__ Acc rowsum[NV][kBlockRows];
  __shared__ Acc wsum[NV][kBlock / kWave];
  __shared__ int rp[kBlockRows + 1];
  const int tid = threadIdx.x;
  const int grid = gridDim.x;
  double sums[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) sums[q] = 0.0;

  int b = blockIdx.x;
  BlockStage<V, Pre> cur;
  if (b < A.nblocks) stage_block(A, op, A.meta[b], A.meta[b + 1], cur);
  int bn = b + grid;
  int2 n0 = make_int2(0, 0), n1 = make_int2(0, 0);
  if (bn < A.nblocks) {
    n0 = A.meta[bn];
    n1 = A.meta[bn + 1];
  }
  while (b < A.nblocks) {
    // Gathers of this block (their loads are the oldest in flight after the streams).
    Acc xs[NV][kPer];
    if (!cur.long_row) {
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const bool live = tid + u * kBlock < cur.nnz;
#pragma unroll
        for (int v = 0; v < NV; ++v) xs[v][u] = live ? op.gather(v, cur.c[u]) : Acc(0);
      }
    }
    // The next block's streams and epilogue inputs, and the schedule entry after it.
    BlockStage<V, Pre> nxt;
    const bool more = bn < A.nblocks;
    if (more) stage_block(A, op, n0, n1, nxt);
    const int bnn = bn + grid;
    int2 q0 = make_int2(0, 0), q1 = make_int2(0, 0);
    if (bnn < A.nblocks) {
      q0 = A.meta[bnn];
      q1 = A.meta[bnn + 1];
    }

    if (!cur.long_row) {
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int k = tid + u * kBlock;
        if (k < cur.nnz) {
#pragma unroll
          for (int v = 0; v < NV; ++v) prod[v][k] = static_cast<Acc>(cur.a[u]) * xs[v][u];
        }
      }
      if (tid < cur.nrows) rp[tid] = cur.rp;
      if (tid == 0) rp[cur.nrows] = cur.nnz;
      __syncthreads();
      // Lanes per row: the largest power of two <= 64 that still gives every row a group.
      int g = kWave;
      while (g > 1 && g * cur.nrows > kBlock) g >>= 1;
      const int grp = tid / g;
      const int lane = tid & (g - 1);
      Acc s[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) s[v] = Acc(0);
      if (grp < cur.nrows) {
        const int e = rp[grp + 1];
        for (int q = rp[grp] + lane; q < e; q += g) {
#pragma unroll
          for (int v = 0; v < NV; ++v) s[v] += prod[v][q];
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) s[v] = group_sum_dyn(s[v], g);
      if (grp < cur.nrows && lane == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) rowsum[v][grp] = s[v];
      }
    } else {
      // One long row: the whole workgroup strides over it, kPer entries per thread per round
      // with every load of a round issued before the first product.
      Acc s[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) s[v] = Acc(0);
      const int end = cur.p0 + cur.nnz;
      for (int p = cur.p0 + tid; p < end; p += kPer * kBlock) {
        V a[kPer];
        int c[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
          const int q = p + u * kBlock;
          a[u] = q < end ? __builtin_nontemporal_load(A.val + q) : V(0);
          c[u] = q < end ? __builtin_nontemporal_load(A.idx + q) : 0;
        }
        Acc g[NV][kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
#pragma unroll
          for (int v = 0; v < NV; ++v) g[v][u] = p + u * kBlock < end ? op.gather(v, c[u]) : Acc(0);
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
#pragma unroll
          for (int v = 0; v < NV; ++v) s[v] += static_cast<Acc>(a[u]) * g[v][u];
        }
      }
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
    }
    __syncthreads();
    if (tid < cur.nrows) {
      Acc acc[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] = rowsum[v][tid];
      op.row(cur.r0 + tid, acc, cur.pre, sums);
    }
    __syncthreads();  // every LDS read of this block precedes the next block's writes
    if (!more) break;
    cur = nxt;
    b = bn;
    bn = bnn;
    n0 = q0;
    n1 = q1;
  }
  if constexpr (Op::kSums > 0) block_partials<Op::kSums>(sums, op.partial);
}

// ------------------------------------------------------------------------------------------
// Epilogues.
// ------------------------------------------------------------------------------------------

struct NoPre {};

// Plain product y = A x (the standalone SpMV entry point).
template <typename V, typename X>
struct SpmvOp {
  using Acc = X;
  using Pre = NoPre;
  static constexpr int kVec = 1;
  static constexpr int kSums = 0;
  const X* __restrict__ x;
  X* __restrict__ y;
  __device__ Pre pre(int) const { return {}; }
  __device__ Acc gather(int, int c) const { return x[c]; }
  __device__ void row(int r, const Acc* acc, const Pre&, double*) const { y[r] = acc[0]; }
};

template <typename T>
struct DualPre {
  T y, y_erg, vbar, v, v_erg, cost, lower, upper;
};

// Rows of A: the dual update of row i and the primal update of its logical variable n + i.
//   y_new_i = y_i + sigma (A vbar_x - vbar_{n+i})_i,  y_erg_i += alpha (y_new_i - y_erg_i)
//   v_new_{n+i} = clamp(v_{n+i} - tau (cost_{n+i} - y_new_i)),  its average and next vbar.
template <typename T>
struct DualRowsOp {
  using Acc = T;
  using Pre = DualPre<T>;
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
  __device__ Pre pre(int i) const {
    const int k = n + i;
    return Pre{y[i], y_erg[i], vbar[k], v[k], v_erg[k], cost[k], lower[k], upper[k]};
  }
  __device__ Acc gather(int, int c) const { return vbar[c]; }
  __device__ void row(int i, const Acc* acc, const Pre& p, double*) const {
    const T sigma = static_cast<T>(sc->sigma);
    const T tau = static_cast<T>(sc->tau);
    const T theta = static_cast<T>(sc->theta);
    const T alpha = static_cast<T>(1.0 / static_cast<double>(sc->k0 + j + 1));
    const int k = n + i;
    const T yn = p.y + sigma * (acc[0] - p.vbar);
    y_new[i] = yn;
    y_erg[i] = (T(1) - alpha) * p.y_erg + alpha * yn;
    const T grad = p.cost + (-yn);
    const T vn = clamp_to(p.v - tau * grad, p.lower, p.upper);
    v_new[k] = vn;
    v_erg[k] = (T(1) - alpha) * p.v_erg + alpha * vn;
    vbar_next[k] = vn + theta * (vn - p.v);
  }
};

template <typename T>
struct PrimalPre {
  T v, v_erg, cost, lower, upper;
};

// Rows of A' (structural variables j < n): v_new_j = clamp(v_j - tau (cost_j + (A' y_new)_j)),
// its running average and the next extrapolation vbar_j = v_new_j + theta (v_new_j - v_j).
template <typename T>
struct PrimalColsOp {
  using Acc = T;
  using Pre = PrimalPre<T>;
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
  __device__ Pre pre(int col) const {
    return Pre{v[col], v_erg[col], cost[col], lower[col], upper[col]};
  }
  __device__ Acc gather(int, int c) const { return y_new[c]; }
  __device__ void row(int col, const Acc* acc, const Pre& p, double*) const {
    const T tau = static_cast<T>(sc->tau);
    const T theta = static_cast<T>(sc->theta);
    const T alpha = static_cast<T>(1.0 / static_cast<double>(sc->k0 + j + 1));
    const T grad = p.cost + acc[0];
    const T vn = clamp_to(p.v - tau * grad, p.lower, p.upper);
    v_new[col] = vn;
    v_erg[col] = (T(1) - alpha) * p.v_erg + alpha * vn;
    vbar_next[col] = vn + theta * (vn - p.v);
  }
};

// lambda: the part of a reduced cost g that the bounds of its variable can absorb (any sign when
// both bounds are finite, >= 0 with only a lower bound, <= 0 with only an upper one, 0 if free).
__device__ inline double absorbable(double g, double lo, double up) {
  const bool l = lo > -INFINITY;
  const bool u = up < INFINITY;
  return l && u ? g : l ? fmax(g, 0.0) : u ? fmin(g, 0.0) : 0.0;
}

template <int NV>
struct MeasurePre {
  double scale, cost, lower, upper;
  double v[NV];
  double y[NV];
};

// Evaluation sums (7 per iterate, in the model space before equilibration): 0 ||A x - r||^2,
// 1 dual residual^2, 2 primal objective, 3 dual objective, 4 ||x||^2, 5 ||r||^2, 6 ||y||^2.
// x = cs v_x, r = v_r / rs, y_o = rs y; A x = (A_eq v_x) / rs and A' y_o = (A_eq' y) / cs.
// Rows of A carry the logical variables' terms.
template <typename T, int NV>
struct MeasureRowsOp {
  using Acc = double;
  using Pre = MeasurePre<NV>;
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
  __device__ Pre pre(int i) const {
    const int k = n + i;
    Pre p;
    p.scale = rs[i];
    p.cost = cost[k];
    p.lower = lower[k];
    p.upper = upper[k];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      p.v[v] = static_cast<double>(vv[v][k]);
      p.y[v] = static_cast<double>(yy[v][i]);
    }
    return p;
  }
  __device__ Acc gather(int v, int c) const { return static_cast<double>(vv[v][c]); }
  __device__ void row(int, const Acc* acc, const Pre& p, double* s) const {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double value = p.v[v] / p.scale;
      const double y_o = p.y[v] * p.scale;
      const double g = p.cost - y_o;
      const double e = acc[v] / p.scale - value;
      s[7 * v + 0] += e * e;
      s[7 * v + 5] += value * value;
      s[7 * v + 6] += y_o * y_o;
      s[7 * v + 2] += p.cost * value;
      const double lam = absorbable(g, p.lower, p.upper);
      s[7 * v + 1] += (g - lam) * (g - lam);
      if (lam > 0.0) s[7 * v + 3] += lam * p.lower;
      if (lam < 0.0) s[7 * v + 3] += lam * p.upper;
    }
  }
};

// Rows of A' carry the structural variables' terms.
template <typename T, int NV>
struct MeasureColsOp {
  using Acc = double;
  using Pre = MeasurePre<NV>;
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
  __device__ Pre pre(int col) const {
    Pre p;
    p.scale = cs[col];
    p.cost = cost[col];
    p.lower = lower[col];
    p.upper = upper[col];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      p.v[v] = static_cast<double>(vv[v][col]);
      p.y[v] = 0.0;
    }
    return p;
  }
  __device__ Acc gather(int v, int c) const { return static_cast<double>(yy[v][c]); }
  __device__ void row(int, const Acc* acc, const Pre& p, double* s) const {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double value = p.v[v] * p.scale;
      const double g = p.cost + (m > 0 ? acc[v] / p.scale : 0.0);
      s[7 * v + 4] += value * value;
      s[7 * v + 2] += p.cost * value;
      const double lam = absorbable(g, p.lower, p.upper);
      s[7 * v + 1] += (g - lam) * (g - lam);
      if (lam > 0.0) s[7 * v + 3] += lam * p.lower;
      if (lam < 0.0) s[7 * v + 3] += lam * p.upper;
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

// out[q] = sum over blocks of partial[q * nblocks + b], in a fixed order: workgroup q sums row q
// (launch with one workgroup per sum).
__global__ __launch_bounds__(kBlock) void finish_sums_kernel(int nblocks,
                                                             const double* __restrict__ partial,
                                                             double* __restrict__ out) {
  __shared__ double wsum[kBlock / kWave];
  const int tid = threadIdx.x;
  const int q = blockIdx.x;
  double acc = 0.0;
  for (int b = tid; b < nblocks; b += kBlock) acc += partial[static_cast<size_t>(q) * nblocks + b];
  acc = group_sum<kWave>(acc);
  if ((tid & (kWave - 1)) == 0) wsum[tid / kWave] = acc;
  __syncthreads();
  if (tid == 0) out[q] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

// A restart at an iterate (cv, cy): it becomes the restart point, the current and next iterate,
// and both extrapolation points (so the next step starts without momentum).
template <typename T>
__global__ void restart_to_kernel(int nt, int m, const T* __restrict__ cv,
                                   const T* __restrict__ cy, T* __restrict__ v_rst,
                                   T* __restrict__ v, T* __restrict__ v_new,
                                   T* __restrict__ vbar0, T* __restrict__ vbar1,
                                   T* __restrict__ y_rst, T* __restrict__ y,
                                   T* __restrict__ y_new) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nt) {
    const T c = cv[k];
    v_rst[k] = c;
    v[k] = c;
    v_new[k] = c;
    vbar0[k] = c;
    vbar1[k] = c;
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
