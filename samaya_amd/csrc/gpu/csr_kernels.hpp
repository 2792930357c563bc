// Device kernels for the PDHG hot path on gfx950 (CDNA4, wavefront 64).
//
// Every product with A or A' runs through one kernel shape, csr_rows_kernel: one 256-thread
// workgroup per row block of the schedule (row_blocks.hpp). The workgroup
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
//
// Tiles. A launch may give each workgroup 1, 2 or 4 consecutive blocks of its list, and may
// order the workgroups so that each XCD takes one contiguous range of blocks; a row's sum depends
// only on its block, so every tile gives the same bits and a timing-based tuner can pick freely.
//
// Batches. Several LPs stacked block-diagonally share the kernels: every block belongs to one
// LP (its instance, binst[b]), whose step scalars the epilogue reads, and a launch covers a list
// of blocks (the blocks of the instances that step in that precision). A single LP is the batch
// of one. All sums are formed in an order fixed by the schedule: repeated runs are bit-identical,
// and an LP's sums do not depend on the other LPs of its batch.
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
  const int2* __restrict__ meta = nullptr;   // nblocks + 1 entries (first row, first nonzero)
  const int* __restrict__ binst = nullptr;   // instance of each block (null: all instance 0)
  const int* __restrict__ list = nullptr;    // blocks of this launch (null: block = workgroup)
  // The instance is blockIdx.y (one matrix shared by the instances, each launched over all
  // blocks) instead of binst[b].
  bool inst_from_y = false;
  // The launch's tile: `count` blocks (list entries, or blocks 0..count-1), `per_wg` consecutive
  // ones per workgroup; with `xcd`, each of the 8 XCDs (which take workgroups round-robin) gets
  // one contiguous range of the blocks, so rows that gather nearby columns share that XCD's L2.
  // No tile changes a single sum: every row is summed within its own block.
  int count = 0;
  int per_wg = 1;
  bool xcd = false;
};

// Workgroup w's place when the 8 XCDs each take one contiguous range of the n workgroups
// (hardware dispatches workgroup w to XCD w % 8).
__device__ inline int xcd_order(int w, int n) {
  constexpr int X = 8;
  const int x = w % X, i = w / X, q = n / X, r = n % X;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// Scalars of one PDHG step for one instance, read by the step kernels from device memory so that
// the same kernel nodes can be replayed from a HIP graph with new values. alpha = 1 / (k0 + j + 1)
// for the j-th step of a run whose first step is the k0-th since the instance's last restart.
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
// totals in index order); thread q < K writes total q to partial[q * stride].
// Slots q >= MaxFrom are maxima instead of sums.
template <int K, int MaxFrom = K>
__device__ inline void block_partials(const double (&local)[K], double* __restrict__ partial,
                                      size_t stride) {
  __shared__ double wsum[K][kBlock / kWave];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const double t = q < MaxFrom ? group_sum<kWave>(local[q]) : group_max<kWave>(local[q]);
    if ((tid & (kWave - 1)) == 0) wsum[q][tid / kWave] = t;
  }
  __syncthreads();
  if (tid < K) {
    partial[static_cast<size_t>(tid) * stride] =
        tid < MaxFrom ? ((wsum[tid][0] + wsum[tid][1]) + wsum[tid][2]) + wsum[tid][3]
                      : fmax(fmax(wsum[tid][0], wsum[tid][1]), fmax(wsum[tid][2], wsum[tid][3]));
  }
}

// What an epilogue's pre / gather / row receive about the workgroup's instance: the instance
// index, or, for an op that defines context(inst), whatever it derives from it once per
// workgroup (step sizes and the buffers of the instance's current slot).
template <typename Op>
__device__ inline auto op_context(const Op& op, int inst) {
  if constexpr (requires { op.context(inst); })
    return op.context(inst);
  else
    return inst;
}

// One row block b of instance inst (see the file comment).
template <typename V, typename Op>
__device__ __attribute__((always_inline)) inline void csr_rows_block(const CsrRows<V>& A, const Op& op, int b,
                                                                     int inst) {
  using Acc = typename Op::Acc;
  using Pre = typename Op::Pre;
  constexpr int NV = Op::kVec;
  constexpr int NS = Op::kSums > 0 ? Op::kSums : 1;
  __shared__ Acc prod[NV][kBlockNnz];
  __shared__ Acc rowsum[NV][kBlockRows];
  __shared__ Acc wsum[NV][kBlock / kWave];
  __shared__ int rp[kBlockRows + 1];
  const int tid = threadIdx.x;
  const int2 m0 = A.meta[b];
  const int2 m1 = A.meta[b + 1];
  if (op.skip(inst)) return;  // uniform over the workgroup, before any barrier
  const auto ctx = op_context(op, inst);
  const int r0 = m0.x, p0 = m0.y;
  const int nrows = m1.x - m0.x;
  const int nnz = m1.y - m0.y;
  const bool long_row = nrows == 1 && nnz > kLongRow;

  // Streams of values and indices, row offsets and epilogue inputs: all issued before any use.
  // No branch guards the loads: a lane past the block's end reads the block's last entry again
  // (a valid column to gather; its product is never stored), so the loads and, below, the
  // gathers go out back to back instead of one conditional block at a time.
  V a[kPer];
  int c[kPer];
  if (!long_row) {
    if (nnz > 0) {
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int k = min(tid + u * kBlock, nnz - 1);
        a[u] = __builtin_nontemporal_load(A.val + p0 + k);
        c[u] = __builtin_nontemporal_load(A.idx + p0 + k);
      }
    } else {
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        a[u] = V(0);
        c[u] = 0;
      }
    }
  }
  Pre pre{};
  if (tid < nrows) {
    rp[tid] = A.ptr[r0 + tid] - p0;
    pre = op.pre(r0 + tid, ctx);
  }
  if (tid == 0) rp[nrows] = nnz;

  double sums[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) sums[q] = 0.0;

  if (!long_row) {
    Acc xs[NV][kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
#pragma unroll
      for (int v = 0; v < NV; ++v) xs[v][u] = op.gather(v, c[u], ctx);
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int k = tid + u * kBlock;
      if (k < nnz) {
#pragma unroll
        for (int v = 0; v < NV; ++v) prod[v][k] = static_cast<Acc>(a[u]) * xs[v][u];
      }
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
      const int e = rp[grp + 1];
      for (int q = rp[grp] + lane; q < e; q += g) {
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
  } else {
    // One long row: the whole workgroup strides over it, kPer entries per thread per round
    // with every load of a round issued before the first product.
    Acc s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = Acc(0);
    const int end = p0 + nnz;
    for (int p = p0 + tid; p < end; p += kPer * kBlock) {
      V al[kPer];
      int cl[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int q = min(p + u * kBlock, end - 1);  // as above: loads without branches
        const V av = __builtin_nontemporal_load(A.val + q);
        cl[u] = __builtin_nontemporal_load(A.idx + q);
        al[u] = p + u * kBlock < end ? av : V(0);
      }
      Acc gv[NV][kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const Acc g = op.gather(v, cl[u], ctx);
          gv[v][u] = p + u * kBlock < end ? g : Acc(0);
        }
      }
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
#pragma unroll
        for (int v = 0; v < NV; ++v) s[v] += static_cast<Acc>(al[u]) * gv[v][u];
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
  if (tid < nrows) {
    Acc acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = rowsum[v][tid];
    op.row(r0 + tid, ctx, acc, pre, sums);
  }
  if constexpr (Op::kSums > 0)
    block_partials<Op::kSums, Op::kMaxFrom>(sums, op.partial + static_cast<size_t>(inst) * op.inst_stride + b,
                                           op.stride);
}

// Tiled: the launch's tile (CsrRows::per_wg, xcd) maps workgroups to blocks; otherwise workgroup
// w takes block w of the launch (the plain tile, compiled apart so that its block's schedule
// entries stay scalar loads and its value and index loads all go out before the first wait).
template <typename V, typename Op, bool Tiled>
__global__ __launch_bounds__(kBlock) void csr_rows_kernel(CsrRows<V> A, Op op) {
  if constexpr (!Tiled) {
    const int b = A.list != nullptr ? A.list[blockIdx.x] : static_cast<int>(blockIdx.x);
    const int inst = A.inst_from_y ? static_cast<int>(blockIdx.y) : A.binst != nullptr ? A.binst[b] : 0;
    csr_rows_block(A, op, b, inst);
  } else {
    const int w = A.xcd ? xcd_order(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x))
                        : static_cast<int>(blockIdx.x);
    for (int t = 0; t < A.per_wg; ++t) {
      const int slot = w * A.per_wg + t;
      if (slot >= A.count) return;  // uniform over the workgroup
      if (t > 0) __syncthreads();   // the last block's LDS reads are done
      const int b = A.list != nullptr ? A.list[slot] : slot;
      const int inst = A.inst_from_y ? static_cast<int>(blockIdx.y) : A.binst != nullptr ? A.binst[b] : 0;
      csr_rows_block(A, op, b, inst);
    }
  }
}

// ------------------------------------------------------------------------------------------
// Epilogues.
// ------------------------------------------------------------------------------------------

struct NoPre {};

// The tile tuner's stand-in for a step kernel: the same product, plus per row what a step's
// epilogue moves (five vectors read, three written), into scratch so that timing it leaves the
// iterates alone.
template <typename T>
struct TuneOp {
  using Acc = T;
  struct Pre {
    T a, b, c, d, e;
  };
  static constexpr int kVec = 1;
  static constexpr int kSums = 0;
  const T* __restrict__ x;
  const T* __restrict__ in[5];
  T* __restrict__ out[3];
  __device__ bool skip(int) const { return false; }
  __device__ Pre pre(int r, int) const { return Pre{in[0][r], in[1][r], in[2][r], in[3][r], in[4][r]}; }
  __device__ Acc gather(int, int c, int) const { return x[c]; }
  __device__ void row(int r, int, const Acc* acc, const Pre& p, double*) const {
    const T v = clamp_to(p.a - p.b * (p.c + acc[0]), p.d, p.e);
    out[0][r] = v;
    out[1][r] = v + p.a;
    out[2][r] = v - p.b;
  }
};

// Plain product y = A x (the standalone SpMV entry point).
template <typename V, typename X>
struct SpmvOp {
  using Acc = X;
  using Pre = NoPre;
  static constexpr int kVec = 1;
  static constexpr int kSums = 0;
  const X* __restrict__ x;
  X* __restrict__ y;
  __device__ bool skip(int) const { return false; }
  __device__ Pre pre(int, int) const { return {}; }
  __device__ Acc gather(int, int c, int) const { return x[c]; }
  __device__ void row(int r, int, const Acc* acc, const Pre&, double*) const { y[r] = acc[0]; }
};

template <typename T>
struct DualPre {
  T y, y_erg, vbar, v, v_erg, cost, lower, upper;
};

// Rows of A: the dual update of row i and the primal update of its logical variable n + i.
//   y_new_i = y_i + sigma (A vbar_x - vbar_{n+i})_i,  y_erg_i += alpha (y_new_i - y_erg_i)
//   v_new_{n+i} = clamp(v_{n+i} - tau (cost_{n+i} - y_new_i)),  its average and next vbar.
// (n: the number of structural variables of the whole stack; the logical variables follow them.)
// A: the accumulator. A wider one (FP32 steps, FP64 accumulation) also forms the update that uses
// the product in A and rounds it once to T; A = T is the plain step.
template <typename T, typename A = T>
struct DualRowsOp {
  using Acc = A;
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
  __device__ bool skip(int) const { return false; }
  __device__ Pre pre(int i, int) const {
    const int k = n + i;
    return Pre{y[i], y_erg[i], vbar[k], v[k], v_erg[k], cost[k], lower[k], upper[k]};
  }
  __device__ Acc gather(int, int c, int) const { return vbar[c]; }
  __device__ void row(int i, int inst, const Acc* acc, const Pre& p, double*) const {
    const StepScalars& s = sc[inst];
    const T sigma = static_cast<T>(s.sigma);
    const T tau = static_cast<T>(s.tau);
    const T theta = static_cast<T>(s.theta);
    const T alpha = static_cast<T>(1.0 / static_cast<double>(s.k0 + j + 1));
    const int k = n + i;
    T yn;
    if constexpr (std::is_same_v<A, T>)
      yn = p.y + sigma * (acc[0] - p.vbar);
    else
      yn = static_cast<T>(static_cast<A>(p.y) + static_cast<A>(sigma) * (acc[0] - static_cast<A>(p.vbar)));
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
// A: the accumulator, as for DualRowsOp.
template <typename T, typename A = T>
struct PrimalColsOp {
  using Acc = A;
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
  __device__ bool skip(int) const { return false; }
  __device__ Pre pre(int col, int) const {
    return Pre{v[col], v_erg[col], cost[col], lower[col], upper[col]};
  }
  __device__ Acc gather(int, int c, int) const { return y_new[c]; }
  __device__ void row(int col, int inst, const Acc* acc, const Pre& p, double*) const {
    const StepScalars& s = sc[inst];
    const T tau = static_cast<T>(s.tau);
    const T theta = static_cast<T>(s.theta);
    const T alpha = static_cast<T>(1.0 / static_cast<double>(s.k0 + j + 1));
    T vn;
    if constexpr (std::is_same_v<A, T>) {
      const T grad = p.cost + acc[0];
      vn = clamp_to(p.v - tau * grad, p.lower, p.upper);
    } else {
      vn = clamp_to(static_cast<T>(static_cast<A>(p.v) - static_cast<A>(tau) * (static_cast<A>(p.cost) + acc[0])),
                    p.lower, p.upper);
    }
    v_new[col] = vn;
    v_erg[col] = (T(1) - alpha) * p.v_erg + alpha * vn;
    vbar_next[col] = vn + theta * (vn - p.v);
  }
};

// ------------------------------------------------------------------------------------------
// Restarted Halpern PDHG with reflection (no reference counterpart; Lu and Yang, "Restarted
// Halpern PDHG for linear programming", 2024, the iteration of cuPDLPx). With z = (v, y), the
// PDHG operator in the primal-first order
//   T(z):  v+ = clamp(v - tau (c + K'y)),  y+ = y + sigma K (2 v+ - v),   K = [A -I],
// and the anchor z0 (the last restart point), a step is
//   z' = lambda ((1 + rho) T(z) - rho z) + (1 - lambda) z0,   lambda = (k + 1) / (k + 2),
// k the steps since the restart and rho in [0, 1] the reflection. No averages: a restart moves
// to T(z) and makes it the anchor. Buffers of the stacked state: z in slot p of v, y; z' into
// slot 1 - p; the anchor in v_rst / y_rst; w = 2 v+ - v of the structural variables in vbar[0]
// (the dual kernel's gathered vector); T(z) itself into v_erg / y_erg at the last step of a run
// only (`keep`), where the evaluation and the restart read it.
// The primal kernel (rows of A': structural variables) runs first, then the dual kernel (rows of
// A: the duals and, from y_i alone, the logical variables).
// ------------------------------------------------------------------------------------------

template <typename T>
struct HalpernColsPre {
  T v, anchor, cost, lower, upper;
};

template <typename T, typename A = T>
struct HalpernColsOp {
  using Acc = A;
  using Pre = HalpernColsPre<T>;
  static constexpr int kVec = 1;
  static constexpr int kSums = 0;
  int j;       // step index within the run
  bool keep;   // write T(z) (last step of the run)
  double rho;  // reflection
  const StepScalars* __restrict__ sc;
  const T* __restrict__ y;
  const T* __restrict__ v;
  const T* __restrict__ v_anchor;
  T* __restrict__ v_next;
  T* __restrict__ w;
  T* __restrict__ v_plus;
  const T* __restrict__ cost;
  const T* __restrict__ lower;
  const T* __restrict__ upper;
  __device__ bool skip(int) const { return false; }
  __device__ Pre pre(int col, int) const {
    return Pre{v[col], v_anchor[col], cost[col], lower[col], upper[col]};
  }
  __device__ Acc gather(int, int c, int) const { return y[c]; }
  __device__ void row(int col, int inst, const Acc* acc, const Pre& p, double*) const {
    const StepScalars& s = sc[inst];
    const T tau = static_cast<T>(s.tau);
    const double k = static_cast<double>(s.k0 + j);
    const T lam = static_cast<T>((k + 1.0) / (k + 2.0));
    const T r = static_cast<T>(rho);
    T vp;
    if constexpr (std::is_same_v<A, T>)
      vp = clamp_to(p.v - tau * (p.cost + acc[0]), p.lower, p.upper);
    else
      vp = clamp_to(static_cast<T>(static_cast<A>(p.v) - static_cast<A>(tau) * (static_cast<A>(p.cost) + acc[0])),
                    p.lower, p.upper);
    w[col] = vp + (vp - p.v);
    v_next[col] = lam * ((T(1) + r) * vp - r * p.v) + (T(1) - lam) * p.anchor;
    if (keep) v_plus[col] = vp;
  }
};

template <typename T>
struct HalpernRowsPre {
  T y, y_anchor, v, v_anchor, cost, lower, upper;
};

template <typename T, typename A = T>
struct HalpernRowsOp {
  using Acc = A;
  using Pre = HalpernRowsPre<T>;
  static constexpr int kVec = 1;
  static constexpr int kSums = 0;
  int n;
  int j;
  bool keep;
  double rho;
  const StepScalars* __restrict__ sc;
  const T* __restrict__ w;  // 2 v+ - v of the structural variables
  const T* __restrict__ y;
  const T* __restrict__ y_anchor;
  T* __restrict__ y_next;
  T* __restrict__ y_plus;
  const T* __restrict__ v;
  const T* __restrict__ v_anchor;
  T* __restrict__ v_next;
  T* __restrict__ v_plus;
  const T* __restrict__ cost;
  const T* __restrict__ lower;
  const T* __restrict__ upper;
  __device__ bool skip(int) const { return false; }
  __device__ Pre pre(int i, int) const {
    const int k = n + i;
    return Pre{y[i], y_anchor[i], v[k], v_anchor[k], cost[k], lower[k], upper[k]};
  }
  __device__ Acc gather(int, int c, int) const { return w[c]; }
  __device__ void row(int i, int inst, const Acc* acc, const Pre& p, double*) const {
    const StepScalars& s = sc[inst];
    const T tau = static_cast<T>(s.tau);
    const T sigma = static_cast<T>(s.sigma);
    const double kk = static_cast<double>(s.k0 + j);
    const T lam = static_cast<T>((kk + 1.0) / (kk + 2.0));
    const T r = static_cast<T>(rho);
    const int k = n + i;
    // The logical variable: its gradient is cost - y_i.
    const T vp = clamp_to(p.v - tau * (p.cost + (-p.y)), p.lower, p.upper);
    const T wr = vp + (vp - p.v);
    T yp;
    if constexpr (std::is_same_v<A, T>)
      yp = p.y + sigma * (acc[0] - wr);
    else
      yp = static_cast<T>(static_cast<A>(p.y) + static_cast<A>(sigma) * (acc[0] - static_cast<A>(wr)));
    y_next[i] = lam * ((T(1) + r) * yp - r * p.y) + (T(1) - lam) * p.y_anchor;
    v_next[k] = lam * ((T(1) + r) * vp - r * p.v) + (T(1) - lam) * p.v_anchor;
    if (keep) {
      y_plus[i] = yp;
      v_plus[k] = vp;
    }
  }
};

// Squared distances of T(z) (v_erg, y_erg) from z (slot p): the fixed-point residual of a
// Halpern run's last step. 0: ||v+ - v||^2, 1: ||y+ - y||^2.
template <typename T>
struct ResidualOp {
  static constexpr int kSums = 2;
  static constexpr int kMaxFrom = 2;
  const int* __restrict__ which;  // per instance: -1 none, else measured
  const T* __restrict__ v;
  const T* __restrict__ y;
  const T* __restrict__ v_plus;
  const T* __restrict__ y_plus;
  double* __restrict__ partial;
  size_t stride;
  int inst_stride = 0;
  __device__ int flag(int inst) const { return which[inst]; }
  __device__ void row_a(int i, int k, int, double* s, int) const {
    const double dv = static_cast<double>(v_plus[k]) - static_cast<double>(v[k]);
    const double dy = static_cast<double>(y_plus[i]) - static_cast<double>(y[i]);
    s[0] += dv * dv;
    s[1] += dy * dy;
  }
  __device__ void col(int k, int, double* s, int) const {
    const double dv = static_cast<double>(v_plus[k]) - static_cast<double>(v[k]);
    s[0] += dv * dv;
  }
};

// ------------------------------------------------------------------------------------------
// Adaptive step sizes: PDLP's rule (Applegate et al., "Practical large-scale linear programming
// using primal-dual hybrid gradient", NeurIPS 2021, sec. 3.1). The reference has no counterpart:
// its adaptive step is disabled code (src/lp/pdlp.cpp:1051-1067). An attempt takes one PDHG step
// with step size eta from z = (v, y) in the primal-first order
//   v+ = clamp(v - tau (c + K'y)),  y+ = y + sigma K (2 v+ - v),   K = [A -I],
// tau = eta / omega, sigma = eta omega. With dv = v+ - v, dy = y+ - y,
//   movement = omega ||dv||^2 / 2 + ||dy||^2 / (2 omega),  interaction = |dy' K dv|,
//   limit = movement / interaction (infinite when the interaction is 0).
// The attempt is accepted when eta <= limit; either way the next attempt uses
//   eta' = min((1 - (k + 1)^-0.3) limit, (1 + (k + 1)^-0.6) eta),   k = attempts so far.
// A rejected attempt is retried from the same z. The averages weight each accepted iterate by its
// step size. Every quantity lives on the device, so a run of attempts replays from one graph:
// each instance's current iterate sits in slot `cur` of the two buffers of v, y and A x (and of
// the averages in slot `erg`), an attempt writes the other slot, and acceptance flips the slot.
// K dv comes without another product: the dual kernel keeps h = A x of the current iterate, and
// A (2 x+ - x) = 2 A x+ - h gives A x+ and A dx from the one product it forms anyway.
// ------------------------------------------------------------------------------------------

struct AdaptState {
  double eta;        // step size of the next attempt
  double omega;      // primal weight
  double weight;     // sum of the accepted step sizes since the restart (the averages' weight)
  long long total;   // attempts since the start of the solve (k of the rule)
  long long rejected;
  int cur;           // slot of the current v, y, h
  int erg;           // slot of the current averages
};

// A x of the current iterate, for the instances flagged (after a start, restart or change of
// precision; the others keep the A x their own steps formed, so a batch stays bit-equal to
// single handles).
template <typename T>
struct AxRefreshOp {
  using Acc = T;
  using Pre = NoPre;
  static constexpr int kVec = 1;
  static constexpr int kSums = 0;
  const int* __restrict__ flags;
  const T* __restrict__ x;
  T* __restrict__ y;
  __device__ bool skip(int inst) const { return flags[inst] == 0; }
  __device__ Pre pre(int, int) const { return {}; }
  __device__ Acc gather(int, int c, int) const { return x[c]; }
  __device__ void row(int r, int, const Acc* acc, const Pre&, double*) const { y[r] = acc[0]; }
};

template <typename T>
struct AdaptPrimalCtx {
  T tau, alpha;
  const T* y;   // gathered: the current dual
  const T* v;
  T* v_out;
  const T* e;
  T* e_out;
};

template <typename T>
struct AdaptPrimalPre {
  T v, v_erg, cost, lower, upper;
};

// Rows of A' (structural variables): v+_j = clamp(v_j - tau (c_j + (A'y)_j)), the point
// 2 v+ - v the dual kernel multiplies, the weighted average, and ||dx||^2 (partial slot 0).
template <typename T, typename A = T>
struct AdaptPrimalOp {
  using Acc = A;  // a wider accumulator as for DualRowsOp
  using Pre = AdaptPrimalPre<T>;
  using Ctx = AdaptPrimalCtx<T>;
  static constexpr int kVec = 1;
  static constexpr int kSums = 1;
  static constexpr int kMaxFrom = 1;
  const AdaptState* __restrict__ st;
  const T* y0;
  const T* y1;
  T* v0;
  T* v1;
  T* e0;
  T* e1;
  T* __restrict__ xbar;
  const T* __restrict__ cost;
  const T* __restrict__ lower;
  const T* __restrict__ upper;
  double* __restrict__ partial;
  size_t stride;
  int inst_stride = 0;
  __device__ bool skip(int) const { return false; }
  __device__ Ctx context(int inst) const {
    const AdaptState s = st[inst];
    const bool c = s.cur != 0, g = s.erg != 0;
    return Ctx{static_cast<T>(s.eta / s.omega), static_cast<T>(s.eta / (s.weight + s.eta)),
               c ? y1 : y0, c ? v1 : v0, c ? v0 : v1, g ? e1 : e0, g ? e0 : e1};
  }
  __device__ Pre pre(int col, const Ctx& x) const {
    return Pre{x.v[col], x.e[col], cost[col], lower[col], upper[col]};
  }
  __device__ Acc gather(int, int c, const Ctx& x) const { return x.y[c]; }
  __device__ void row(int col, const Ctx& x, const Acc* acc, const Pre& p, double* s) const {
    T vn;
    if constexpr (std::is_same_v<A, T>)
      vn = clamp_to(p.v - x.tau * (p.cost + acc[0]), p.lower, p.upper);
    else
      vn = clamp_to(static_cast<T>(static_cast<A>(p.v) - static_cast<A>(x.tau) * (static_cast<A>(p.cost) + acc[0])),
                    p.lower, p.upper);
    x.v_out[col] = vn;
    xbar[col] = vn + (vn - p.v);
    x.e_out[col] = (T(1) - x.alpha) * p.v_erg + x.alpha * vn;
    const double d = static_cast<double>(vn) - static_cast<double>(p.v);
    s[0] += d * d;
  }
};

template <typename T>
struct AdaptDualCtx {
  T tau, sigma, alpha;
  const T* y;
  T* y_out;
  const T* ye;
  T* ye_out;
  const T* v;
  T* v_out;
  const T* ve;
  T* ve_out;
  const T* h;
  T* h_out;
};

template <typename T>
struct AdaptDualPre {
  T y, y_erg, r, r_erg, cost, lower, upper, h;
};

// Rows of A: row i's logical variable r = v_{n+i} steps first (r+ = clamp(r - tau (c_r - y_i))),
// then y+_i = y_i + sigma ((A xbar)_i - (2 r+ - r)), with xbar = 2 x+ - x from the primal kernel.
// A x+ = ((A xbar)_i + h_i) / 2 replaces h, and (A dx)_i = ((A xbar)_i - h_i) / 2 gives row i's
// share of dy' K dv. Partial slots: 0 ||dr||^2, 1 ||dy||^2, 2 dy' K dv.
template <typename T, typename A = T>
struct AdaptDualOp {
  using Acc = A;  // a wider accumulator as for DualRowsOp
  using Pre = AdaptDualPre<T>;
  using Ctx = AdaptDualCtx<T>;
  static constexpr int kVec = 1;
  static constexpr int kSums = 3;
  static constexpr int kMaxFrom = 3;
  int n;
  const AdaptState* __restrict__ st;
  const T* __restrict__ xbar;
  T* y0;
  T* y1;
  T* ye0;
  T* ye1;
  T* v0;
  T* v1;
  T* ve0;
  T* ve1;
  T* h0;
  T* h1;
  const T* __restrict__ cost;
  const T* __restrict__ lower;
  const T* __restrict__ upper;
  double* __restrict__ partial;
  size_t stride;
  int inst_stride = 0;
  __device__ bool skip(int) const { return false; }
  __device__ Ctx context(int inst) const {
    const AdaptState s = st[inst];
    const bool c = s.cur != 0, g = s.erg != 0;
    return Ctx{static_cast<T>(s.eta / s.omega), static_cast<T>(s.eta * s.omega),
               static_cast<T>(s.eta / (s.weight + s.eta)),
               c ? y1 : y0, c ? y0 : y1, g ? ye1 : ye0, g ? ye0 : ye1,
               c ? v1 : v0, c ? v0 : v1, g ? ve1 : ve0, g ? ve0 : ve1,
               c ? h1 : h0, c ? h0 : h1};
  }
  __device__ Pre pre(int i, const Ctx& x) const {
    const int k = n + i;
    return Pre{x.y[i], x.ye[i], x.v[k], x.ve[k], cost[k], lower[k], upper[k], x.h[i]};
  }
  __device__ Acc gather(int, int c, const Ctx&) const { return xbar[c]; }
  __device__ void row(int i, const Ctx& x, const Acc* acc, const Pre& p, double* s) const {
    const int k = n + i;
    const T rn = clamp_to(p.r - x.tau * (p.cost - p.y), p.lower, p.upper);
    const T rbar = rn + (rn - p.r);
    T yn, adx, hn;
    if constexpr (std::is_same_v<A, T>) {
      yn = p.y + x.sigma * (acc[0] - rbar);
      adx = (acc[0] - p.h) * T(0.5);
      hn = (acc[0] + p.h) * T(0.5);
    } else {
      yn = static_cast<T>(static_cast<A>(p.y) + static_cast<A>(x.sigma) * (acc[0] - static_cast<A>(rbar)));
      adx = static_cast<T>((acc[0] - static_cast<A>(p.h)) * A(0.5));
      hn = static_cast<T>((acc[0] + static_cast<A>(p.h)) * A(0.5));
    }
    x.y_out[i] = yn;
    x.v_out[k] = rn;
    x.h_out[i] = hn;
    x.ye_out[i] = (T(1) - x.alpha) * p.y_erg + x.alpha * yn;
    x.ve_out[k] = (T(1) - x.alpha) * p.r_erg + x.alpha * rn;
    const double dr = static_cast<double>(rn) - static_cast<double>(p.r);
    const double dy = static_cast<double>(yn) - static_cast<double>(p.y);
    s[0] += dr * dr;
    s[1] += dy * dy;
    s[2] += dy * (static_cast<double>(adx) - dr);
  }
};

// The attempt's totals from the block partials (the instance's A blocks hold slots 0-2, its A'
// blocks slot 0), then the rule. It runs after every attempt, so the partials (tens of thousands
// for a large LP) are spread over `nslices` 1024-thread workgroups per instance (grid: nslices x
// instances), each thread with many loads in flight. Thread t of slice r adds entries
// r 1024 + t, then every nslices 1024 further, in order. With one slice the workgroup applies the
// rule itself. With more, each writes its slice totals and the last to arrive adds the slices in
// slice order, applies the rule and sets the arrival count back to 0 for the next attempt; the
// arrival needs a device-scope fence, which writes back the XCD's L2, so the host asks for more
// slices only for an instance of many blocks. The totals are fixed either way for a given
// nslices. Instances that do not step (frozen) keep their state.
constexpr int kDecideThreads = 1024;
constexpr int kDecideUnroll = 12;
constexpr int kDecideSlices = 4;            // at most
constexpr int kDecideBlocksPerSlice = 4096;  // partials per slice the host aims for

static __global__ __launch_bounds__(kDecideThreads) void adapt_decide_kernel(
    int stride, const double* __restrict__ partial, const int2* __restrict__ ranges,
    const int* __restrict__ prec, AdaptState* __restrict__ st, int nslices,
    double* __restrict__ slices, unsigned* __restrict__ arrivals) {
  constexpr int W = kDecideThreads / kWave;
  const int step = nslices * kDecideThreads;
  __shared__ double wsum[3][W];
  __shared__ int last;
  const int slice = blockIdx.x, i = blockIdx.y;
  if (prec[i] < 0) return;  // frozen (the same for every slice of the instance)
  const int tid = threadIdx.x;
  const int2 ra = ranges[2 * i];
  const int2 rt = ranges[2 * i + 1];
  // Both ranges in one loop, so every round has the loads of both in flight.
  const int len_a = ra.y - ra.x, len_t = rt.y - rt.x;
  const int span = max(len_a, len_t);
  double acc[3] = {0.0, 0.0, 0.0};
  double at = 0.0;
  for (int o = slice * kDecideThreads + tid; o < span; o += kDecideUnroll * step) {
    double v[kDecideUnroll][3];
    double w[kDecideUnroll];
#pragma unroll
    for (int u = 0; u < kDecideUnroll; ++u) {
      const int b = o + u * step;
#pragma unroll
      for (int q = 0; q < 3; ++q)
        v[u][q] = b < len_a ? partial[static_cast<size_t>(q) * stride + ra.x + b] : 0.0;
      w[u] = b < len_t ? partial[rt.x + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kDecideUnroll; ++u) {
#pragma unroll
      for (int q = 0; q < 3; ++q) acc[q] += v[u][q];
      at += w[u];
    }
  }
  acc[0] = acc[0] + at;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double t = group_sum<kWave>(acc[q]);
    if ((tid & (kWave - 1)) == 0) wsum[q][tid / kWave] = t;
  }
  __syncthreads();
  double tot[3] = {0.0, 0.0, 0.0};
  if (nslices == 1) {
    if (tid != 0) return;
#pragma unroll
    for (int q = 0; q < 3; ++q)
      for (int w = 0; w < W; ++w) tot[q] += wsum[q][w];
  } else {
    if (tid == 0) {
      double* mine = slices + (static_cast<size_t>(i) * kDecideSlices + slice) * 3;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double t = 0.0;
        for (int w = 0; w < W; ++w) t += wsum[q][w];
        mine[q] = t;
      }
      __threadfence();  // this slice's totals are visible before its arrival is
      last = atomicAdd(arrivals + i, 1u) == static_cast<unsigned>(nslices - 1);
    }
    __syncthreads();
    if (tid != 0 || !last) return;
    __threadfence();
    for (int r = 0; r < nslices; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        tot[q] += __hip_atomic_load(slices + (static_cast<size_t>(i) * kDecideSlices + r) * 3 + q,
                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    arrivals[i] = 0;
  }
  AdaptState s = st[i];
  const double movement = 0.5 * s.omega * tot[0] + 0.5 / s.omega * tot[1];
  const double interaction = fabs(tot[2]);
  const double limit = interaction > 0.0 ? movement / interaction : INFINITY;
  const bool ok = s.eta <= limit;
  s.total += 1;
  const double k1 = static_cast<double>(s.total + 1);
  const double next = fmin((1.0 - pow(k1, -0.3)) * limit, (1.0 + pow(k1, -0.6)) * s.eta);
  if (ok) {
    s.cur ^= 1;
    s.erg ^= 1;
    s.weight += s.eta;
  } else {
    s.rejected += 1;
  }
  s.eta = next;
  st[i] = s;
}

// After a run of attempts: every stepping instance's current iterate moves to the slot the
// caller expects (`want`; the previous point, accepted or tried, to the other one), and its
// averages to slot 0, so that the rest of the handle sees the layout of constant steps.
template <typename T>
struct AdaptSettleOp {
  static constexpr int kSums = 0;
  const AdaptState* __restrict__ st;
  const int* __restrict__ prec;  // per instance: precision of its steps (frozen < 0)
  int q;                         // the precision of this launch's buffers
  int want;
  T* v0;
  T* v1;
  T* y0;
  T* y1;
  T* h0;
  T* h1;
  T* ve0;
  const T* ve1;
  T* ye0;
  const T* ye1;
  __device__ int flag(int inst) const {
    if (prec[inst] != q) return -1;
    const int f = (st[inst].cur != want ? 1 : 0) | (st[inst].erg != 0 ? 2 : 0);
    return f == 0 ? -1 : f;
  }
  __device__ void col(int k, int f, double*, int) const {
    if (f & 1) {
      const T a = v0[k];
      v0[k] = v1[k];
      v1[k] = a;
    }
    if (f & 2) ve0[k] = ve1[k];
  }
  __device__ void row_a(int i, int k, int f, double* s, int inst) const {
    col(k, f, s, inst);
    if (f & 1) {
      const T a = y0[i];
      y0[i] = y1[i];
      y1[i] = a;
      const T b = h0[i];
      h0[i] = h1[i];
      h1[i] = b;
    }
    if (f & 2) ye0[i] = ye1[i];
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
// Rows of A carry the logical variables' terms. Block b leaves its totals at
// partial[q * stride + b].
template <typename T, int NV>
struct MeasureRowsOp {
  using Acc = double;
  using Pre = MeasurePre<NV>;
  static constexpr int kVec = NV;
  static constexpr int kSums = 7 * NV;
  static constexpr int kMaxFrom = kSums;
  int n;
  const T* __restrict__ vv[NV];
  const T* __restrict__ yy[NV];
  const double* __restrict__ rs;
  const double* __restrict__ cost;
  const double* __restrict__ lower;
  const double* __restrict__ upper;
  double* __restrict__ partial;
  size_t stride;
  int inst_stride = 0;
  __device__ bool skip(int) const { return false; }
  __device__ Pre pre(int i, int) const {
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
  __device__ Acc gather(int v, int c, int) const { return static_cast<double>(vv[v][c]); }
  __device__ void row(int, int, const Acc* acc, const Pre& p, double* s) const {
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
  static constexpr int kMaxFrom = kSums;
  const T* __restrict__ vv[NV];
  const T* __restrict__ yy[NV];
  const double* __restrict__ cs;
  const double* __restrict__ cost;
  const double* __restrict__ lower;
  const double* __restrict__ upper;
  double* __restrict__ partial;
  size_t stride;
  int inst_stride = 0;
  __device__ bool skip(int) const { return false; }
  __device__ Pre pre(int col, int) const {
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
  __device__ Acc gather(int v, int c, int) const { return static_cast<double>(yy[v][c]); }
  __device__ void row(int, int, const Acc* acc, const Pre& p, double* s) const {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double value = p.v[v] * p.scale;
      const double g = p.cost + acc[v] / p.scale;
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
// Row-wise kernels over both schedules (no products): workgroup w < nba takes block w of the A
// schedule (rows i: y_i and the logical variable v_{n+i}), w >= nba takes block w - nba of the A'
// schedule (columns j: v_j). Every entry of every vector belongs to exactly one block, hence to
// one instance, whose flag says whether the workgroup works at all.
// ------------------------------------------------------------------------------------------

struct BlockRows {
  const int2* __restrict__ meta_a;
  const int* __restrict__ binst_a;
  int nba;
  const int2* __restrict__ meta_t;
  const int* __restrict__ binst_t;
  int n;  // structural variables of the stack: v_{n+i} is row i's logical variable
  bool inst_from_y = false;  // the instance is blockIdx.y (a shared matrix's schedules)
  int first = 0;             // workgroup w covers block first + w of the combined schedule
};

template <typename Op>
__global__ __launch_bounds__(kBlock) void block_rows_kernel(BlockRows R, Op op) {
  constexpr int NS = Op::kSums > 0 ? Op::kSums : 1;
  const int w = R.first + static_cast<int>(blockIdx.x);
  const bool is_a = w < R.nba;
  const int b = is_a ? w : w - R.nba;
  const int2* meta = is_a ? R.meta_a : R.meta_t;
  const int inst = R.inst_from_y ? static_cast<int>(blockIdx.y) : (is_a ? R.binst_a : R.binst_t)[b];
  const int r0 = meta[b].x;
  const int nrows = meta[b + 1].x - r0;
  const int tid = threadIdx.x;
  double local[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) local[q] = 0.0;
  // Blocks of instances without a flag leave their partials untouched: launches for the two
  // precisions cover disjoint instances and must not overwrite each other's totals.
  const int flag = op.flag(inst);  // uniform over the workgroup
  if (flag < 0) return;
  if (tid < nrows) {
    if (is_a)
      op.row_a(r0 + tid, R.n + r0 + tid, flag, local, inst);
    else
      op.col(r0 + tid, flag, local, inst);
  }
  if constexpr (Op::kSums > 0)
    block_partials<Op::kSums, Op::kMaxFrom>(local, op.partial + static_cast<size_t>(inst) * op.inst_stride + w,
                                           op.stride);
}

// Squared distances of the candidate (current or average, by the instance's flag) from the
// restart point: 0: ||v - v_rst||^2, 1: ||y - y_rst||^2.
template <typename T>
struct DistanceOp {
  static constexpr int kSums = 2;
  static constexpr int kMaxFrom = 2;
  const int* __restrict__ which;  // per instance: -1 none, 0 current (v_new, y_new), 1 average
  const T* __restrict__ v_new;
  const T* __restrict__ y_new;
  const T* __restrict__ v_erg;
  const T* __restrict__ y_erg;
  const T* __restrict__ v_rst;
  const T* __restrict__ y_rst;
  double* __restrict__ partial;
  size_t stride;
  int inst_stride = 0;
  __device__ int flag(int inst) const { return which[inst]; }
  __device__ void row_a(int i, int k, int f, double* s, int) const {
    const double dv = static_cast<double>(f == 0 ? v_new[k] : v_erg[k]) - static_cast<double>(v_rst[k]);
    const double dy = static_cast<double>(f == 0 ? y_new[i] : y_erg[i]) - static_cast<double>(y_rst[i]);
    s[0] += dv * dv;
    s[1] += dy * dy;
  }
  __device__ void col(int k, int f, double* s, int) const {
    const double dv = static_cast<double>(f == 0 ? v_new[k] : v_erg[k]) - static_cast<double>(v_rst[k]);
    s[0] += dv * dv;
  }
};

// A restart at the candidate (current or average): it becomes the restart point, the current and
// next iterate, and both extrapolation points (so the next step starts without momentum). The
// candidate may be v_new / y_new itself: each entry is read before it is written, by one thread.
template <typename T>
struct RestartOp {
  static constexpr int kSums = 0;
  const int* __restrict__ which;
  T* v;
  T* v_new;
  T* y;
  T* y_new;
  const T* v_erg;
  const T* y_erg;
  T* __restrict__ v_rst;
  T* __restrict__ y_rst;
  T* __restrict__ vbar0;
  T* __restrict__ vbar1;
  __device__ int flag(int inst) const { return which[inst]; }
  __device__ void set_v(int k, int f) const {
    const T c = f == 0 ? v_new[k] : v_erg[k];
    v_rst[k] = c;
    v[k] = c;
    v_new[k] = c;
    vbar0[k] = c;
    vbar1[k] = c;
  }
  __device__ void row_a(int i, int k, int f, double*, int) const {
    set_v(k, f);
    const T c = f == 0 ? y_new[i] : y_erg[i];
    y_rst[i] = c;
    y[i] = c;
    y_new[i] = c;
  }
  __device__ void col(int k, int f, double*, int) const { set_v(k, f); }
};

// The iterate state of the flagged instances from one precision to the other.
template <typename From, typename To>
struct ConvertOp {
  static constexpr int kSums = 0;
  const int* __restrict__ which;
  const From* vin[6];  // v[2], vbar[2], v_erg, v_rst
  To* vout[6];
  const From* yin[4];  // y[2], y_erg, y_rst
  To* yout[4];
  __device__ int flag(int inst) const { return which[inst]; }
  __device__ void col(int k, int, double*, int) const {
#pragma unroll
    for (int q = 0; q < 6; ++q) vout[q][k] = static_cast<To>(vin[q][k]);
  }
  __device__ void row_a(int i, int k, int f, double* s, int inst) const {
    col(k, f, s, inst);
#pragma unroll
    for (int q = 0; q < 4; ++q) yout[q][i] = static_cast<To>(yin[q][i]);
  }
};

// ------------------------------------------------------------------------------------------
// Reductions and elementwise kernels.
// ------------------------------------------------------------------------------------------

// Per-instance totals of block partials, in a fixed order. Workgroup (q, i * parts + part) sums
// partial[q * stride + b] over b in ranges[i * parts + part] = [first, last) and writes
// out[(i * parts + part) * nsums + q].
static __global__ __launch_bounds__(kBlock) void finish_ranges_kernel(int stride, int nsums,
                                                               const double* __restrict__ partial,
                                                               const int2* __restrict__ ranges,
                                                               double* __restrict__ out) {
  __shared__ double wsum[kBlock / kWave];
  const int tid = threadIdx.x;
  const int q = blockIdx.x;
  const int seg = blockIdx.y;
  const int2 r = ranges[seg];
  const double* row = partial + static_cast<size_t>(q) * stride;
  double acc = 0.0;
  for (int b = r.x + tid; b < r.y; b += kBlock) acc += row[b];
  acc = group_sum<kWave>(acc);
  if ((tid & (kWave - 1)) == 0) wsum[tid / kWave] = acc;
  __syncthreads();
  if (tid == 0) out[static_cast<size_t>(seg) * nsums + q] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

// vbar = v + theta (v - v_prev): only when theta changed since vbar was formed.
template <typename T>
__global__ void extrapolate_kernel(int nt, const T* __restrict__ v, const T* __restrict__ v_prev,
                                   T theta, T* __restrict__ vbar) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nt) vbar[k] = v[k] + theta * (v[k] - v_prev[k]);
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
