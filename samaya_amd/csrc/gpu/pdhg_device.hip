// The PDHG device layer for gfx950 behind the C ABI of include/samaya_pdhg.h.
//
// A handle holds a batch of K LPs (K = 1 for the single-LP interface), stacked block-diagonally
// into one LP: the structural variables of all instances first (instance i's at x offset ox_i),
// then their logical variables / rows (instance i's rows at y offset oy_i, their logical variables
// at n + oy_i, n the total structural count). Device state (all resident in HBM for the solve):
//   A by rows (CSR: ptr int32[m+1], idx int32[nnz], val) = the CSC of each A' read as rows, and
//   A' by rows (ptr int32[n+1], ...) = the CSC of each A read as rows, with global indices; FP64
//   values always (the evaluation sums and the FP64 steps), FP32 values too once any instance
//   steps in FP32; each with its row-block schedule ((first row, first nonzero) int32 pairs,
//   nblocks + 1), the instance of every block, and per precision the list of blocks whose
//   instances step in it. No block holds rows of two instances.
//   Iterates in the step precision T: v[2] and y[2] (current / next, swapped by a parity bit
//   instead of pointers), vbar[2] (the extrapolated point of this and of the next step), the
//   running averages v_erg, y_erg, the restart point v_rst, y_rst, and cost/lower/upper.
//   Per instance: step scalars (tau, sigma, theta, steps since restart) and the precision of its
//   steps (or frozen: no further steps), so instances restart, change precision and stop on their
//   own schedule. The model-space data of evaluate (scales, cost and bounds before
//   equilibration), FP64.
//
// One step is two launches of csr_rows_kernel per precision in use: over the rows of A (dual
// update, fused with the primal update of the row's logical variable) and over the rows of A'
// (primal update of the structural variables). Each also forms the next step's extrapolated
// point, so the reference's separate extrapolate / dual / primal / average kernels and the ax /
// aty vectors disappear. An instance's arithmetic does not depend on the rest of its batch: a
// batch of K gives each LP the bits it gets alone.

#include <random>

#include "pdhg_handle.hpp"
#include "scaling_kernels.hpp"

namespace smy {
namespace {

// The stacked matrix by rows from each instance's CSC of the transpose (`rows_of` = the CSC whose
// columns are the rows), with the column indices moved by the instances' offsets, and its
// row-block schedule cut per instance.
void upload_stack(smy_pdhg* h, const smy_csc* parts, bool rows_are_a, DevMatrix& M) {
  const int K = h->K;
  long long nnz = 0;
  int rows = 0;
  for (int i = 0; i < K; ++i) {
    nnz += parts[i].start[parts[i].cols];
    rows += parts[i].cols;
  }
  M.rows = rows;
  M.nnz = nnz;
  std::vector<int> ptr(static_cast<size_t>(rows) + 1);
  std::vector<int> idx(static_cast<size_t>(nnz));
  std::vector<double> val(static_cast<size_t>(nnz));
  std::vector<RowBlockMeta> meta;
  std::vector<int> binst;
  int r = 0;
  long long z = 0;
  for (int i = 0; i < K; ++i) {
    const smy_csc& c = parts[i];
    Layout& L = h->inst[i];
    const int col_shift = rows_are_a ? L.ox : L.oy;
    const int row_first = r;
    std::vector<int> local(static_cast<size_t>(c.cols) + 1);
    for (int q = 0; q <= c.cols; ++q) local[q] = static_cast<int>(c.start[q]);
    for (int q = 0; q < c.cols; ++q) ptr[r + q] = static_cast<int>(z) + local[q];
    const long long cnt = c.start[c.cols];
    for (long long e = 0; e < cnt; ++e) {
      idx[z + e] = c.index[e] + col_shift;
      val[z + e] = c.value[e];
    }
    std::vector<RowBlockMeta> bl = build_row_blocks(local.data(), c.cols);
    const int first_block = static_cast<int>(meta.size());
    for (size_t q = 0; q + 1 < bl.size(); ++q) {
      meta.push_back({bl[q].row + row_first, bl[q].nz + static_cast<int>(z)});
      binst.push_back(i);
    }
    if (rows_are_a) {
      L.ba0 = first_block;
      L.ba1 = static_cast<int>(meta.size());
    } else {
      L.bt0 = first_block;
      L.bt1 = static_cast<int>(meta.size());
    }
    r += c.cols;
    z += cnt;
  }
  ptr[rows] = static_cast<int>(nnz);
  meta.push_back({rows, static_cast<int>(nnz)});
  M.nblocks = static_cast<int>(meta.size()) - 1;
  M.ptr = h->upload(ptr.data(), static_cast<long long>(ptr.size()));
  M.idx = h->upload(idx.data(), nnz);
  M.val64 = h->upload(val.data(), nnz);
  static_assert(sizeof(RowBlockMeta) == sizeof(int2), "schedule entries are read as int2");
  M.meta = h->upload(reinterpret_cast<const int2*>(meta.data()), static_cast<long long>(meta.size()));
  M.binst = h->upload(binst.data(), static_cast<long long>(binst.size()));
  M.single = K == 1;
  M.list[0] = h->alloc<int>(M.nblocks);
  M.list[1] = h->alloc<int>(M.nblocks);
}

// The block lists of both precisions from the instances' current precisions.
void rebuild_lists(smy_pdhg* h) {
  SMY_HIP(hipStreamSynchronize(h->stream));
  for (DevMatrix* M : {&h->a, &h->at}) {
    const bool is_a = M == &h->a;
    for (int q = 0; q < 2; ++q) {
      // The instances' blocks, size bin after size bin.
      std::vector<int> list;
      for (int bin = 0; bin < kTileBins; ++bin) {
        M->bin_start[q][bin] = static_cast<int>(list.size());
        for (const Layout& L : h->inst) {
          if (L.prec != q) continue;
          const int b0 = is_a ? L.ba0 : L.bt0, b1 = is_a ? L.ba1 : L.bt1;
          if (tile_bin(b1 - b0) != bin) continue;
          for (int b = b0; b < b1; ++b) list.push_back(b);
        }
      }
      M->bin_start[q][kTileBins] = static_cast<int>(list.size());
      M->nlist[q] = static_cast<int>(list.size());
      if (!list.empty())
        SMY_HIP(hipMemcpy(M->list[q], list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  std::vector<int> prec(static_cast<size_t>(h->K));
  for (int i = 0; i < h->K; ++i) prec[i] = h->inst[i].prec;
  SMY_HIP(hipMemcpy(h->step_prec, prec.data(), prec.size() * sizeof(int), hipMemcpyHostToDevice));
  ++h->list_gen;
  h->lists_dirty = false;
}

void refresh_lists(smy_pdhg* h) {
  if (h->lists_dirty) rebuild_lists(h);
}

template <typename T>
Iter<T>& state(smy_pdhg* h);
template <>
Iter<double>& state<double>(smy_pdhg* h) { return h->d; }
template <>
Iter<float>& state<float>(smy_pdhg* h) { return h->f; }

template <typename T>
const T* avals(const DevMatrix& M);
template <>
const double* avals<double>(const DevMatrix& M) { return M.val64; }
template <>
const float* avals<float>(const DevMatrix& M) { return M.val32; }

template <typename T>
void alloc_iter(smy_pdhg* h, Iter<T>& S) {
  const long long nt = h->nt, m = h->m;
  for (int q = 0; q < 2; ++q) {
    S.v[q] = h->zeros<T>(nt);
    S.y[q] = h->zeros<T>(m);
    S.vbar[q] = h->zeros<T>(nt);
  }
  S.v_erg = h->zeros<T>(nt);
  S.y_erg = h->zeros<T>(m);
  S.v_rst = h->zeros<T>(nt);
  S.y_rst = h->zeros<T>(m);
  S.cost = h->alloc<T>(nt);
  S.lower = h->alloc<T>(nt);
  S.upper = h->alloc<T>(nt);
  S.ready = true;
}

// FP32 values of both matrices, FP32 iterates, cost and bounds (from the FP64 ones).
void ensure_fp32(smy_pdhg* h) {
  if (h->f.ready) return;
  for (DevMatrix* M : {&h->a, &h->at}) {
    M->val32 = h->alloc<float>(M->nnz);
    if (M->nnz > 0)
      hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(M->nnz)), dim3(kBlock), 0,
                         h->stream, M->nnz, M->val64, M->val32);
  }
  alloc_iter<float>(h, h->f);
  for (auto [src, dst] : {std::pair{h->d.cost, h->f.cost}, std::pair{h->d.lower, h->f.lower},
                          std::pair{h->d.upper, h->f.upper}})
    if (h->nt > 0)
      hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(h->nt)), dim3(kBlock), 0,
                         h->stream, static_cast<long long>(h->nt), src, dst);
  SMY_HIP(hipGetLastError());
}

BlockRows block_rows(const smy_pdhg* h) {
  return BlockRows{h->a.meta, h->a.binst, h->a.nblocks, h->at.meta, h->at.binst, h->n};
}

int block_rows_grid(const smy_pdhg* h) { return h->a.nblocks + h->at.nblocks; }

int stride(const smy_pdhg* h) { return h->a.nblocks + h->at.nblocks; }

bool uses(const smy_pdhg* h, int q) {
  for (const Layout& L : h->inst)
    if (L.prec == q) return true;
  return false;
}

// A: the products' accumulator (T, or double for FP32 steps with FP64 accumulation).
template <typename T, typename A = T>
void step_kernels(smy_pdhg* h, int j) {
  constexpr int q = prec_of<T>();
  Iter<T>& S = state<T>(h);
  const int p = h->p;
  DualRowsOp<T, A> dual;
  dual.n = h->n;
  dual.j = j;
  dual.sc = h->scalars;
  dual.vbar = S.vbar[p];
  dual.y = S.y[p];
  dual.y_new = S.y[1 - p];
  dual.y_erg = S.y_erg;
  dual.v = S.v[p];
  dual.v_new = S.v[1 - p];
  dual.v_erg = S.v_erg;
  dual.vbar_next = S.vbar[1 - p];
  dual.cost = S.cost;
  dual.lower = S.lower;
  dual.upper = S.upper;
  launch_rows(h->a, avals<T>(h->a), h->a.list[q], h->a.nlist[q], dual, h->stream);
  PrimalColsOp<T, A> primal;
  primal.j = j;
  primal.sc = h->scalars;
  primal.y_new = S.y[1 - p];
  primal.v = S.v[p];
  primal.v_new = S.v[1 - p];
  primal.v_erg = S.v_erg;
  primal.vbar_next = S.vbar[1 - p];
  primal.cost = S.cost;
  primal.lower = S.lower;
  primal.upper = S.upper;
  launch_rows(h->at, avals<T>(h->at), h->at.list[q], h->at.nlist[q], primal, h->stream);
}

// The two kernels of a restarted Halpern step (csr_kernels.hpp): the primal one over the rows of
// A' first, then the dual one over the rows of A. `keep`: also write T(z) (the run's last step).
template <typename T, typename A = T>
struct HalpernOps {
  HalpernColsOp<T, A> primal;
  HalpernRowsOp<T, A> dual;
};

template <typename T, typename A = T>
HalpernOps<T, A> halpern_ops(smy_pdhg* h, int j, bool keep) {
  Iter<T>& S = state<T>(h);
  const int p = h->p;
  HalpernOps<T, A> o;
  o.primal.j = j;
  o.primal.keep = keep;
  o.primal.rho = h->rho;
  o.primal.sc = h->scalars;
  o.primal.y = S.y[p];
  o.primal.v = S.v[p];
  o.primal.v_anchor = S.v_rst;
  o.primal.v_next = S.v[1 - p];
  o.primal.w = S.vbar[0];
  o.primal.v_plus = S.v_erg;
  o.primal.cost = S.cost;
  o.primal.lower = S.lower;
  o.primal.upper = S.upper;
  o.dual.n = h->n;
  o.dual.j = j;
  o.dual.keep = keep;
  o.dual.rho = h->rho;
  o.dual.sc = h->scalars;
  o.dual.w = S.vbar[0];
  o.dual.y = S.y[p];
  o.dual.y_anchor = S.y_rst;
  o.dual.y_next = S.y[1 - p];
  o.dual.y_plus = S.y_erg;
  o.dual.v = S.v[p];
  o.dual.v_anchor = S.v_rst;
  o.dual.v_next = S.v[1 - p];
  o.dual.v_plus = S.v_erg;
  o.dual.cost = S.cost;
  o.dual.lower = S.lower;
  o.dual.upper = S.upper;
  return o;
}

template <typename T, typename A = T>
void halpern_kernels(smy_pdhg* h, int j, bool keep) {
  constexpr int q = prec_of<T>();
  const HalpernOps<T, A> o = halpern_ops<T, A>(h, j, keep);
  launch_rows(h->at, avals<T>(h->at), h->at.list[q], h->at.nlist[q], o.primal, h->stream);
  launch_rows(h->a, avals<T>(h->a), h->a.list[q], h->a.nlist[q], o.dual, h->stream);
}

// ---- Tile tuner ----------------------------------------------------------------------------
// The first run in a precision times the candidate tiles of both matrices with the values of that
// precision (over the blocks of the instances stepping in it) on a stand-in for the step kernels
// (TuneOp: the product plus an epilogue's reads and writes, into scratch), and keeps the fastest;
// the plain one-block-per-workgroup tile unless another is at least 5% faster. Every tile sums
// every row in the same order, so the choice changes speed only: answers do not depend on the
// timing.
constexpr Tile kTiles[] = {{1, false}, {1, true}, {2, false}, {2, true}, {4, false}, {4, true}};

template <typename T>
void tune_matrix(smy_pdhg* h, DevMatrix& M, const T* x, const T* const (&in)[5]) {
  constexpr int q = prec_of<T>();
  if (M.tuned[q] || M.nlist[q] == 0) return;
  // A launch of fewer blocks than this lasts a few microseconds, where timing noise outweighs any
  // tile's gain and fewer workgroups only lose parallelism: such a matrix keeps the plain tile.
  constexpr int kTuneMinBlocks = 2048;
  if (M.nlist[q] < kTuneMinBlocks) {
    M.tile[q] = Tile{};
    M.tuned[q] = true;
    return;
  }
  const long long rows = std::max(h->m, h->n);
  if (h->tune_out == nullptr) h->tune_out = h->alloc<double>(3 * std::max(rows, 1LL));
  TuneOp<T> op;
  op.x = x;
  for (int k = 0; k < 5; ++k) op.in[k] = in[k];
  for (int k = 0; k < 3; ++k) op.out[k] = reinterpret_cast<T*>(h->tune_out) + k * rows;
  // 5 launches per tile, and a tile must be 5% faster than the plain one (or the bins than one
  // launch): closer calls are within the timing noise of this stand-in, and a pick by noise would
  // let two handles of the same LP launch differently.
  constexpr int kReps = 5;
  constexpr float kMargin = 0.95f;
  // The fastest tile on blocks [first, first + count) of the list: the plain one unless another
  // is faster by the margin. Returns its time (ms for kReps launches).
  const auto best_tile = [&](int first, int count, Tile& pick) {
    float base = 0.0f, best = 0.0f;
    pick = Tile{};
    for (const Tile& t : kTiles) {
      launch_rows(M, avals<T>(M), M.list[q] + first, count, op, h->stream, &t);  // warm
      SMY_HIP(hipEventRecord(h->ev_start, h->stream));
      for (int r = 0; r < kReps; ++r) launch_rows(M, avals<T>(M), M.list[q] + first, count, op, h->stream, &t);
      SMY_HIP(hipEventRecord(h->ev_stop, h->stream));
      SMY_HIP(hipEventSynchronize(h->ev_stop));
      float ms = 0.0f;
      SMY_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
      if (&t == kTiles) {
        base = best = ms;
      } else if (ms < best && ms < kMargin * base) {
        best = ms;
        pick = t;
      }
    }
    return best;
  };
  Tile pick;
  const float whole = best_tile(0, M.nlist[q], pick);
  M.tile[q] = pick;
  M.split[q] = false;
  // Size bins launched apart, each with its own fastest tile (ranked on the bin alone), when
  // the bins launched one after the other beat one launch of all blocks by the margin. That comparison
  // times the whole sequence: a bin timed alone keeps its smaller data in cache between repeats.
  int filled = 0;
  for (int b = 0; b < kTileBins; ++b) filled += M.bin_start[q][b + 1] > M.bin_start[q][b];
  if (filled > 1) {
    for (int b = 0; b < kTileBins; ++b) {
      M.bin_tile[q][b] = Tile{};
      const int first = M.bin_start[q][b], count = M.bin_start[q][b + 1] - first;
      if (count > 0) best_tile(first, count, M.bin_tile[q][b]);
    }
    M.split[q] = true;
    launch_rows(M, avals<T>(M), M.list[q], M.nlist[q], op, h->stream);  // warm
    SMY_HIP(hipEventRecord(h->ev_start, h->stream));
    for (int r = 0; r < kReps; ++r) launch_rows(M, avals<T>(M), M.list[q], M.nlist[q], op, h->stream);
    SMY_HIP(hipEventRecord(h->ev_stop, h->stream));
    SMY_HIP(hipEventSynchronize(h->ev_stop));
    float apart = 0.0f;
    SMY_HIP(hipEventElapsedTime(&apart, h->ev_start, h->ev_stop));
    M.split[q] = apart < kMargin * whole;
  }
  M.tuned[q] = true;
}

// Before any capture: tiles of the precisions in use that have none yet.
void tune(smy_pdhg* h) {
  if (!h->autotune) return;
  const bool any = (uses(h, SMY_FP32) && !(h->a.tuned[SMY_FP32] && h->at.tuned[SMY_FP32])) ||
                   (uses(h, SMY_FP64) && !(h->a.tuned[SMY_FP64] && h->at.tuned[SMY_FP64]));
  if (!any) return;
  // Epilogue inputs: any five vectors of at least max(m, n) entries of the state.
  const auto go = [&](auto& S) {
    using T = std::remove_pointer_t<decltype(S.cost)>;
    const T* const in[5] = {S.v[0], S.v[1], S.cost, S.lower, S.upper};
    tune_matrix<T>(h, h->a, S.v[h->p], in);
    tune_matrix<T>(h, h->at, S.y[h->p], in);
  };
  if (uses(h, SMY_FP32)) go(h->f);
  if (uses(h, SMY_FP64)) go(h->d);
  ++h->list_gen;  // graphs captured before hold the old launch shapes
  h->timed = false;
}

// Step j of a run of `count` (Halpern steps write T(z) at the run's last step).
void step_all(smy_pdhg* h, int j, int count = 1) {
  if (h->halpern) {
    const bool keep = j + 1 == count;
    if (uses(h, SMY_FP32)) {
      if (h->acc64)
        halpern_kernels<float, double>(h, j, keep);
      else
        halpern_kernels<float>(h, j, keep);
    }
    if (uses(h, SMY_FP64)) halpern_kernels<double>(h, j, keep);
    return;
  }
  if (uses(h, SMY_FP32)) {
    if (h->acc64)
      step_kernels<float, double>(h, j);
    else
      step_kernels<float>(h, j);
  }
  if (uses(h, SMY_FP64)) step_kernels<double>(h, j);
}

// Re-forms vbar for a theta other than the one it was formed with (the reference recomputes it
// from v and v_prev at every step; here only a change of theta needs it).
int refresh_vbar(smy_pdhg* h, double theta) {
  if (h->vbar_any || theta == h->theta_vbar) return SMY_OK;
  if (h->stepped) return fail(SMY_STATE, "theta changed between two steps without advance()");
  const auto go = [&](auto& S) {
    using T = std::remove_pointer_t<decltype(S.cost)>;
    hipLaunchKernelGGL((extrapolate_kernel<T>), dim3(blocks_for(h->nt)), dim3(kBlock), 0, h->stream,
                       h->nt, S.v[h->p], S.v[1 - h->p], static_cast<T>(theta), S.vbar[h->p]);
  };
  if (h->nt > 0) {
    if (h->d.ready) go(h->d);
    if (h->f.ready) go(h->f);
  }
  return SMY_OK;
}

void set_scalars(smy_pdhg* h, const double* tau, const double* sigma, double theta,
                 const long long* k0) {
  staging_wait(h);
  for (int i = 0; i < h->K; ++i) h->scalars_host[i] = StepScalars{tau[i], sigma[i], theta, k0[i]};
  SMY_HIP(hipMemcpyAsync(h->scalars, h->scalars_host, h->K * sizeof(StepScalars),
                         hipMemcpyHostToDevice, h->stream));
  staging_sent(h);
}

int step_impl(smy_pdhg* h, const double* tau, const double* sigma, double theta,
              const long long* k0) {
  h->fixed_fresh = false;
  refresh_lists(h);
  const int rc = refresh_vbar(h, theta);
  if (rc != SMY_OK) return rc;
  set_scalars(h, tau, sigma, theta, k0);
  step_all(h, 0);
  SMY_HIP(hipGetLastError());
  h->theta_vbar = theta;
  std::fill(h->ax_stale.begin(), h->ax_stale.end(), 1);
  h->vbar_any = false;
  h->stepped = true;
  return SMY_OK;
}

int run_impl(smy_pdhg* h, const double* tau, const double* sigma, double theta,
             const long long* k0, int count) {
  if (count <= 0) return SMY_OK;
  h->fixed_fresh = false;
  refresh_lists(h);
  tune(h);
  const int rc = refresh_vbar(h, theta);
  if (rc != SMY_OK) return rc;
  SMY_HIP(hipEventRecord(h->ev_start, h->stream));
  set_scalars(h, tau, sigma, theta, k0);
  const int mask = (uses(h, SMY_FP64) ? 1 : 0) | (uses(h, SMY_FP32) ? 2 : 0);
  if (count == smy_pdhg::kGraphSteps && mask != 0) {
    const int p0 = h->p;
    smy_pdhg::Captured& c = (h->halpern ? h->hgraphs : h->graphs)[mask][p0];
    if (c.exec == nullptr || c.gen != h->list_gen) {
      h->drop_graph(c);
      SMY_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      for (int j = 0; j < count; ++j) {
        step_all(h, j, count);
        if (j + 1 < count) h->p ^= 1;
      }
      SMY_HIP(hipStreamEndCapture(h->stream, &c.graph));
      h->p = p0;
      SMY_HIP(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
      c.gen = h->list_gen;
    }
    SMY_HIP(hipGraphLaunch(c.exec, h->stream));
    if ((count - 1) & 1) h->p ^= 1;
  } else {
    for (int j = 0; j < count; ++j) {
      step_all(h, j, count);
      if (j + 1 < count) h->p ^= 1;
    }
  }
  SMY_HIP(hipGetLastError());
  SMY_HIP(hipEventRecord(h->ev_stop, h->stream));
  h->timed = true;
  h->theta_vbar = theta;
  std::fill(h->ax_stale.begin(), h->ax_stale.end(), 1);
  h->vbar_any = false;
  h->stepped = true;
  return SMY_OK;
}

// ---- Adaptive steps ----------------------------------------------------------------------

template <typename T>
void ensure_adaptive_state(smy_pdhg* h, Iter<T>& S) {
  if (!S.ready || S.ax[0] != nullptr) return;
  S.ax[0] = h->zeros<T>(h->m);
  S.ax[1] = h->zeros<T>(h->m);
  S.v_erg2 = h->zeros<T>(h->nt);
  S.y_erg2 = h->zeros<T>(h->m);
}

void ensure_adaptive(smy_pdhg* h) {
  if (h->adapt == nullptr) {
    h->adapt = h->alloc<AdaptState>(h->K);
    h->ax_flags = h->alloc<int>(h->K);
    h->decide_slices = h->alloc<double>(3LL * kDecideSlices * h->K);
    h->decide_arrivals = h->zeros<unsigned>(h->K);
    // Slices of the decision: about kDecideBlocksPerSlice partials each for the largest instance.
    int most = 0;
    for (const Layout& L : h->inst) most = std::max({most, L.ba1 - L.ba0, L.bt1 - L.bt0});
    h->decide_nslices =
        std::clamp((most + kDecideBlocksPerSlice - 1) / kDecideBlocksPerSlice, 1, kDecideSlices);
    SMY_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->adapt_host), sizeof(AdaptState) * h->K,
                          hipHostMallocDefault));
  }
  ensure_adaptive_state(h, h->d);
  ensure_adaptive_state(h, h->f);
}

// ax[p] = A x of the current iterates of the flagged instances (after a start, a restart or a
// change of precision).
template <typename T>
void refresh_ax(smy_pdhg* h) {
  constexpr int q = prec_of<T>();
  Iter<T>& S = state<T>(h);
  AxRefreshOp<T> op{h->ax_flags, S.v[h->p], S.ax[h->p]};
  launch_rows(h->a, avals<T>(h->a), h->a.list[q], h->a.nlist[q], op, h->stream);
}

template <typename T, typename A = T>
void adapt_primal(smy_pdhg* h) {
  constexpr int q = prec_of<T>();
  Iter<T>& S = state<T>(h);
  AdaptPrimalOp<T, A> primal;
  primal.st = h->adapt;
  primal.y0 = S.y[0];
  primal.y1 = S.y[1];
  primal.v0 = S.v[0];
  primal.v1 = S.v[1];
  primal.e0 = S.v_erg;
  primal.e1 = S.v_erg2;
  primal.xbar = S.vbar[0];
  primal.cost = S.cost;
  primal.lower = S.lower;
  primal.upper = S.upper;
  primal.partial = h->partial + h->a.nblocks;
  primal.stride = stride(h);
  launch_rows(h->at, avals<T>(h->at), h->at.list[q], h->at.nlist[q], primal, h->stream);
}

template <typename T, typename A = T>
void adapt_dual(smy_pdhg* h) {
  constexpr int q = prec_of<T>();
  Iter<T>& S = state<T>(h);
  AdaptDualOp<T, A> dual;
  dual.n = h->n;
  dual.st = h->adapt;
  dual.xbar = S.vbar[0];
  dual.y0 = S.y[0];
  dual.y1 = S.y[1];
  dual.ye0 = S.y_erg;
  dual.ye1 = S.y_erg2;
  dual.v0 = S.v[0];
  dual.v1 = S.v[1];
  dual.ve0 = S.v_erg;
  dual.ve1 = S.v_erg2;
  dual.h0 = S.ax[0];
  dual.h1 = S.ax[1];
  dual.cost = S.cost;
  dual.lower = S.lower;
  dual.upper = S.upper;
  dual.partial = h->partial;
  dual.stride = stride(h);
  launch_rows(h->a, avals<T>(h->a), h->a.list[q], h->a.nlist[q], dual, h->stream);
}

void adapt_decide(smy_pdhg* h) {
  hipLaunchKernelGGL(adapt_decide_kernel, dim3(h->decide_nslices, h->K), dim3(kDecideThreads), 0,
                     h->stream, stride(h), h->partial, h->ranges, h->step_prec, h->adapt,
                     h->decide_nslices, h->decide_slices, h->decide_arrivals);
}

void adapt_attempt(smy_pdhg* h) {
  if (uses(h, SMY_FP32)) {
    if (h->acc64) {
      adapt_primal<float, double>(h);
      adapt_dual<float, double>(h);
    } else {
      adapt_primal<float>(h);
      adapt_dual<float>(h);
    }
  }
  if (uses(h, SMY_FP64)) {
    adapt_primal<double>(h);
    adapt_dual<double>(h);
  }
  adapt_decide(h);
}

// A x of the instances whose iterate changed outside the attempts.
void refresh_stale_ax(smy_pdhg* h, int mask) {
  if (std::none_of(h->ax_stale.begin(), h->ax_stale.end(), [](int f) { return f != 0; })) return;
  SMY_HIP(hipMemcpyAsync(h->ax_flags, h->ax_stale.data(), h->K * sizeof(int), hipMemcpyHostToDevice,
                         h->stream));
  if (mask & 2) refresh_ax<float>(h);
  if (mask & 1) refresh_ax<double>(h);
  SMY_HIP(hipStreamSynchronize(h->stream));  // the flags' host copy may change after this
  std::fill(h->ax_stale.begin(), h->ax_stale.end(), 0);
}

// Measurement helper: average milliseconds of the adaptive primal (rows of A'), dual (rows of A)
// and decision kernels of the first precision in use, from a small step size. Changes the state.
template <typename T, typename A = T>
int time_adaptive_impl(smy_pdhg* h, int reps, double* dual_ms, double* primal_ms, double* decide_ms) {
  ensure_adaptive(h);
  staging_wait(h);
  for (int i = 0; i < h->K; ++i) h->adapt_host[i] = AdaptState{1e-3, 1.0, 0.0, 0, 0, h->p, 0};
  SMY_HIP(hipMemcpyAsync(h->adapt, h->adapt_host, h->K * sizeof(AdaptState), hipMemcpyHostToDevice,
                         h->stream));
  staging_sent(h);
  std::fill(h->ax_stale.begin(), h->ax_stale.end(), 1);
  refresh_stale_ax(h, prec_of<T>() == SMY_FP32 ? 2 : 1);
  *primal_ms = time_repeated(h, reps, [&] { adapt_primal<T, A>(h); });
  *dual_ms = time_repeated(h, reps, [&] { adapt_dual<T, A>(h); });
  *decide_ms = time_repeated(h, reps, [&] { adapt_decide(h); });
  SMY_HIP(hipGetLastError());
  std::fill(h->ax_stale.begin(), h->ax_stale.end(), 1);
  h->timed = false;
  return SMY_OK;
}

template <typename T>
void adapt_settle(smy_pdhg* h, int want) {
  Iter<T>& S = state<T>(h);
  AdaptSettleOp<T> op;
  op.st = h->adapt;
  op.prec = h->step_prec;
  op.q = prec_of<T>();
  op.want = want;
  op.v0 = S.v[0];
  op.v1 = S.v[1];
  op.y0 = S.y[0];
  op.y1 = S.y[1];
  op.h0 = S.ax[0];
  op.h1 = S.ax[1];
  op.ve0 = S.v_erg;
  op.ve1 = S.v_erg2;
  op.ye0 = S.y_erg;
  op.ye1 = S.y_erg2;
  hipLaunchKernelGGL((block_rows_kernel<AdaptSettleOp<T>>), dim3(block_rows_grid(h)), dim3(kBlock), 0,
                     h->stream, block_rows(h), op);
}

// `count` attempts, then every instance's current iterate moved to slot 1 - p of the final
// parity (as after `count` constant steps). The state goes up before and comes back after.
int run_adaptive_impl(smy_pdhg* h, smy_pdhg_adaptive* st, int count) {
  if (count <= 0) return SMY_OK;
  for (int i = 0; i < h->K; ++i)
    if (h->inst[i].prec != kFrozen && !(st[i].eta > 0.0 && st[i].omega > 0.0))
      return fail(SMY_INVALID, "adaptive steps need eta > 0 and omega > 0");
  refresh_lists(h);
  ensure_adaptive(h);
  tune(h);
  const int mask = (uses(h, SMY_FP64) ? 1 : 0) | (uses(h, SMY_FP32) ? 2 : 0);
  refresh_stale_ax(h, mask);
  staging_wait(h);
  const int p0 = h->p;
  for (int i = 0; i < h->K; ++i)
    h->adapt_host[i] = AdaptState{st[i].eta, st[i].omega, st[i].weight, st[i].attempts, st[i].rejected, p0, 0};
  SMY_HIP(hipMemcpyAsync(h->adapt, h->adapt_host, h->K * sizeof(AdaptState), hipMemcpyHostToDevice,
                         h->stream));
  staging_sent(h);
  SMY_HIP(hipEventRecord(h->ev_start, h->stream));
  const int p_end = p0 ^ ((count - 1) & 1);
  const auto all = [&] {
    if (h->nt > 0)
      for (int j = 0; j < count; ++j) adapt_attempt(h);
    if (block_rows_grid(h) > 0) {
      if (mask & 2) adapt_settle<float>(h, 1 - p_end);
      if (mask & 1) adapt_settle<double>(h, 1 - p_end);
    }
  };
  if (count == smy_pdhg::kGraphSteps && mask != 0) {
    smy_pdhg::Captured& c = h->agraphs[mask][p0];
    if (c.exec == nullptr || c.gen != h->list_gen) {
      h->drop_graph(c);
      SMY_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      all();
      SMY_HIP(hipStreamEndCapture(h->stream, &c.graph));
      SMY_HIP(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
      c.gen = h->list_gen;
    }
    SMY_HIP(hipGraphLaunch(c.exec, h->stream));
  } else if (mask != 0) {
    all();
  }
  SMY_HIP(hipGetLastError());
  SMY_HIP(hipEventRecord(h->ev_stop, h->stream));
  SMY_HIP(hipMemcpyAsync(h->adapt_host, h->adapt, h->K * sizeof(AdaptState), hipMemcpyDeviceToHost,
                         h->stream));
  SMY_HIP(hipStreamSynchronize(h->stream));
  h->staged = false;
  for (int i = 0; i < h->K; ++i) {
    if (h->inst[i].prec == kFrozen) continue;
    const AdaptState& s = h->adapt_host[i];
    st[i].eta = s.eta;
    st[i].weight = s.weight;
    st[i].attempts = s.total;
    st[i].rejected = s.rejected;
  }
  h->p = p_end;
  h->timed = true;
  // vbar holds 2 v+ - v of the last attempt, not the constant steps' extrapolation.
  h->vbar_any = false;
  h->theta_vbar = std::nan("");
  h->stepped = true;
  return SMY_OK;
}

// Average duration of each of the two step kernels of the first precision in use (FP32 before
// FP64) over `reps` back-to-back launches (HIP events on the handle's stream). The launches repeat
// one step from the current state, so the iterates and averages change: a measurement helper,
// not part of a solve.
template <typename T, typename A = T>
int time_kernels_impl(smy_pdhg* h, int reps, double* dual_ms, double* primal_ms) {
  constexpr int q = prec_of<T>();
  Iter<T>& S = state<T>(h);
  const int p = h->p;
  std::vector<double> tau(h->K, 0.5), sigma(h->K, 0.5);
  std::vector<long long> k0(h->K, 0);
  set_scalars(h, tau.data(), sigma.data(), 1.0, k0.data());
  if (h->halpern) {
    const HalpernOps<T, A> o = halpern_ops<T, A>(h, 0, false);
    *dual_ms = time_repeated(h, reps, [&] {
      launch_rows(h->a, avals<T>(h->a), h->a.list[q], h->a.nlist[q], o.dual, h->stream);
    });
    *primal_ms = time_repeated(h, reps, [&] {
      launch_rows(h->at, avals<T>(h->at), h->at.list[q], h->at.nlist[q], o.primal, h->stream);
    });
    SMY_HIP(hipGetLastError());
    h->timed = false;
    return SMY_OK;
  }
  DualRowsOp<T, A> dual;
  dual.n = h->n;
  dual.j = 0;
  dual.sc = h->scalars;
  dual.vbar = S.vbar[p];
  dual.y = S.y[p];
  dual.y_new = S.y[1 - p];
  dual.y_erg = S.y_erg;
  dual.v = S.v[p];
  dual.v_new = S.v[1 - p];
  dual.v_erg = S.v_erg;
  dual.vbar_next = S.vbar[1 - p];
  dual.cost = S.cost;
  dual.lower = S.lower;
  dual.upper = S.upper;
  PrimalColsOp<T, A> primal;
  primal.j = 0;
  primal.sc = h->scalars;
  primal.y_new = S.y[1 - p];
  primal.v = S.v[p];
  primal.v_new = S.v[1 - p];
  primal.v_erg = S.v_erg;
  primal.vbar_next = S.vbar[1 - p];
  primal.cost = S.cost;
  primal.lower = S.lower;
  primal.upper = S.upper;
  const auto dual_go = [&] { launch_rows(h->a, avals<T>(h->a), h->a.list[q], h->a.nlist[q], dual, h->stream); };
  const auto primal_go = [&] {
    launch_rows(h->at, avals<T>(h->at), h->at.list[q], h->at.nlist[q], primal, h->stream);
  };
  *dual_ms = time_repeated(h, reps, dual_go);
  *primal_ms = time_repeated(h, reps, primal_go);
  SMY_HIP(hipGetLastError());
  h->timed = false;
  return SMY_OK;
}

// Block partials of the evaluation sums of NV iterates of the instances stepping in precision T.
template <typename T, int NV>
void measure_launch(smy_pdhg* h, const T* const (&vv)[NV], const T* const (&yy)[NV]) {
  constexpr int q = prec_of<T>();
  MeasureRowsOp<T, NV> rows;
  rows.n = h->n;
  for (int v = 0; v < NV; ++v) {
    rows.vv[v] = vv[v];
    rows.yy[v] = yy[v];
  }
  rows.rs = h->rs;
  rows.cost = h->cost_o;
  rows.lower = h->lower_o;
  rows.upper = h->upper_o;
  rows.partial = h->partial;
  rows.stride = stride(h);
  launch_rows(h->a, h->a.val64, h->a.list[q], h->a.nlist[q], rows, h->stream);
  MeasureColsOp<T, NV> cols;
  for (int v = 0; v < NV; ++v) {
    cols.vv[v] = vv[v];
    cols.yy[v] = yy[v];
  }
  cols.cs = h->cs;
  cols.cost = h->cost_o;
  cols.lower = h->lower_o;
  cols.upper = h->upper_o;
  cols.partial = h->partial + h->a.nblocks;
  cols.stride = stride(h);
  launch_rows(h->at, h->at.val64, h->at.list[q], h->at.nlist[q], cols, h->stream);
}

// Per-instance totals of `nsums` block partials into out[i * nsums + q]: the instance's A blocks
// first, then its A' blocks, added on the host in that order. Synchronizes. Entries of frozen
// instances are meaningless.
void finish(smy_pdhg* h, int nsums, double* out) {
  hipLaunchKernelGGL(finish_ranges_kernel, dim3(nsums, 2 * h->K), dim3(kBlock), 0, h->stream,
                     stride(h), nsums, h->partial, h->ranges, h->sums_dev);
  SMY_HIP(hipGetLastError());
  SMY_HIP(hipMemcpyAsync(h->sums_host, h->sums_dev, 2 * h->K * nsums * sizeof(double),
                         hipMemcpyDeviceToHost, h->stream));
  SMY_HIP(hipStreamSynchronize(h->stream));
  for (int i = 0; i < h->K; ++i)
    for (int q = 0; q < nsums; ++q)
      out[i * nsums + q] = h->sums_host[(2 * i) * nsums + q] + h->sums_host[(2 * i + 1) * nsums + q];
}

template <typename T>
void measure_both(smy_pdhg* h) {
  Iter<T>& S = state<T>(h);
  const T* const vv[2] = {S.v[1 - h->p], S.v_erg};
  const T* const yy[2] = {S.y[1 - h->p], S.y_erg};
  measure_launch<T, 2>(h, vv, yy);
}

template <typename T>
void measure_one(smy_pdhg* h, int which);
void stage_flags(smy_pdhg* h, const int* value);
template <typename T>
void residual_launch(smy_pdhg* h, int row0 = 0);

int evaluate_batch_impl(smy_pdhg* h, double* out) {
  for (const Layout& L : h->inst)
    if (L.prec != kFrozen && !L.has_original)
      return fail(SMY_STATE, "set_original() must be called before evaluate()");
  refresh_lists(h);
  if (h->halpern) {
    // Halpern steps: only T(z) (the averages' slot) is an answer (the iterate z' itself may lie
    // outside the bounds); its sums fill both halves. The fixed-point residual of the run's last
    // step is summed in the same pass (partial rows 7-8) and kept for smy_pdhg_fixed_point_batch,
    // so the restart test costs no second synchronization.
    if (uses(h, SMY_FP32)) measure_one<float>(h, SMY_AVERAGE);
    if (uses(h, SMY_FP64)) measure_one<double>(h, SMY_AVERAGE);
    std::vector<int> all(static_cast<size_t>(h->K), 0);
    stage_flags(h, all.data());
    if (uses(h, SMY_FP32)) residual_launch<float>(h, 7);
    if (uses(h, SMY_FP64)) residual_launch<double>(h, 7);
    SMY_HIP(hipGetLastError());
    std::vector<double> nine(9 * static_cast<size_t>(h->K));
    finish(h, 9, nine.data());
    h->fixed_cache.resize(2 * static_cast<size_t>(h->K));
    for (int i = 0; i < h->K; ++i) {
      for (int q = 0; q < 7; ++q) out[14 * i + q] = out[14 * i + 7 + q] = nine[9 * i + q];
      h->fixed_cache[2 * i] = nine[9 * i + 7];
      h->fixed_cache[2 * i + 1] = nine[9 * i + 8];
    }
    h->fixed_fresh = true;
    return SMY_OK;
  }
  if (uses(h, SMY_FP32)) measure_both<float>(h);
  if (uses(h, SMY_FP64)) measure_both<double>(h);
  SMY_HIP(hipGetLastError());
  finish(h, 14, out);
  return SMY_OK;
}

template <typename T>
void measure_one(smy_pdhg* h, int which) {
  Iter<T>& S = state<T>(h);
  const T* const vv[1] = {which == SMY_CURRENT ? S.v[1 - h->p] : S.v_erg};
  const T* const yy[1] = {which == SMY_CURRENT ? S.y[1 - h->p] : S.y_erg};
  measure_launch<T, 1>(h, vv, yy);
}

// Stages per-precision instance flags: flags[q * K + i] = value[i] if instance i's state is in
// precision q (and it is not frozen), else -1.
void stage_flags(smy_pdhg* h, const int* value) {
  staging_wait(h);
  for (int q = 0; q < 2; ++q)
    for (int i = 0; i < h->K; ++i)
      h->flags_host[q * h->K + i] = h->inst[i].prec == q ? value[i] : -1;
  SMY_HIP(hipMemcpyAsync(h->flags, h->flags_host, 2 * h->K * sizeof(int), hipMemcpyHostToDevice,
                         h->stream));
  staging_sent(h);
}

bool any_flag(const smy_pdhg* h, const int* which, int q) {
  for (int i = 0; i < h->K; ++i)
    if (which[i] >= 0 && h->inst[i].prec == q) return true;
  return false;
}

template <typename T>
void distance_launch(smy_pdhg* h) {
  constexpr int q = prec_of<T>();
  Iter<T>& S = state<T>(h);
  DistanceOp<T> op;
  op.which = h->flags + q * h->K;
  op.v_new = S.v[1 - h->p];
  op.y_new = S.y[1 - h->p];
  op.v_erg = S.v_erg;
  op.y_erg = S.y_erg;
  op.v_rst = S.v_rst;
  op.y_rst = S.y_rst;
  op.partial = h->partial;
  op.stride = stride(h);
  hipLaunchKernelGGL((block_rows_kernel<DistanceOp<T>>), dim3(block_rows_grid(h)), dim3(kBlock), 0,
                     h->stream, block_rows(h), op);
}

// Partial rows `row0` and row0 + 1 of the block partials (the evaluation passes 7: its own sums
// take rows 0-6).
template <typename T>
void residual_launch(smy_pdhg* h, int row0) {
  constexpr int q = prec_of<T>();
  Iter<T>& S = state<T>(h);
  ResidualOp<T> op;
  op.which = h->flags + q * h->K;
  op.v = S.v[h->p];
  op.y = S.y[h->p];
  op.v_plus = S.v_erg;
  op.y_plus = S.y_erg;
  op.partial = h->partial + static_cast<size_t>(row0) * stride(h);
  op.stride = stride(h);
  hipLaunchKernelGGL((block_rows_kernel<ResidualOp<T>>), dim3(block_rows_grid(h)), dim3(kBlock), 0,
                     h->stream, block_rows(h), op);
}

// After a run of Halpern steps: per instance ||v+ - v||^2, ||y+ - y||^2 of its last step.
int fixed_point_batch_impl(smy_pdhg* h, double* out) {
  if (h->fixed_fresh) {  // formed with the evaluation since the last run
    std::copy(h->fixed_cache.begin(), h->fixed_cache.end(), out);
    return SMY_OK;
  }
  if (block_rows_grid(h) == 0) {
    std::fill(out, out + 2 * h->K, 0.0);
    return SMY_OK;
  }
  std::vector<int> all(static_cast<size_t>(h->K), 0);
  stage_flags(h, all.data());
  if (uses(h, SMY_FP32)) residual_launch<float>(h);
  if (uses(h, SMY_FP64)) residual_launch<double>(h);
  SMY_HIP(hipGetLastError());
  finish(h, 2, out);
  return SMY_OK;
}

int distance_batch_impl(smy_pdhg* h, const int* which, double* out) {
  if (block_rows_grid(h) == 0) {
    std::fill(out, out + 2 * h->K, 0.0);
    return SMY_OK;
  }
  stage_flags(h, which);
  if (any_flag(h, which, SMY_FP32)) distance_launch<float>(h);
  if (any_flag(h, which, SMY_FP64)) distance_launch<double>(h);
  SMY_HIP(hipGetLastError());
  finish(h, 2, out);
  return SMY_OK;
}

template <typename T>
void restart_launch(smy_pdhg* h) {
  constexpr int q = prec_of<T>();
  Iter<T>& S = state<T>(h);
  const int p = h->p;
  RestartOp<T> op;
  op.which = h->flags + q * h->K;
  op.v = S.v[p];
  op.v_new = S.v[1 - p];
  op.y = S.y[p];
  op.y_new = S.y[1 - p];
  op.v_erg = S.v_erg;
  op.y_erg = S.y_erg;
  op.v_rst = S.v_rst;
  op.y_rst = S.y_rst;
  op.vbar0 = S.vbar[0];
  op.vbar1 = S.vbar[1];
  hipLaunchKernelGGL((block_rows_kernel<RestartOp<T>>), dim3(block_rows_grid(h)), dim3(kBlock), 0,
                     h->stream, block_rows(h), op);
}

int restart_to_batch_impl(smy_pdhg* h, const int* which) {
  h->fixed_fresh = false;
  if (block_rows_grid(h) == 0) return SMY_OK;
  stage_flags(h, which);
  if (any_flag(h, which, SMY_FP32)) restart_launch<float>(h);
  if (any_flag(h, which, SMY_FP64)) restart_launch<double>(h);
  SMY_HIP(hipGetLastError());
  h->vbar_any = true;
  for (int i = 0; i < h->K; ++i)
    if (which[i] >= 0) h->ax_stale[i] = 1;
  h->stepped = false;
  return SMY_OK;
}

// ---- Restart rule on the device ------------------------------------------------------------

int rule_init_impl(smy_pdhg* h, const smy_pdhg_rule_setup* init) {
  if (h->rule == nullptr) {
    h->rule = h->alloc<RuleState>(h->K);
    h->rule_in = h->alloc<RuleInput>(h->K);
    h->rule_dist = h->alloc<double>(4LL * h->K);
    SMY_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->rule_host), sizeof(RuleState) * h->K,
                          hipHostMallocDefault));
    SMY_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->rule_in_host), sizeof(RuleInput) * h->K,
                          hipHostMallocDefault));
  }
  SMY_HIP(hipStreamSynchronize(h->stream));
  for (int i = 0; i < h->K; ++i) {
    const smy_pdhg_rule_setup& r = init[i];
    if (!(r.omega > 0.0)) return fail(SMY_INVALID, "the primal weight must be positive");
    h->rule_host[i] = RuleState{r.a_colmax, r.at_rowmax, r.c_norm, r.has_rows, -1, r.omega,
                                INFINITY, INFINITY, 0.0, 0};
  }
  SMY_HIP(hipMemcpy(h->rule, h->rule_host, sizeof(RuleState) * h->K, hipMemcpyHostToDevice));
  return SMY_OK;
}

// Decision, distances, primal weight and restart of every due instance, on the device, from the
// totals the last evaluation left in sums_dev (the A part, then the A' part of each instance).
int rule_decide_impl(smy_pdhg* h, const smy_pdhg_rule_in* in, long long iter1, smy_pdhg_rule_out* out) {
  if (h->rule == nullptr) return fail(SMY_STATE, "smy_pdhg_rule_init() must come first");
  refresh_lists(h);  // the instances' step precisions, for the per-precision flags
  staging_wait(h);
  for (int i = 0; i < h->K; ++i) h->rule_in_host[i] = RuleInput{in[i].due, in[i].force, in[i].since};
  SMY_HIP(hipMemcpyAsync(h->rule_in, h->rule_in_host, sizeof(RuleInput) * h->K, hipMemcpyHostToDevice,
                         h->stream));
  staging_sent(h);
  const int grid = (h->K + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(rule_decide_kernel, dim3(grid), dim3(kBlock), 0, h->stream, h->K, h->K, h->sums_dev,
                     iter1, h->rule_in, h->step_prec, h->rule, h->flags);
  const bool rows = block_rows_grid(h) > 0;
  if (rows) {
    if (uses(h, SMY_FP32)) distance_launch<float>(h);
    if (uses(h, SMY_FP64)) distance_launch<double>(h);
  }
  hipLaunchKernelGGL(finish_ranges_kernel, dim3(2, 2 * h->K), dim3(kBlock), 0, h->stream, stride(h), 2,
                     h->partial, h->ranges, h->rule_dist);
  hipLaunchKernelGGL(rule_weight_kernel, dim3(grid), dim3(kBlock), 0, h->stream, h->K, 2, h->rule_dist,
                     h->rule);
  if (rows) {
    if (uses(h, SMY_FP32)) restart_launch<float>(h);
    if (uses(h, SMY_FP64)) restart_launch<double>(h);
  }
  SMY_HIP(hipGetLastError());
  SMY_HIP(hipMemcpyAsync(h->rule_host, h->rule, sizeof(RuleState) * h->K, hipMemcpyDeviceToHost, h->stream));
  SMY_HIP(hipStreamSynchronize(h->stream));
  h->staged = false;
  bool any = false;
  for (int i = 0; i < h->K; ++i) {
    const RuleState& s = h->rule_host[i];
    out[i] = smy_pdhg_rule_out{s.which, s.kkt_candidate, s.omega};
    if (s.which >= 0) {
      any = true;
      h->ax_stale[i] = 1;
    }
  }
  if (any) {
    h->vbar_any = true;
    h->stepped = false;
  }
  return SMY_OK;
}

template <typename From, typename To>
void convert_launch(smy_pdhg* h) {
  Iter<From>& A = state<From>(h);
  Iter<To>& B = state<To>(h);
  ConvertOp<From, To> op;
  op.which = h->flags + prec_of<From>() * h->K;
  const From* vin[6] = {A.v[0], A.v[1], A.vbar[0], A.vbar[1], A.v_erg, A.v_rst};
  To* vout[6] = {B.v[0], B.v[1], B.vbar[0], B.vbar[1], B.v_erg, B.v_rst};
  const From* yin[4] = {A.y[0], A.y[1], A.y_erg, A.y_rst};
  To* yout[4] = {B.y[0], B.y[1], B.y_erg, B.y_rst};
  for (int q = 0; q < 6; ++q) {
    op.vin[q] = vin[q];
    op.vout[q] = vout[q];
  }
  for (int q = 0; q < 4; ++q) {
    op.yin[q] = yin[q];
    op.yout[q] = yout[q];
  }
  hipLaunchKernelGGL((block_rows_kernel<ConvertOp<From, To>>), dim3(block_rows_grid(h)),
                     dim3(kBlock), 0, h->stream, block_rows(h), op);
}

// New step precisions (SMY_FP64 / SMY_FP32; -1 keeps the instance as it is); the state of each
// instance that changes is converted.
int set_precision_batch_impl(smy_pdhg* h, const int* precision) {
  h->fixed_fresh = false;
  std::vector<int> flag(h->K, -1);
  bool any = false, to32 = false;
  for (int i = 0; i < h->K; ++i) {
    const int want = precision[i];
    if (want < 0 || want == h->inst[i].prec) continue;
    if (h->inst[i].prec == kFrozen) return fail(SMY_STATE, "instance " + std::to_string(i) + " is frozen");
    flag[i] = 1;
    any = true;
    to32 = to32 || want == SMY_FP32;
  }
  if (!any) return SMY_OK;
  if (to32) ensure_fp32(h);
  if (block_rows_grid(h) > 0) {
    stage_flags(h, flag.data());
    if (any_flag(h, flag.data(), SMY_FP64)) convert_launch<double, float>(h);
    if (any_flag(h, flag.data(), SMY_FP32)) convert_launch<float, double>(h);
    SMY_HIP(hipGetLastError());
  }
  for (int i = 0; i < h->K; ++i)
    if (flag[i] > 0) {
      h->inst[i].prec = h->inst[i].last_prec = precision[i];
      h->ax_stale[i] = 1;
    }
  h->lists_dirty = true;
  return SMY_OK;
}

template <typename T>
void download_slice(smy_pdhg* h, const T* src, long long count, double* dst) {
  if (dst == nullptr || count == 0) return;
  if constexpr (std::is_same_v<T, double>) {
    SMY_HIP(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  } else {
    std::vector<float> tmp(static_cast<size_t>(count));
    SMY_HIP(hipMemcpyAsync(tmp.data(), src, count * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    SMY_HIP(hipStreamSynchronize(h->stream));
    for (long long k = 0; k < count; ++k) dst[k] = static_cast<double>(tmp[k]);
  }
}

// Instance i's v (its n structural entries, then its m logical ones) and y.
template <typename T>
int download_impl(smy_pdhg* h, int i, double* v, double* v_new, double* y, double* y_new,
                  double* v_erg, double* y_erg) {
  Iter<T>& S = state<T>(h);
  const Layout& L = h->inst[i];
  const int p = h->p;
  const auto vec = [&](const T* src, double* dst) {
    if (dst == nullptr) return;
    download_slice(h, src + L.ox, L.n, dst);
    download_slice(h, src + h->n + L.oy, L.m, dst + L.n);
  };
  const auto row = [&](const T* src, double* dst) {
    if (dst != nullptr) download_slice(h, src + L.oy, L.m, dst);
  };
  vec(S.v[p], v);
  vec(S.v[1 - p], v_new);
  row(S.y[p], y);
  row(S.y[1 - p], y_new);
  vec(S.v_erg, v_erg);
  row(S.y_erg, y_erg);
  SMY_HIP(hipStreamSynchronize(h->stream));
  return SMY_OK;
}

template <typename T>
void set_start_impl(smy_pdhg* h, int i, const double* v, const double* y) {
  h->fixed_fresh = false;
  Iter<T>& S = state<T>(h);
  const Layout& L = h->inst[i];
  const auto up = [&](const double* src, T* dst, long long count) {
    if (count == 0) return;
    if constexpr (std::is_same_v<T, double>) {
      SMY_HIP(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyHostToDevice));
    } else {
      std::vector<float> tmp(static_cast<size_t>(count));
      for (long long k = 0; k < count; ++k) tmp[k] = static_cast<float>(src[k]);
      SMY_HIP(hipMemcpy(dst, tmp.data(), count * sizeof(float), hipMemcpyHostToDevice));
    }
  };
  const auto vec = [&](T* dst) {
    up(v, dst + L.ox, L.n);
    up(v + L.n, dst + h->n + L.oy, L.m);
  };
  SMY_HIP(hipStreamSynchronize(h->stream));
  for (int q = 0; q < 2; ++q) {
    vec(S.v[q]);
    vec(S.vbar[q]);
    up(y, S.y[q] + L.oy, L.m);
  }
  vec(S.v_erg);
  vec(S.v_rst);
  up(y, S.y_erg + L.oy, L.m);
  up(y, S.y_rst + L.oy, L.m);
  h->vbar_any = true;
  h->ax_stale[i] = 1;
  h->stepped = false;
}

// Instance i's cost / bounds into the stacked FP64 arrays (x part at ox, logical part at n + oy).
void put_instance_vec(smy_pdhg* h, double* dst, int i, const double* src) {
  const Layout& L = h->inst[i];
  if (L.n > 0) SMY_HIP(hipMemcpy(dst + L.ox, src, L.n * sizeof(double), hipMemcpyHostToDevice));
  if (L.m > 0) SMY_HIP(hipMemcpy(dst + h->n + L.oy, src + L.n, L.m * sizeof(double), hipMemcpyHostToDevice));
}

int create_impl(int count, const smy_csc* A, const smy_csc* At, const double* const* cost,
                const double* const* lower, const double* const* upper, int precision,
                smy_pdhg** out) {
  long long nnz_a = 0, nnz_t = 0, n = 0, m = 0;
  for (int i = 0; i < count; ++i) {
    if (A[i].start == nullptr || At[i].start == nullptr) return fail(SMY_INVALID, "null argument");
    if (A[i].rows + A[i].cols > 0 && (cost[i] == nullptr || lower[i] == nullptr || upper[i] == nullptr))
      return fail(SMY_INVALID, "null argument");
    if (A[i].rows != At[i].cols || A[i].cols != At[i].rows || A[i].rows < 0 || A[i].cols < 0)
      return fail(SMY_INVALID, "A and At dimensions do not match");
    if (A[i].start[A[i].cols] != At[i].start[At[i].cols]) return fail(SMY_INVALID, "A and At nnz differ");
    nnz_a += At[i].start[At[i].cols];
    nnz_t += A[i].start[A[i].cols];
    n += A[i].cols;
    m += A[i].rows;
  }
  if (!smy_gpu_available()) return fail(SMY_NO_DEVICE, "no usable gfx950 GPU");
  if (nnz_a >= INT_MAX || nnz_t >= INT_MAX || n + m >= INT_MAX)
    return fail(SMY_TOO_LARGE, "matrix too large for 32-bit CSR offsets");
  return guarded([&]() -> int {
    auto h = std::make_unique<smy_pdhg>();
    h->K = count;
    h->n = static_cast<int>(n);
    h->m = static_cast<int>(m);
    h->nt = h->n + h->m;
    h->inst.resize(count);
    int ox = 0, oy = 0;
    for (int i = 0; i < count; ++i) {
      Layout& L = h->inst[i];
      L.n = A[i].cols;
      L.m = A[i].rows;
      L.ox = ox;
      L.oy = oy;
      L.prec = L.last_prec = precision;
      ox += L.n;
      oy += L.m;
    }
    SMY_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    SMY_HIP(hipEventCreate(&h->ev_start));
    SMY_HIP(hipEventCreate(&h->ev_stop));
    SMY_HIP(hipEventCreateWithFlags(&h->ev_staged, hipEventDisableTiming));
    // A by rows is the CSC of each A' read as CSR; A' by rows is the CSC of each A.
    upload_stack(h.get(), At, true, h->a);
    upload_stack(h.get(), A, false, h->at);
    alloc_iter(h.get(), h->d);
    for (int i = 0; i < count; ++i) {
      put_instance_vec(h.get(), h->d.cost, i, cost[i]);
      put_instance_vec(h.get(), h->d.lower, i, lower[i]);
      put_instance_vec(h.get(), h->d.upper, i, upper[i]);
    }
    if (precision == SMY_FP32) ensure_fp32(h.get());
    h->partial = h->alloc<double>(14LL * std::max(1, stride(h.get())));
    std::vector<int2> ranges(2 * static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) {
      const Layout& L = h->inst[i];
      ranges[2 * i] = make_int2(L.ba0, L.ba1);
      ranges[2 * i + 1] = make_int2(h->a.nblocks + L.bt0, h->a.nblocks + L.bt1);
    }
    h->ranges = h->upload(ranges.data(), static_cast<long long>(ranges.size()));
    h->sums_dev = h->alloc<double>(28LL * count);
    h->scalars = h->alloc<StepScalars>(count);
    h->flags = h->alloc<int>(2LL * count);
    h->step_prec = h->alloc<int>(count);
    h->ax_stale.assign(static_cast<size_t>(count), 1);
    h->cs = h->alloc<double>(h->n);
    h->rs = h->alloc<double>(h->m);
    h->cost_o = h->alloc<double>(h->nt);
    h->lower_o = h->alloc<double>(h->nt);
    h->upper_o = h->alloc<double>(h->nt);
    SMY_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->sums_host), 28 * sizeof(double) * count,
                          hipHostMallocDefault));
    SMY_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->scalars_host), sizeof(StepScalars) * count,
                          hipHostMallocDefault));
    SMY_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->flags_host), 2 * sizeof(int) * count,
                          hipHostMallocDefault));
    rebuild_lists(h.get());
    SMY_HIP(hipStreamSynchronize(h->stream));
    *out = h.release();
    return SMY_OK;
  });
}

// ---- Preconditioning on the device (single-LP handles created with the unscaled LP) ----

// The enhanced algorithm's preconditioner of the handle's LP, computed and applied on the device
// (the host's ruiz_pock_chambolle, csc_scaled of both copies, the cost and bound scaling of
// Lane::prepare, operator_norm_bound and the termination norms, csrc/host/pdlp.cpp): `passes`
// infinity-norm Ruiz rounds, then (pock_chambolle) one Pock-Chambolle step with alpha = 1.
// out = {L, a_colmax, at_rowmax}.
// reference2: the reference's 2-norm variant (passes of full-step updates with the stop rule)
// instead of the infinity-norm Ruiz passes and Pock-Chambolle step.
int equilibrate_impl(smy_pdhg* h, int passes, int pock_chambolle, double* col_scale,
                     double* row_scale, double* out, bool reference2 = false) {
  const int n = h->n, m = h->m;
  const hipStream_t s = h->stream;
  DevMatrix& A = h->a;    // A by rows: row i, indices j
  DevMatrix& At = h->at;  // A' by rows: row j, indices i
  double* cs = h->cs;
  double* rs = h->rs;
  double* cn = h->alloc<double>(n);
  double* rn = h->alloc<double>(m);
  double* scalar = h->alloc<double>(2);
  const auto fill_ones = [&](double* x, int count) {
    if (count == 0) return;
    std::vector<double> ones(static_cast<size_t>(count), 1.0);
    SMY_HIP(hipMemcpyAsync(x, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice, s));
    SMY_HIP(hipStreamSynchronize(s));
  };
  fill_ones(cs, n);
  fill_ones(rs, m);
  const auto cols_reduce = [&](auto term, auto max_tag, double* dst) {
    constexpr bool kMax = decltype(max_tag)::value;
    hipLaunchKernelGGL((segment_wave_kernel<decltype(term), kMax>), dim3(wave_blocks(n)), dim3(kBlock), 0,
                       s, n, At.ptr, term, dst);
  };
  const auto rows_reduce = [&](auto term, auto max_tag, double* dst) {
    constexpr bool kMax = decltype(max_tag)::value;
    hipLaunchKernelGGL((segment_wave_kernel<decltype(term), kMax>), dim3(wave_blocks(m)), dim3(kBlock), 0,
                       s, m, A.ptr, term, dst);
  };
  using Sum = std::false_type;
  using Max = std::true_type;
  // Norms of the termination test from the unscaled matrix (src/lp/pdlp.cpp:377-394).
  double a_colmax = 0.0, at_rowmax = 0.0;
  if (m > 0 && n > 0) {
    cols_reduce(Square{At.val64}, Sum{}, cn);
    hipLaunchKernelGGL(sqrt_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, cn);
    hipLaunchKernelGGL((reduce_finish_kernel<true>), dim3(1), dim3(kReduceThreads), 0, s,
                       static_cast<long long>(n), cn, scalar);
    rows_reduce(Square{A.val64}, Sum{}, rn);
    hipLaunchKernelGGL(sqrt_kernel, dim3(blocks_for(m)), dim3(kBlock), 0, s, m, rn);
    hipLaunchKernelGGL((reduce_finish_kernel<true>), dim3(1), dim3(kReduceThreads), 0, s,
                       static_cast<long long>(m), rn, scalar + 1);
    SMY_HIP(hipGetLastError());
    double two[2];
    SMY_HIP(hipMemcpyAsync(two, scalar, sizeof two, hipMemcpyDeviceToHost, s));
    SMY_HIP(hipStreamSynchronize(s));
    a_colmax = two[0];
    at_rowmax = two[1];
  }
  if (m > 0 && n > 0) {
    const ScaledAbs col_term{At.idx, At.val64, rs, cs, false};
    const ScaledAbs row_term{A.idx, A.val64, rs, cs, true};
    const auto round = [&](bool one_norm) {
      if (one_norm) {
        cols_reduce(col_term, Sum{}, cn);
        rows_reduce(row_term, Sum{}, rn);
      } else {
        cols_reduce(col_term, Max{}, cn);
        rows_reduce(row_term, Max{}, rn);
      }
      hipLaunchKernelGGL(divide_by_sqrt_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, cs, cn);
      hipLaunchKernelGGL(divide_by_sqrt_kernel, dim3(blocks_for(m)), dim3(kBlock), 0, s, m, rs, rn);
    };
    if (reference2) {
      // Both norms from the current factors, then both updates; the host reads the flag after
      // each pass and stops when no factor moved by more than 10%.
      int* moved = reinterpret_cast<int*>(scalar);
      for (int pass = 0; pass < passes; ++pass) {
        cols_reduce(ScaledSquare{At.idx, At.val64, rs, cs, false}, Sum{}, cn);
        hipLaunchKernelGGL(sqrt_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, cn);
        rows_reduce(ScaledSquare{A.idx, A.val64, rs, cs, true}, Sum{}, rn);
        hipLaunchKernelGGL(sqrt_kernel, dim3(blocks_for(m)), dim3(kBlock), 0, s, m, rn);
        SMY_HIP(hipMemsetAsync(moved, 0, sizeof(int), s));
        hipLaunchKernelGGL(ruiz2_update_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, cs, cn, moved);
        hipLaunchKernelGGL(ruiz2_update_kernel, dim3(blocks_for(m)), dim3(kBlock), 0, s, m, rs, rn, moved);
        SMY_HIP(hipGetLastError());
        int any = 0;
        SMY_HIP(hipMemcpyAsync(&any, moved, sizeof(int), hipMemcpyDeviceToHost, s));
        SMY_HIP(hipStreamSynchronize(s));
        if (!any) break;
      }
    } else {
      for (int pass = 0; pass < passes; ++pass) round(false);
      if (pock_chambolle) round(true);
    }
  }
  // 20 power iterations on A_eq'A_eq from the host's seeded start (src/lp/pdlp.cpp:36-103), on the
  // unscaled values and the factors in the host's order, norms on the host:
  // L = sqrt((1.01 sigma)^2 + 1).
  double L = 1.0;
  if (m > 0 && n > 0) {
    std::vector<double> hu(static_cast<size_t>(m)), hw(static_cast<size_t>(n));
    std::mt19937_64 gen(12345);
    std::uniform_real_distribution<double> uniform(-1.0, 1.0);
    for (double& t : hu) t = uniform(gen);
    const auto norm = [](const std::vector<double>& z) {
      double acc = 0.0;
      for (double t : z) acc += t * t;
      return std::sqrt(acc);
    };
    double un = norm(hu);
    for (double& t : hu) t /= un;
    double* u = rn;  // reused: the row norms are no longer needed
    double* w = cn;
    SMY_HIP(hipMemcpyAsync(u, hu.data(), hu.size() * sizeof(double), hipMemcpyHostToDevice, s));
    double sigma = 0.0;
    bool zero = false;
    for (int it = 0; it < 20; ++it) {
      cols_reduce(PowerW{At.idx, At.val64, rs, cs, u}, Sum{}, w);
      SMY_HIP(hipMemcpyAsync(hw.data(), w, hw.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      SMY_HIP(hipStreamSynchronize(s));
      const double wn = norm(hw);
      if (wn == 0.0) {
        zero = true;
        break;
      }
      rows_reduce(PowerU{A.idx, A.val64, rs, cs, w}, Sum{}, u);
      SMY_HIP(hipMemcpyAsync(hu.data(), u, hu.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      SMY_HIP(hipStreamSynchronize(s));
      un = norm(hu);
      if (un == 0.0) {
        zero = true;
        break;
      }
      sigma = un / wn;
      for (double& t : hu) t /= un;
      SMY_HIP(hipMemcpyAsync(u, hu.data(), hu.size() * sizeof(double), hipMemcpyHostToDevice, s));
    }
    if (!zero) {
      const double la = sigma * 1.01;
      L = std::sqrt(la * la + 1.0);
    }
  }
  if (m > 0 && n > 0) {
    // The scaled copies: A by rows (a * cs_j) * rs_i, A' by rows (a * rs_i) * cs_j.
    hipLaunchKernelGGL((scale_values_wave_kernel<true>), dim3(wave_blocks(m)), dim3(kBlock), 0, s,
                       m, A.ptr, A.idx, A.val64, rs, cs);
    hipLaunchKernelGGL((scale_values_wave_kernel<false>), dim3(wave_blocks(n)), dim3(kBlock), 0, s,
                       n, At.ptr, At.idx, At.val64, rs, cs);
    SMY_HIP(hipGetLastError());
  }
  if (h->nt > 0)
    hipLaunchKernelGGL(scale_vectors_kernel, dim3(blocks_for(h->nt)), dim3(kBlock), 0, s, n, m, cs, rs,
                       h->d.cost, h->d.lower, h->d.upper);
  if (h->f.ready) {  // the FP32 copies follow the scaled FP64 ones
    for (DevMatrix* M : {&h->a, &h->at})
      if (M->nnz > 0)
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(M->nnz)), dim3(kBlock), 0, s,
                           M->nnz, M->val64, M->val32);
    for (auto [src, dst] : {std::pair{h->d.cost, h->f.cost}, std::pair{h->d.lower, h->f.lower},
                            std::pair{h->d.upper, h->f.upper}})
      if (h->nt > 0)
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(h->nt)), dim3(kBlock), 0, s,
                           static_cast<long long>(h->nt), src, dst);
  }
  SMY_HIP(hipGetLastError());
  if (n > 0) SMY_HIP(hipMemcpyAsync(col_scale, cs, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (m > 0) SMY_HIP(hipMemcpyAsync(row_scale, rs, m * sizeof(double), hipMemcpyDeviceToHost, s));
  SMY_HIP(hipStreamSynchronize(s));
  out[0] = L;
  out[1] = a_colmax;
  out[2] = at_rowmax;
  h->vbar_any = true;
  return SMY_OK;
}

int set_original_impl(smy_pdhg* h, int i, const double* col_scale, const double* row_scale,
                      const double* cost, const double* lower, const double* upper) {
  const Layout& L = h->inst[i];
  // Empty vectors may come as null pointers (an LP without rows has no row scales).
  if ((L.n > 0 && col_scale == nullptr) || (L.m > 0 && row_scale == nullptr) ||
      (L.n + L.m > 0 && (cost == nullptr || lower == nullptr || upper == nullptr)))
    return fail(SMY_INVALID, "null argument");
  return guarded([&]() -> int {
    SMY_HIP(hipStreamSynchronize(h->stream));
    if (L.n > 0) SMY_HIP(hipMemcpy(h->cs + L.ox, col_scale, L.n * sizeof(double), hipMemcpyHostToDevice));
    if (L.m > 0) SMY_HIP(hipMemcpy(h->rs + L.oy, row_scale, L.m * sizeof(double), hipMemcpyHostToDevice));
    put_instance_vec(h, h->cost_o, i, cost);
    put_instance_vec(h, h->lower_o, i, lower);
    put_instance_vec(h, h->upper_o, i, upper);
    h->inst[i].has_original = true;
    h->inst[i].boxed = all_boxed(L.n, lower, upper);
    return SMY_OK;
  });
}

// ---- Device-side infeasibility / unboundedness tests (stacked layout) ----

void ensure_detect(smy_pdhg* h) {
  if (h->det_partial != nullptr) return;
  h->det_partial = h->alloc<double>(static_cast<long long>(kDetectSlots) * std::max(1, stride(h)));
  h->det_scratch = h->alloc<double>(std::max(1, h->n));
  h->det_ray = h->alloc<double>(std::max(1, h->n));
  h->det_prm = h->zeros<double>(3LL * h->K);
  h->det_out = h->alloc<double>(static_cast<long long>(kDetectSlots) * h->K);
  SMY_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->det_prm_host), 3 * sizeof(double) * h->K,
                        hipHostMallocDefault));
  SMY_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->det_out_host), kDetectSlots * sizeof(double) * h->K,
                        hipHostMallocDefault));
}

template <typename T>
T*& det_y(smy_pdhg* h);
template <>
double*& det_y<double>(smy_pdhg* h) { return h->det_y64; }
template <>
float*& det_y<float>(smy_pdhg* h) { return h->det_y32; }

template <typename T, template <typename, bool> class Op>
Op<T, false> stacked_detect_op(smy_pdhg* h, int slot) {
  Iter<T>& S = state<T>(h);
  Op<T, false> op{};
  op.n = h->n;
  op.kp = 1;
  op.v = S.v[h->p];
  op.v_new = S.v[1 - h->p];
  op.y_erg = h->det_projected ? det_y<T>(h) : S.y_erg;
  op.y_rst = S.y_rst;
  op.rs = h->rs;
  op.cs = h->cs;
  op.cost = h->cost_o;
  op.lower = h->lower_o;
  op.upper = h->upper_o;
  op.row_lower = nullptr;
  op.row_upper = nullptr;
  op.scratch = h->det_scratch;
  op.ray = h->det_ray;
  op.prm = h->det_prm;
  op.flags = h->flags + prec_of<T>() * h->K;
  op.partial = h->det_partial + static_cast<size_t>(slot) * stride(h);
  op.stride = static_cast<size_t>(stride(h));
  op.inst_stride = 0;
  return op;
}

template <typename Op>
void detect_csr(smy_pdhg* h, const DevMatrix& M, const Op& op) {
  if (M.nblocks == 0) return;
  CsrRows<double> A;
  A.rows = M.rows;
  A.nblocks = M.nblocks;
  A.ptr = M.ptr;
  A.idx = M.idx;
  A.val = M.val64;
  A.meta = M.meta;
  A.binst = M.binst;
  A.count = M.nblocks;
  hipLaunchKernelGGL((csr_rows_kernel<double, Op, false>), dim3(M.nblocks), dim3(kBlock), 0, h->stream, A, op);
}

template <typename Op>
void detect_rows(smy_pdhg* h, const Op& op, int first, int count) {
  if (count == 0) return;
  BlockRows R = block_rows(h);
  R.first = first;
  hipLaunchKernelGGL((block_rows_kernel<Op>), dim3(count), dim3(kBlock), 0, h->stream, R, op);
}

template <typename T>
void stacked_detect_phase(smy_pdhg* h, int phase) {
  const int nba = h->a.nblocks, nbt = h->at.nblocks;
  if (phase == 1 && h->det_projected) {
    // The tested duals (y_erg or its sign projection) into det_y, which the Farkas sums read.
    if (det_y<T>(h) == nullptr) det_y<T>(h) = h->alloc<T>(std::max(1, h->m));
    h->det_projected = false;  // the copy reads y_erg
    auto project = stacked_detect_op<T, FarkasProjectOp>(h, 0);
    h->det_projected = true;
    project.ydst = det_y<T>(h);
    detect_rows(h, project, 0, nba);
  }
  if (phase == 1) {
    // Block partials of A' rows follow those of A (the instance ranges' layout).
    auto cols = stacked_detect_op<T, FarkasColsOp>(h, 0);
    cols.partial += nba;
    detect_csr(h, h->at, cols);
    detect_rows(h, stacked_detect_op<T, FarkasRowsOp>(h, 0), 0, nba);
    if (h->det_ray_test) detect_rows(h, stacked_detect_op<T, RayMaxOp>(h, 7), 0, nba + nbt);
  } else {
    detect_rows(h, stacked_detect_op<T, FarkasBoundsOp>(h, 2), 0, nba + nbt);
    if (h->det_ray_test) {
      detect_csr(h, h->a, stacked_detect_op<T, RayRowsOp>(h, 8));
      detect_rows(h, stacked_detect_op<T, RayColsOp>(h, 9), nba, nbt);
      detect_csr(h, h->a, stacked_detect_op<T, RayCheckOp>(h, 10));
    }
  }
}

void zero_slots(smy_pdhg* h, double* partial, size_t slot_stride, int first, int count) {
  SMY_HIP(hipMemsetAsync(partial + static_cast<size_t>(first) * slot_stride, 0,
                         static_cast<size_t>(count) * slot_stride * sizeof(double), h->stream));
}

// Both phases; the host forms drop, y_max and d_max between them.
int stacked_detect(smy_pdhg* h, const int* flags, double* out) {
  refresh_lists(h);
  ensure_detect(h);
  stage_flags(h, flags);
  h->det_projected = std::any_of(flags, flags + h->K, [](int f) { return f >= 2; });
  if (h->det_projected) h->det_y_ready = true;
  h->det_ray_test = false;
  for (int i = 0; i < h->K; ++i) h->det_ray_test = h->det_ray_test || (flags[i] == 1 && !h->inst[i].boxed);
  const bool f32 = any_flag(h, flags, SMY_FP32), f64 = any_flag(h, flags, SMY_FP64);
  const size_t st = static_cast<size_t>(stride(h));
  const auto finish_det = [&]() {
    hipLaunchKernelGGL(finish_detect_kernel, dim3(kDetectSlots, h->K), dim3(kBlock), 0, h->stream, st,
                       h->det_partial, h->ranges, h->det_out);
    SMY_HIP(hipGetLastError());
    SMY_HIP(hipMemcpyAsync(h->det_out_host, h->det_out, kDetectSlots * sizeof(double) * h->K,
                           hipMemcpyDeviceToHost, h->stream));
    SMY_HIP(hipStreamSynchronize(h->stream));
  };
  if (st > 0) {
    zero_slots(h, h->det_partial, st, 0, 2);
    zero_slots(h, h->det_partial, st, 7, 1);
    if (f32) stacked_detect_phase<float>(h, 1);
    if (f64) stacked_detect_phase<double>(h, 1);
  }
  finish_det();
  staging_wait(h);
  for (int i = 0; i < h->K; ++i) {
    const double* o = h->det_out_host + static_cast<size_t>(i) * kDetectSlots;
    h->det_prm_host[3 * i] = 1e-9 * o[0];
    h->det_prm_host[3 * i + 1] = o[1];
    h->det_prm_host[3 * i + 2] = o[7];
    for (int q = 0; q < kDetectSlots; ++q) out[static_cast<size_t>(i) * kDetectSlots + q] = o[q];
  }
  SMY_HIP(hipMemcpyAsync(h->det_prm, h->det_prm_host, 3 * sizeof(double) * h->K, hipMemcpyHostToDevice,
                         h->stream));
  staging_sent(h);
  if (st > 0) {
    zero_slots(h, h->det_partial, st, 2, 5);
    zero_slots(h, h->det_partial, st, 8, 3);
    if (f32) stacked_detect_phase<float>(h, 2);
    if (f64) stacked_detect_phase<double>(h, 2);
  }
  finish_det();
  h->det_projected = false;
  for (int i = 0; i < h->K; ++i)
    for (int q : {2, 3, 4, 5, 6, 8, 9, 10})
      out[static_cast<size_t>(i) * kDetectSlots + q] = h->det_out_host[static_cast<size_t>(i) * kDetectSlots + q];
  return SMY_OK;
}

bool single(smy_pdhg* h) { return h->K == 1; }

}  // namespace
}  // namespace smy

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

const char* smy_last_error(void) { return g_last_error; }

int smy_gpu_available(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return 0;
  return device_is_gfx950(device) ? 1 : 0;
}

int smy_gpu_device_name(char* buf, size_t len) {
  if (buf == nullptr || len == 0) return fail(SMY_INVALID, "null buffer");
  buf[0] = '\0';
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SMY_NO_DEVICE;
  int device = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess)
    return SMY_NO_DEVICE;
  std::snprintf(buf, len, "%s (%s)", prop.name, prop.gcnArchName);
  return SMY_OK;
}

int smy_pdhg_create(const smy_csc* A, const smy_csc* At, const double* cost, const double* lower,
                    const double* upper, int precision, smy_pdhg** out) {
  if (out == nullptr) return fail(SMY_INVALID, "null output handle");
  *out = nullptr;
  if (A == nullptr || At == nullptr) return fail(SMY_INVALID, "null argument");
  if (precision != SMY_FP64 && precision != SMY_FP32) return fail(SMY_INVALID, "bad precision");
  return create_impl(1, A, At, &cost, &lower, &upper, precision, out);
}

int smy_pdhg_create_batch(int count, const smy_csc* A, const smy_csc* At,
                          const double* const* cost, const double* const* lower,
                          const double* const* upper, int precision, smy_pdhg** out) {
  if (out == nullptr) return fail(SMY_INVALID, "null output handle");
  *out = nullptr;
  if (count <= 0 || count > kMaxBatch || A == nullptr || At == nullptr || cost == nullptr ||
      lower == nullptr || upper == nullptr)
    return fail(SMY_INVALID, "bad argument");
  if (precision != SMY_FP64 && precision != SMY_FP32) return fail(SMY_INVALID, "bad precision");
  return create_impl(count, A, At, cost, lower, upper, precision, out);
}

void smy_pdhg_destroy(smy_pdhg* h) { delete h; }

int smy_pdhg_batch_size(smy_pdhg* h) { return h == nullptr ? 0 : h->K; }

#define SMY_SINGLE(h)                                                                   \
  do {                                                                                  \
    if ((h) == nullptr) return fail(SMY_INVALID, "null handle");                        \
    if (!single(h) || (h)->shared != nullptr)                                           \
      return fail(SMY_STATE, "a batch handle needs the batch functions");               \
  } while (0)

#define SMY_INSTANCE(h, i)                                                   \
  do {                                                                       \
    if ((h) == nullptr) return fail(SMY_INVALID, "null handle");             \
    if ((i) < 0 || (i) >= (h)->K) return fail(SMY_INVALID, "bad instance");  \
  } while (0)

int smy_pdhg_set_original(smy_pdhg* h, const double* col_scale, const double* row_scale,
                          const double* cost, const double* lower, const double* upper) {
  SMY_SINGLE(h);
  return set_original_impl(h, 0, col_scale, row_scale, cost, lower, upper);
}

int smy_pdhg_set_original_at(smy_pdhg* h, int instance, const double* col_scale,
                             const double* row_scale, const double* cost, const double* lower,
                             const double* upper) {
  SMY_INSTANCE(h, instance);
  if (h->shared) {
    const int n = h->n, m = h->m;
    if ((n > 0 && col_scale == nullptr) || (m > 0 && row_scale == nullptr) ||
        (n + m > 0 && (cost == nullptr || lower == nullptr || upper == nullptr)))
      return fail(SMY_INVALID, "null argument");
    return guarded([&]() -> int { return shared_set_original(h, instance, col_scale, row_scale, cost, lower, upper); });
  }
  return set_original_impl(h, instance, col_scale, row_scale, cost, lower, upper);
}

// The test hook's check before `count` more steps: SMY_ERROR (nothing launched) once they would
// pass the injected step, else the count is taken.
static int take_steps(smy_pdhg* h, int count) {
  if (h->fault_at >= 0 && h->steps_run + count > h->fault_at) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "injected fault at step %lld (test hook)", h->fault_at);
    return fail(SMY_ERROR, msg);
  }
  h->steps_run += count > 0 ? count : 0;
  return SMY_OK;
}

int smy_pdhg_equilibrate(smy_pdhg* h, int passes, int pock_chambolle, double* col_scale,
                         double* row_scale, double out[3]) {
  SMY_SINGLE(h);
  if (passes < 0 || out == nullptr || (h->n > 0 && col_scale == nullptr) ||
      (h->m > 0 && row_scale == nullptr))
    return fail(SMY_INVALID, "bad argument");
  if (h->stepped || h->steps_run > 0) return fail(SMY_STATE, "equilibrate before the first step");
  return guarded([&]() -> int { return equilibrate_impl(h, passes, pock_chambolle, col_scale, row_scale, out); });
}

int smy_pdhg_equilibrate_reference(smy_pdhg* h, int passes, double* col_scale, double* row_scale,
                                   double out[3]) {
  SMY_SINGLE(h);
  if (passes < 0 || out == nullptr || (h->n > 0 && col_scale == nullptr) ||
      (h->m > 0 && row_scale == nullptr))
    return fail(SMY_INVALID, "bad argument");
  if (h->stepped || h->steps_run > 0) return fail(SMY_STATE, "equilibrate before the first step");
  return guarded([&]() -> int { return equilibrate_impl(h, passes, 0, col_scale, row_scale, out, true); });
}

int smy_pdhg_inject_fault(smy_pdhg* h, long long step_no) {
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  h->fault_at = step_no < 0 ? -1 : step_no;
  return SMY_OK;
}

int smy_pdhg_step(smy_pdhg* h, double tau, double sigma, double theta, long long iter) {
  SMY_SINGLE(h);
  if (take_steps(h, 1) != SMY_OK) return SMY_ERROR;
  return guarded([&]() -> int { return step_impl(h, &tau, &sigma, theta, &iter); });
}

int smy_pdhg_run(smy_pdhg* h, double tau, double sigma, double theta, long long iter0, int count) {
  SMY_SINGLE(h);
  if (take_steps(h, count) != SMY_OK) return SMY_ERROR;
  return guarded([&]() -> int { return run_impl(h, &tau, &sigma, theta, &iter0, count); });
}

int smy_pdhg_run_batch(smy_pdhg* h, const double* tau, const double* sigma, double theta,
                       const long long* iter0, int count) {
  if (h == nullptr || tau == nullptr || sigma == nullptr || iter0 == nullptr)
    return fail(SMY_INVALID, "null argument");
  if (take_steps(h, count) != SMY_OK) return SMY_ERROR;
  if (h->shared) return guarded([&]() -> int { return shared_run(h, tau, sigma, theta, iter0, count); });
  return guarded([&]() -> int { return run_impl(h, tau, sigma, theta, iter0, count); });
}

int smy_pdhg_set_halpern(smy_pdhg* h, int on, double reflection) {
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  if (!(reflection >= 0.0 && reflection <= 1.0)) return fail(SMY_INVALID, "reflection outside [0, 1]");
  if (h->stepped) return fail(SMY_STATE, "the step kind changed between two steps without advance()");
  if (h->rho != reflection) {
    for (auto& row : h->hgraphs)
      for (auto& c : row) h->drop_graph(c);
  }
  h->halpern = on != 0;
  h->rho = reflection;
  return SMY_OK;
}

int smy_pdhg_fixed_point_batch(smy_pdhg* h, double* out) {
  if (h == nullptr || out == nullptr) return fail(SMY_INVALID, "null argument");
  if (!h->halpern) return fail(SMY_STATE, "no Halpern steps on this handle");
  if (h->shared) return guarded([&]() -> int { return shared_fixed_point(h, out); });
  return guarded([&]() -> int {
    refresh_lists(h);
    return fixed_point_batch_impl(h, out);
  });
}

int smy_pdhg_run_adaptive(smy_pdhg* h, smy_pdhg_adaptive* state, int count) {
  if (h == nullptr || state == nullptr) return fail(SMY_INVALID, "null argument");
  if (h->shared) return fail(SMY_INVALID, "adaptive steps need the stacked batch layout");
  if (take_steps(h, count) != SMY_OK) return SMY_ERROR;
  return guarded([&]() -> int { return run_adaptive_impl(h, state, count); });
}

int smy_pdhg_time_kernels(smy_pdhg* h, int reps, double* dual_ms, double* primal_ms) {
  if (h == nullptr || dual_ms == nullptr || primal_ms == nullptr || reps <= 0)
    return fail(SMY_INVALID, "bad argument");
  if (h->shared) return guarded([&]() -> int { return shared_time_kernels(h, reps, dual_ms, primal_ms); });
  return guarded([&]() -> int {
    refresh_lists(h);
    tune(h);
    if (uses(h, SMY_FP32))
      return h->acc64 ? time_kernels_impl<float, double>(h, reps, dual_ms, primal_ms)
                      : time_kernels_impl<float>(h, reps, dual_ms, primal_ms);
    if (uses(h, SMY_FP64)) return time_kernels_impl<double>(h, reps, dual_ms, primal_ms);
    return fail(SMY_STATE, "no instance steps");
  });
}

int smy_pdhg_time_adaptive_kernels(smy_pdhg* h, int reps, double* dual_ms, double* primal_ms,
                                   double* decide_ms) {
  if (h == nullptr || dual_ms == nullptr || primal_ms == nullptr || decide_ms == nullptr || reps <= 0)
    return fail(SMY_INVALID, "bad argument");
  if (h->shared) return fail(SMY_INVALID, "adaptive steps need the stacked batch layout");
  return guarded([&]() -> int {
    refresh_lists(h);
    tune(h);
    if (uses(h, SMY_FP32))
      return h->acc64 ? time_adaptive_impl<float, double>(h, reps, dual_ms, primal_ms, decide_ms)
                      : time_adaptive_impl<float>(h, reps, dual_ms, primal_ms, decide_ms);
    if (uses(h, SMY_FP64)) return time_adaptive_impl<double>(h, reps, dual_ms, primal_ms, decide_ms);
    return fail(SMY_STATE, "no instance steps");
  });
}

int smy_pdhg_rule_init(smy_pdhg* h, const smy_pdhg_rule_setup* init) {
  if (h == nullptr || init == nullptr) return fail(SMY_INVALID, "null argument");
  return guarded([&]() -> int { return rule_init_impl(h, init); });
}

int smy_pdhg_rule_decide(smy_pdhg* h, const smy_pdhg_rule_in* in, long long iter1, smy_pdhg_rule_out* out) {
  if (h == nullptr || in == nullptr || out == nullptr) return fail(SMY_INVALID, "null argument");
  if (h->rule == nullptr) return fail(SMY_STATE, "smy_pdhg_rule_init() must come first");
  if (h->shared) return guarded([&]() -> int { return shared_rule_decide(h, in, iter1, out); });
  return guarded([&]() -> int { return rule_decide_impl(h, in, iter1, out); });
}

int smy_pdhg_get_tile(smy_pdhg* h, int matrix, int precision, int* per_wg, int* xcd) {
  if (h == nullptr || per_wg == nullptr || xcd == nullptr || (matrix != 0 && matrix != 1) ||
      (precision != SMY_FP64 && precision != SMY_FP32))
    return fail(SMY_INVALID, "bad argument");
  if (h->shared) return fail(SMY_INVALID, "the shared-matrix batch has no product tiles");
  const DevMatrix& M = matrix == 0 ? h->a : h->at;
  *per_wg = M.tile[precision].per_wg;
  *xcd = M.tile[precision].xcd ? 1 : 0;
  return M.tuned[precision] ? SMY_OK : SMY_STATE;
}

int smy_pdhg_set_tile(smy_pdhg* h, int matrix, int precision, int per_wg, int xcd) {
  if (h == nullptr || (matrix != 0 && matrix != 1) || (precision != SMY_FP64 && precision != SMY_FP32))
    return fail(SMY_INVALID, "bad argument");
  if (h->shared) return fail(SMY_INVALID, "the shared-matrix batch has no product tiles");
  if (per_wg != 1 && per_wg != 2 && per_wg != 4) return fail(SMY_INVALID, "per_wg must be 1, 2 or 4");
  DevMatrix& M = matrix == 0 ? h->a : h->at;
  M.tile[precision] = Tile{per_wg, xcd != 0};
  M.split[precision] = false;
  M.tuned[precision] = true;
  ++h->list_gen;
  return SMY_OK;
}

int smy_pdhg_set_fp64_accumulation(smy_pdhg* h, int on) {
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  if (h->shared && on) return fail(SMY_INVALID, "the shared-matrix batch accumulates in its step precision");
  if (h->acc64 != (on != 0)) {
    h->acc64 = on != 0;
    ++h->list_gen;  // graphs captured before hold the other kernels
  }
  return SMY_OK;
}

int smy_pdhg_get_bin_tile(smy_pdhg* h, int matrix, int precision, int bin, int* per_wg, int* xcd,
                          int* split) {
  if (h == nullptr || per_wg == nullptr || xcd == nullptr || split == nullptr ||
      (matrix != 0 && matrix != 1) || (precision != SMY_FP64 && precision != SMY_FP32) || bin < 0 ||
      bin >= kTileBins)
    return fail(SMY_INVALID, "bad argument");
  if (h->shared) return fail(SMY_INVALID, "the shared-matrix batch has no product tiles");
  const DevMatrix& M = matrix == 0 ? h->a : h->at;
  *per_wg = M.bin_tile[precision][bin].per_wg;
  *xcd = M.bin_tile[precision][bin].xcd ? 1 : 0;
  *split = M.split[precision] ? 1 : 0;
  return M.tuned[precision] ? SMY_OK : SMY_STATE;
}

int smy_pdhg_set_bin_tile(smy_pdhg* h, int matrix, int precision, int bin, int per_wg, int xcd) {
  if (h == nullptr || (matrix != 0 && matrix != 1) || (precision != SMY_FP64 && precision != SMY_FP32) ||
      bin < 0 || bin >= kTileBins)
    return fail(SMY_INVALID, "bad argument");
  if (h->shared) return fail(SMY_INVALID, "the shared-matrix batch has no product tiles");
  if (per_wg != 1 && per_wg != 2 && per_wg != 4) return fail(SMY_INVALID, "per_wg must be 1, 2 or 4");
  DevMatrix& M = matrix == 0 ? h->a : h->at;
  M.bin_tile[precision][bin] = Tile{per_wg, xcd != 0};
  M.split[precision] = true;
  M.tuned[precision] = true;
  ++h->list_gen;
  return SMY_OK;
}

int smy_pdhg_last_run_ms(smy_pdhg* h, double* ms) {
  if (h == nullptr || ms == nullptr) return fail(SMY_INVALID, "null argument");
  if (!h->timed) return fail(SMY_STATE, "no timed run");
  return guarded([&]() -> int {
    SMY_HIP(hipEventSynchronize(h->ev_stop));
    float t = 0.0f;
    SMY_HIP(hipEventElapsedTime(&t, h->ev_start, h->ev_stop));
    *ms = t;
    return SMY_OK;
  });
}

int smy_pdhg_advance(smy_pdhg* h) {
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  h->p ^= 1;
  h->stepped = false;
  return SMY_OK;
}

int smy_pdhg_restart(smy_pdhg* h) {
  SMY_SINGLE(h);
  return guarded([&]() -> int {
    const auto go = [&](auto& S) {
      using T = std::remove_pointer_t<decltype(S.cost)>;
      hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks_for(h->nt)), dim3(kBlock), 0, h->stream,
                         static_cast<long long>(h->nt), S.v[h->p], S.v[1 - h->p]);
      hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks_for(h->nt)), dim3(kBlock), 0, h->stream,
                         static_cast<long long>(h->nt), S.v[h->p], S.vbar[h->p]);
    };
    if (h->nt > 0) {
      if (h->inst[0].last_prec == SMY_FP64) go(h->d); else go(h->f);
    }
    SMY_HIP(hipGetLastError());
    h->vbar_any = true;
    return SMY_OK;
  });
}

int smy_pdhg_download(smy_pdhg* h, double* v, double* v_new, double* y, double* y_new,
                      double* v_erg, double* y_erg) {
  SMY_SINGLE(h);
  return smy_pdhg_download_at(h, 0, v, v_new, y, y_new, v_erg, y_erg);
}

int smy_pdhg_download_at(smy_pdhg* h, int instance, double* v, double* v_new, double* y,
                         double* y_new, double* v_erg, double* y_erg) {
  SMY_INSTANCE(h, instance);
  if (h->shared)
    return guarded([&]() -> int { return shared_download(h, instance, v, v_new, y, y_new, v_erg, y_erg); });
  return guarded([&]() -> int {
    return h->inst[instance].last_prec == SMY_FP64
               ? download_impl<double>(h, instance, v, v_new, y, y_new, v_erg, y_erg)
               : download_impl<float>(h, instance, v, v_new, y, y_new, v_erg, y_erg);
  });
}

int smy_pdhg_evaluate(smy_pdhg* h, int which, double out[7]) {
  SMY_SINGLE(h);
  if (out == nullptr) return fail(SMY_INVALID, "null argument");
  if (!h->inst[0].has_original) return fail(SMY_STATE, "set_original() must be called before evaluate()");
  if (h->inst[0].prec == kFrozen) return fail(SMY_STATE, "instance 0 is frozen");
  return guarded([&]() -> int {
    refresh_lists(h);
    if (h->inst[0].prec == SMY_FP64) measure_one<double>(h, which); else measure_one<float>(h, which);
    SMY_HIP(hipGetLastError());
    finish(h, 7, out);
    return SMY_OK;
  });
}

int smy_pdhg_evaluate_both(smy_pdhg* h, double out[14]) {
  SMY_SINGLE(h);
  if (out == nullptr) return fail(SMY_INVALID, "null argument");
  return guarded([&]() -> int { return evaluate_batch_impl(h, out); });
}

int smy_pdhg_evaluate_batch(smy_pdhg* h, double* out) {
  if (h == nullptr || out == nullptr) return fail(SMY_INVALID, "null argument");
  if (h->shared) return guarded([&]() -> int { return shared_evaluate(h, out); });
  return guarded([&]() -> int { return evaluate_batch_impl(h, out); });
}

int smy_pdhg_restart_distance(smy_pdhg* h, int which, double* dv_sq, double* dy_sq) {
  SMY_SINGLE(h);
  if (dv_sq == nullptr || dy_sq == nullptr) return fail(SMY_INVALID, "null argument");
  return guarded([&]() -> int {
    double out[2] = {0.0, 0.0};
    const int rc = distance_batch_impl(h, &which, out);
    *dv_sq = out[0];
    *dy_sq = out[1];
    return rc;
  });
}

int smy_pdhg_restart_distance_batch(smy_pdhg* h, const int* which, double* out) {
  if (h == nullptr || which == nullptr || out == nullptr) return fail(SMY_INVALID, "null argument");
  if (h->shared) return guarded([&]() -> int { return shared_distance(h, which, out); });
  return guarded([&]() -> int { return distance_batch_impl(h, which, out); });
}

int smy_pdhg_restart_to(smy_pdhg* h, int which) {
  SMY_SINGLE(h);
  return guarded([&]() -> int { return restart_to_batch_impl(h, &which); });
}

int smy_pdhg_restart_to_batch(smy_pdhg* h, const int* which) {
  if (h == nullptr || which == nullptr) return fail(SMY_INVALID, "null argument");
  if (h->shared) return guarded([&]() -> int { return shared_restart_to(h, which); });
  return guarded([&]() -> int { return restart_to_batch_impl(h, which); });
}

int smy_pdhg_set_precision(smy_pdhg* h, int precision) {
  SMY_SINGLE(h);
  if (precision != SMY_FP64 && precision != SMY_FP32) return fail(SMY_INVALID, "bad precision");
  return guarded([&]() -> int { return set_precision_batch_impl(h, &precision); });
}

int smy_pdhg_set_precision_batch(smy_pdhg* h, const int* precision) {
  if (h == nullptr || precision == nullptr) return fail(SMY_INVALID, "null argument");
  for (int i = 0; i < h->K; ++i)
    if (precision[i] != -1 && precision[i] != SMY_FP64 && precision[i] != SMY_FP32)
      return fail(SMY_INVALID, "bad precision");
  if (h->shared) return guarded([&]() -> int { return shared_set_precision(h, precision); });
  return guarded([&]() -> int { return set_precision_batch_impl(h, precision); });
}

int smy_pdhg_freeze(smy_pdhg* h, int instance) {
  SMY_INSTANCE(h, instance);
  if (h->shared) return shared_freeze(h, instance);
  if (h->inst[instance].prec == kFrozen) return SMY_OK;
  h->inst[instance].prec = kFrozen;
  h->lists_dirty = true;
  return SMY_OK;
}

int smy_pdhg_set_start(smy_pdhg* h, const double* v, const double* y) {
  SMY_SINGLE(h);
  return smy_pdhg_set_start_at(h, 0, v, y);
}

int smy_pdhg_set_start_at(smy_pdhg* h, int instance, const double* v, const double* y) {
  SMY_INSTANCE(h, instance);
  const Layout& L = h->inst[instance];
  if ((v == nullptr && L.n + L.m > 0) || (y == nullptr && L.m > 0)) return fail(SMY_INVALID, "null argument");
  if (h->shared) return guarded([&]() -> int { return shared_set_start(h, instance, v, y); });
  return guarded([&]() -> int {
    if (L.last_prec == SMY_FP64) set_start_impl<double>(h, instance, v, y);
    else set_start_impl<float>(h, instance, v, y);
    return SMY_OK;
  });
}

int smy_pdhg_detect_batch(smy_pdhg* h, const int* flags, double* out) {
  if (h == nullptr || flags == nullptr || out == nullptr) return fail(SMY_INVALID, "null argument");
  for (int i = 0; i < h->K; ++i)
    if (flags[i] == 4 || flags[i] > 7) return fail(SMY_INVALID, "detection flags are 0-3 and 5-7");
  for (int i = 0; i < h->K; ++i)
    if (flags[i] > 0 && (h->inst[i].prec == kFrozen || !h->inst[i].has_original))
      return fail(SMY_STATE, "detection needs a running instance with set_original()");
  if (h->shared) return guarded([&]() -> int { return shared_detect(h, flags, out); });
  return guarded([&]() -> int { return stacked_detect(h, flags, out); });
}

int smy_pdhg_download_ray(smy_pdhg* h, int instance, double* ray) {
  SMY_INSTANCE(h, instance);
  if (ray == nullptr) return fail(SMY_INVALID, "null argument");
  if (h->det_ray == nullptr) return fail(SMY_STATE, "no detection has run");
  if (h->shared) return guarded([&]() -> int { return shared_download_ray(h, instance, ray); });
  return guarded([&]() -> int {
    const Layout& L = h->inst[instance];
    if (L.n > 0)
      SMY_HIP(hipMemcpy(ray, h->det_ray + L.ox, L.n * sizeof(double), hipMemcpyDeviceToHost));
    return SMY_OK;
  });
}

int smy_pdhg_download_farkas(smy_pdhg* h, int instance, double* y) {
  SMY_INSTANCE(h, instance);
  if (y == nullptr) return fail(SMY_INVALID, "null argument");
  if (!h->det_y_ready) return fail(SMY_STATE, "no detection with a flag above 1 has run");
  if (h->shared) return guarded([&]() -> int { return shared_download_farkas(h, instance, y); });
  return guarded([&]() -> int {
    const Layout& L = h->inst[instance];
    if (L.m == 0) return SMY_OK;
    if (L.last_prec == SMY_FP64) {
      if (h->det_y64 == nullptr) return fail(SMY_STATE, "no FP64 detection dual");
      SMY_HIP(hipMemcpy(y, h->det_y64 + L.oy, L.m * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      if (h->det_y32 == nullptr) return fail(SMY_STATE, "no FP32 detection dual");
      std::vector<float> t(static_cast<size_t>(L.m));
      SMY_HIP(hipMemcpy(t.data(), h->det_y32 + L.oy, L.m * sizeof(float), hipMemcpyDeviceToHost));
      std::copy(t.begin(), t.end(), y);
    }
    return SMY_OK;
  });
}

int smy_spmv_threads_per_row(long long nnz, int rows) {
  const double mean = rows > 0 ? static_cast<double>(nnz) / rows : 0.0;
  int t = 2;
  while (t < 32 && t < mean) t *= 2;
  return t;
}

}  // extern "C"

