// The PDHG device layer for gfx950 behind the C ABI of include/samaya_pdhg.h.
//
// Device state of one LP (all resident in HBM for the whole solve):
//   A by rows (CSR: ptr int32[m+1], idx int32[nnz], val) = the CSC of A' read as rows, and
//   A' by rows (ptr int32[n+1], ...) = the CSC of A read as rows; FP64 values always (the
//   evaluation sums and the FP64 steps), FP32 values too when FP32 steps are asked for; each
//   with its row-block schedule ((first row, first nonzero) int32 pairs, nblocks + 1).
//   Iterates in the step precision T: v[2] and y[2] (current / next, swapped by a parity bit
//   instead of pointers), vbar[2] (the extrapolated point of this and of the next step), the
//   running averages v_erg, y_erg, the restart point v_rst, y_rst, and cost/lower/upper.
//   The model-space data of evaluate (scales, cost and bounds before equilibration), FP64.
//
// One step is two launches of csr_rows_kernel: over the rows of A (dual update, fused with the
// primal update of the row's logical variable) and over the rows of A' (primal update of the
// structural variables). Each also forms the next step's extrapolated point, so the reference's
// separate extrapolate / dual / primal / average kernels and the ax / aty vectors disappear.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/samaya_pdhg.h"
#include "csr_kernels.hpp"
#include "hip_common.hpp"
#include "row_blocks.hpp"

namespace smy {
namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& what) {
  g_last_error = what;
  return code;
}

template <typename F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    return f();
  } catch (const HipError& e) {
    return fail(SMY_ERROR, e.what());
  } catch (const std::exception& e) {
    return fail(SMY_ERROR, e.what());
  } catch (...) {
    return fail(SMY_ERROR, "unknown error");
  }
}

bool device_is_gfx950(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return false;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

int blocks_for(long long count) { return static_cast<int>((count + kBlock - 1) / kBlock); }

struct DevMatrix {
  int rows = 0;
  long long nnz = 0;
  int* ptr = nullptr;
  int* idx = nullptr;
  double* val64 = nullptr;
  float* val32 = nullptr;
  int2* meta = nullptr;  // row-block schedule: nblocks + 1 (first row, first nonzero)
  int nblocks = 0;
};

template <typename T>
struct Iter {
  T* v[2] = {nullptr, nullptr};
  T* y[2] = {nullptr, nullptr};
  T* vbar[2] = {nullptr, nullptr};
  T* v_erg = nullptr;
  T* y_erg = nullptr;
  T* v_rst = nullptr;
  T* y_rst = nullptr;
  T* cost = nullptr;
  T* lower = nullptr;
  T* upper = nullptr;
  bool ready = false;
};

// One workgroup per row block: the hardware dispatcher balances blocks of uneven cost (the long
// rows of a power-law column-count matrix) better than a static persistent assignment, which
// measured 10-15% slower on such matrices; the kernel's next-block prefetch then never triggers.
template <typename V, typename Op>
int rows_grid(const DevMatrix& M) {
  return M.nblocks;
}

// Returns the grid used (0: no rows, nothing launched).
template <typename V, typename Op>
int launch_rows(const DevMatrix& M, const V* val, const Op& op, hipStream_t s) {
  if (M.nblocks == 0) return 0;
  CsrRows<V> A;
  A.rows = M.rows;
  A.nblocks = M.nblocks;
  A.ptr = M.ptr;
  A.idx = M.idx;
  A.val = val;
  A.meta = M.meta;
  const int grid = rows_grid<V, Op>(M);
  hipLaunchKernelGGL((csr_rows_kernel<V, Op>), dim3(grid), dim3(kBlock), 0, s, A, op);
  return grid;
}

}  // namespace
}  // namespace smy

using namespace smy;

struct smy_pdhg {
  int n = 0, m = 0, nt = 0;
  int precision = SMY_FP64;
  DevMatrix a, at;  // A by rows, A' by rows
  Iter<double> d;
  Iter<float> f;
  int p = 0;                 // parity: v = v[p], v_new = v[1 - p] (same for y, vbar)
  double theta_vbar = 1.0;   // theta the current vbar was formed with
  bool vbar_any = true;      // vbar == v (valid for every theta): start, restart
  bool stepped = false;      // a step since the last advance / restart_to
  bool has_original = false;
  double* cs = nullptr;
  double* rs = nullptr;
  double* cost_o = nullptr;
  double* lower_o = nullptr;
  double* upper_o = nullptr;
  double* partial = nullptr;  // 14 x (a.nblocks + at.nblocks) or 2 x 1024
  double* sums_dev = nullptr;
  double* sums_host = nullptr;  // pinned
  StepScalars* scalars = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool timed = false;
  // Captured runs of kGraphSteps steps, by (precision, parity).
  static constexpr int kGraphSteps = 64;
  hipGraph_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipGraphExec_t graph_exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  std::vector<void*> allocations;

  template <typename T>
  T* alloc(long long count) {
    void* ptr = nullptr;
    SMY_HIP(hipMalloc(&ptr, static_cast<size_t>(std::max<long long>(count, 1)) * sizeof(T)));
    allocations.push_back(ptr);
    return static_cast<T*>(ptr);
  }
  template <typename T>
  T* upload(const T* host, long long count) {
    T* dev = alloc<T>(count);
    if (count > 0) SMY_HIP(hipMemcpy(dev, host, static_cast<size_t>(count) * sizeof(T), hipMemcpyHostToDevice));
    return dev;
  }
  template <typename T>
  T* zeros(long long count) {
    T* dev = alloc<T>(count);
    SMY_HIP(hipMemsetAsync(dev, 0, static_cast<size_t>(std::max<long long>(count, 1)) * sizeof(T), stream));
    return dev;
  }

  ~smy_pdhg() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& row : graph_exec)
      for (auto& g : row)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto& row : graph)
      for (auto& g : row)
        if (g) (void)hipGraphDestroy(g);
    for (void* ptr : allocations) (void)hipFree(ptr);
    if (sums_host) (void)hipHostFree(sums_host);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace smy {
namespace {

// CSR view of a CSC matrix read as the rows of its transpose (no data movement on the host
// beyond narrowing the offsets to 32 bits).
void upload_rows(smy_pdhg* h, const smy_csc* csc, DevMatrix& M, bool fp32) {
  M.rows = csc->cols;
  M.nnz = csc->start[csc->cols];
  std::vector<int> ptr(static_cast<size_t>(csc->cols) + 1);
  for (int r = 0; r <= csc->cols; ++r) ptr[r] = static_cast<int>(csc->start[r]);
  const std::vector<RowBlockMeta> meta = build_row_blocks(ptr.data(), M.rows);
  M.nblocks = static_cast<int>(meta.size()) - 1;
  M.ptr = h->upload(ptr.data(), static_cast<long long>(ptr.size()));
  M.idx = h->upload(csc->index, M.nnz);
  M.val64 = h->upload(csc->value, M.nnz);
  static_assert(sizeof(RowBlockMeta) == sizeof(int2), "schedule entries are read as int2");
  M.meta = h->upload(reinterpret_cast<const int2*>(meta.data()), static_cast<long long>(meta.size()));
  if (fp32) {
    M.val32 = h->alloc<float>(M.nnz);
    if (M.nnz > 0)
      hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(M.nnz)), dim3(kBlock), 0,
                         h->stream, M.nnz, M.val64, M.val32);
  }
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
void alloc_iter(smy_pdhg* h, Iter<T>& S, const double* cost, const double* lower,
                const double* upper) {
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
  if (cost != nullptr) {
    if constexpr (std::is_same_v<T, double>) {
      SMY_HIP(hipMemcpy(S.cost, cost, nt * sizeof(double), hipMemcpyHostToDevice));
      SMY_HIP(hipMemcpy(S.lower, lower, nt * sizeof(double), hipMemcpyHostToDevice));
      SMY_HIP(hipMemcpy(S.upper, upper, nt * sizeof(double), hipMemcpyHostToDevice));
    } else {
      std::vector<float> tmp(static_cast<size_t>(nt));
      for (const auto& [src, dst] : {std::pair{cost, S.cost}, std::pair{lower, S.lower},
                                     std::pair{upper, S.upper}}) {
        for (long long k = 0; k < nt; ++k) tmp[k] = static_cast<float>(src[k]);
        if (nt > 0) SMY_HIP(hipMemcpy(dst, tmp.data(), nt * sizeof(float), hipMemcpyHostToDevice));
      }
    }
  }
  S.ready = true;
}

template <typename T>
void step_kernels(smy_pdhg* h, int j) {
  Iter<T>& S = state<T>(h);
  const int p = h->p;
  DualRowsOp<T> dual;
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
  launch_rows(h->a, avals<T>(h->a), dual, h->stream);
  PrimalColsOp<T> primal;
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
  launch_rows(h->at, avals<T>(h->at), primal, h->stream);
}

// Re-forms vbar for a theta other than the one it was formed with (the reference recomputes it
// from v and v_prev at every step; here only a change of theta needs it).
template <typename T>
int refresh_vbar(smy_pdhg* h, double theta) {
  if (h->vbar_any || theta == h->theta_vbar) return SMY_OK;
  if (h->stepped) return fail(SMY_STATE, "theta changed between two steps without advance()");
  Iter<T>& S = state<T>(h);
  hipLaunchKernelGGL((extrapolate_kernel<T>), dim3(blocks_for(h->nt)), dim3(kBlock), 0, h->stream,
                     h->nt, S.v[h->p], S.v[1 - h->p], static_cast<T>(theta), S.vbar[h->p]);
  return SMY_OK;
}

template <typename T>
int step_impl(smy_pdhg* h, double tau, double sigma, double theta, long long iter) {
  const int rc = refresh_vbar<T>(h, theta);
  if (rc != SMY_OK) return rc;
  hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(64), 0, h->stream, h->scalars, tau, sigma,
                     theta, iter);
  step_kernels<T>(h, 0);
  SMY_HIP(hipGetLastError());
  h->theta_vbar = theta;
  h->vbar_any = false;
  h->stepped = true;
  return SMY_OK;
}

// Average duration of each of the two step kernels over `reps` back-to-back launches (HIP events
// on the handle's stream). The launches repeat one step from the current state, so the iterates
// and averages change: a measurement helper, not part of a solve.
template <typename T>
int time_kernels_impl(smy_pdhg* h, int reps, double* dual_ms, double* primal_ms) {
  Iter<T>& S = state<T>(h);
  const int p = h->p;
  hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(64), 0, h->stream, h->scalars, 0.5, 0.5,
                     1.0, 0LL);
  DualRowsOp<T> dual;
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
  PrimalColsOp<T> primal;
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
  float t = 0.0f;
  launch_rows(h->a, avals<T>(h->a), dual, h->stream);  // warm
  SMY_HIP(hipEventRecord(h->ev_start, h->stream));
  for (int r = 0; r < reps; ++r) launch_rows(h->a, avals<T>(h->a), dual, h->stream);
  SMY_HIP(hipEventRecord(h->ev_stop, h->stream));
  SMY_HIP(hipEventSynchronize(h->ev_stop));
  SMY_HIP(hipEventElapsedTime(&t, h->ev_start, h->ev_stop));
  *dual_ms = static_cast<double>(t) / reps;
  launch_rows(h->at, avals<T>(h->at), primal, h->stream);
  SMY_HIP(hipEventRecord(h->ev_start, h->stream));
  for (int r = 0; r < reps; ++r) launch_rows(h->at, avals<T>(h->at), primal, h->stream);
  SMY_HIP(hipEventRecord(h->ev_stop, h->stream));
  SMY_HIP(hipEventSynchronize(h->ev_stop));
  SMY_HIP(hipEventElapsedTime(&t, h->ev_start, h->ev_stop));
  *primal_ms = static_cast<double>(t) / reps;
  SMY_HIP(hipGetLastError());
  h->timed = false;
  return SMY_OK;
}

template <typename T>
int run_impl(smy_pdhg* h, double tau, double sigma, double theta, long long iter0, int count) {
  if (count <= 0) return SMY_OK;
  const int rc = refresh_vbar<T>(h, theta);
  if (rc != SMY_OK) return rc;
  SMY_HIP(hipEventRecord(h->ev_start, h->stream));
  hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(64), 0, h->stream, h->scalars, tau, sigma,
                     theta, iter0);
  const int prec = std::is_same_v<T, double> ? 0 : 1;
  if (count == smy_pdhg::kGraphSteps) {
    const int p0 = h->p;
    hipGraphExec_t& exec = h->graph_exec[prec][p0];
    if (exec == nullptr) {
      hipGraph_t& g = h->graph[prec][p0];
      SMY_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      for (int j = 0; j < count; ++j) {
        step_kernels<T>(h, j);
        if (j + 1 < count) h->p ^= 1;
      }
      SMY_HIP(hipStreamEndCapture(h->stream, &g));
      h->p = p0;
      SMY_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    }
    SMY_HIP(hipGraphLaunch(exec, h->stream));
    if ((count - 1) & 1) h->p ^= 1;
  } else {
    for (int j = 0; j < count; ++j) {
      step_kernels<T>(h, j);
      if (j + 1 < count) h->p ^= 1;
    }
  }
  SMY_HIP(hipGetLastError());
  SMY_HIP(hipEventRecord(h->ev_stop, h->stream));
  h->timed = true;
  h->theta_vbar = theta;
  h->vbar_any = false;
  h->stepped = true;
  return SMY_OK;
}

// Evaluation sums of NV iterates into h->sums_host (7 per iterate). Synchronizes.
template <typename T, int NV>
void measure(smy_pdhg* h, const T* const (&vv)[NV], const T* const (&yy)[NV]) {
  // Each persistent workgroup leaves one partial per sum: at most nblocks of them per matrix.
  double* part1 = h->partial;
  double* part2 = h->partial + static_cast<size_t>(7 * NV) * h->a.nblocks;
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
  rows.partial = part1;
  const int nb1 = launch_rows(h->a, h->a.val64, rows, h->stream);
  MeasureColsOp<T, NV> cols;
  cols.m = h->m;
  for (int v = 0; v < NV; ++v) {
    cols.vv[v] = vv[v];
    cols.yy[v] = yy[v];
  }
  cols.cs = h->cs;
  cols.cost = h->cost_o;
  cols.lower = h->lower_o;
  cols.upper = h->upper_o;
  cols.partial = part2;
  const int nb2 = launch_rows(h->at, h->at.val64, cols, h->stream);
  // Both partial arrays in one fixed-order pass: the rows of A first, then the rows of A'.
  constexpr int K = 7 * NV;
  hipLaunchKernelGGL(finish_sums_kernel, dim3(K), dim3(kBlock), 0, h->stream, nb1, part1,
                     h->sums_dev);
  hipLaunchKernelGGL(finish_sums_kernel, dim3(K), dim3(kBlock), 0, h->stream, nb2, part2,
                     h->sums_dev + K);
  SMY_HIP(hipGetLastError());
  SMY_HIP(hipMemcpyAsync(h->sums_host, h->sums_dev, 2 * K * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream));
  SMY_HIP(hipStreamSynchronize(h->stream));
  for (int q = 0; q < K; ++q) h->sums_host[q] = h->sums_host[q] + h->sums_host[K + q];
}

template <typename T>
int evaluate_impl(smy_pdhg* h, int which, double* out) {
  if (!h->has_original) return fail(SMY_STATE, "set_original() must be called before evaluate()");
  Iter<T>& S = state<T>(h);
  const T* v = which == SMY_CURRENT ? S.v[1 - h->p] : S.v_erg;
  const T* y = which == SMY_CURRENT ? S.y[1 - h->p] : S.y_erg;
  const T* const vv[1] = {v};
  const T* const yy[1] = {y};
  measure<T, 1>(h, vv, yy);
  std::memcpy(out, h->sums_host, 7 * sizeof(double));
  return SMY_OK;
}

template <typename T>
int evaluate_both_impl(smy_pdhg* h, double* out) {
  if (!h->has_original) return fail(SMY_STATE, "set_original() must be called before evaluate()");
  Iter<T>& S = state<T>(h);
  const T* const vv[2] = {S.v[1 - h->p], S.v_erg};
  const T* const yy[2] = {S.y[1 - h->p], S.y_erg};
  measure<T, 2>(h, vv, yy);
  std::memcpy(out, h->sums_host, 14 * sizeof(double));
  return SMY_OK;
}

template <typename T>
int distance_impl(smy_pdhg* h, int which, double* dv, double* dy) {
  Iter<T>& S = state<T>(h);
  const T* v = which == SMY_CURRENT ? S.v[1 - h->p] : S.v_erg;
  const T* y = which == SMY_CURRENT ? S.y[1 - h->p] : S.y_erg;
  const int grid = std::max(1, std::min(1024, blocks_for(h->nt)));
  hipLaunchKernelGGL((distance_kernel<T>), dim3(grid), dim3(kBlock), 0, h->stream, h->nt, h->m, v,
                     S.v_rst, y, S.y_rst, h->partial);
  hipLaunchKernelGGL(finish_sums_kernel, dim3(2), dim3(kBlock), 0, h->stream, grid, h->partial,
                     h->sums_dev);
  SMY_HIP(hipGetLastError());
  SMY_HIP(hipMemcpyAsync(h->sums_host, h->sums_dev, 2 * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream));
  SMY_HIP(hipStreamSynchronize(h->stream));
  *dv = h->sums_host[0];
  *dy = h->sums_host[1];
  return SMY_OK;
}

template <typename T>
int restart_to_impl(smy_pdhg* h, int which) {
  Iter<T>& S = state<T>(h);
  const int p = h->p;
  // The candidate may be v_new / y_new itself: each element is read before it is written, by the
  // same thread.
  const T* cv = which == SMY_CURRENT ? S.v[1 - p] : S.v_erg;
  const T* cy = which == SMY_CURRENT ? S.y[1 - p] : S.y_erg;
  const int grid = blocks_for(std::max(h->nt, h->m));
  hipLaunchKernelGGL((restart_to_kernel<T>), dim3(grid), dim3(kBlock), 0, h->stream, h->nt, h->m,
                     cv, cy, S.v_rst, S.v[p], S.v[1 - p], S.vbar[0], S.vbar[1], S.y_rst, S.y[p],
                     S.y[1 - p]);
  SMY_HIP(hipGetLastError());
  h->vbar_any = true;
  h->stepped = false;
  return SMY_OK;
}

template <typename T>
void download_vec(smy_pdhg* h, const T* src, long long count, double* dst) {
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

template <typename T>
int download_impl(smy_pdhg* h, double* v, double* v_new, double* y, double* y_new, double* v_erg,
                  double* y_erg) {
  Iter<T>& S = state<T>(h);
  const int p = h->p;
  download_vec(h, S.v[p], h->nt, v);
  download_vec(h, S.v[1 - p], h->nt, v_new);
  download_vec(h, S.y[p], h->m, y);
  download_vec(h, S.y[1 - p], h->m, y_new);
  download_vec(h, S.v_erg, h->nt, v_erg);
  download_vec(h, S.y_erg, h->m, y_erg);
  SMY_HIP(hipStreamSynchronize(h->stream));
  return SMY_OK;
}

template <typename From, typename To>
void convert_state(smy_pdhg* h) {
  Iter<From>& A = state<From>(h);
  Iter<To>& B = state<To>(h);
  const auto cv = [&](const From* src, To* dst, long long count) {
    if (count > 0)
      hipLaunchKernelGGL((convert_kernel<From, To>), dim3(blocks_for(count)), dim3(kBlock), 0,
                         h->stream, count, src, dst);
  };
  for (int q = 0; q < 2; ++q) {
    cv(A.v[q], B.v[q], h->nt);
    cv(A.y[q], B.y[q], h->m);
    cv(A.vbar[q], B.vbar[q], h->nt);
  }
  cv(A.v_erg, B.v_erg, h->nt);
  cv(A.y_erg, B.y_erg, h->m);
  cv(A.v_rst, B.v_rst, h->nt);
  cv(A.y_rst, B.y_rst, h->m);
  SMY_HIP(hipGetLastError());
}

template <typename T>
void set_start_impl(smy_pdhg* h, const double* v, const double* y) {
  Iter<T>& S = state<T>(h);
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
  SMY_HIP(hipStreamSynchronize(h->stream));
  for (int q = 0; q < 2; ++q) {
    up(v, S.v[q], h->nt);
    up(v, S.vbar[q], h->nt);
    up(y, S.y[q], h->m);
  }
  up(v, S.v_erg, h->nt);
  up(v, S.v_rst, h->nt);
  up(y, S.y_erg, h->m);
  up(y, S.y_rst, h->m);
  h->vbar_any = true;
  h->stepped = false;
}

}  // namespace
}  // namespace smy

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

const char* smy_last_error(void) { return g_last_error.c_str(); }

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
  if (A == nullptr || At == nullptr || A->start == nullptr || At->start == nullptr)
    return fail(SMY_INVALID, "null argument");
  if (A->rows + A->cols > 0 && (cost == nullptr || lower == nullptr || upper == nullptr))
    return fail(SMY_INVALID, "null argument");
  if (precision != SMY_FP64 && precision != SMY_FP32) return fail(SMY_INVALID, "bad precision");
  if (A->rows != At->cols || A->cols != At->rows || A->rows < 0 || A->cols < 0)
    return fail(SMY_INVALID, "A and At dimensions do not match");
  if (!smy_gpu_available()) return fail(SMY_NO_DEVICE, "no usable gfx950 GPU");
  if (A->start[A->cols] >= INT_MAX || At->start[At->cols] >= INT_MAX)
    return fail(SMY_TOO_LARGE, "matrix too large for 32-bit CSR offsets");
  if (A->start[A->cols] != At->start[At->cols]) return fail(SMY_INVALID, "A and At nnz differ");
  return guarded([&]() -> int {
    auto h = std::make_unique<smy_pdhg>();
    h->m = A->rows;
    h->n = A->cols;
    h->nt = h->n + h->m;
    h->precision = precision;
    SMY_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    SMY_HIP(hipEventCreate(&h->ev_start));
    SMY_HIP(hipEventCreate(&h->ev_stop));
    // A by rows is the CSC of A' read as CSR; A' by rows is the CSC of A.
    upload_rows(h.get(), At, h->a, precision == SMY_FP32);
    upload_rows(h.get(), A, h->at, precision == SMY_FP32);
    alloc_iter(h.get(), h->d, cost, lower, upper);
    if (precision == SMY_FP32) alloc_iter(h.get(), h->f, cost, lower, upper);
    const long long nparts = std::max<long long>(14LL * (h->a.nblocks + h->at.nblocks), 2048);
    h->partial = h->alloc<double>(nparts);
    h->sums_dev = h->alloc<double>(32);
    h->scalars = h->alloc<StepScalars>(1);
    SMY_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->sums_host), 32 * sizeof(double),
                          hipHostMallocDefault));
    SMY_HIP(hipStreamSynchronize(h->stream));
    *out = h.release();
    return SMY_OK;
  });
}

void smy_pdhg_destroy(smy_pdhg* h) { delete h; }

int smy_pdhg_set_original(smy_pdhg* h, const double* col_scale, const double* row_scale,
                          const double* cost, const double* lower, const double* upper) {
  // Empty vectors may come as null pointers (an LP without rows has no row scales).
  if (h == nullptr || (h->n > 0 && col_scale == nullptr) || (h->m > 0 && row_scale == nullptr) ||
      (h->nt > 0 && (cost == nullptr || lower == nullptr || upper == nullptr)))
    return fail(SMY_INVALID, "null argument");
  return guarded([&]() -> int {
    h->cs = h->upload(col_scale, h->n);
    h->rs = h->upload(row_scale, h->m);
    h->cost_o = h->upload(cost, h->nt);
    h->lower_o = h->upload(lower, h->nt);
    h->upper_o = h->upload(upper, h->nt);
    h->has_original = true;
    return SMY_OK;
  });
}

int smy_pdhg_step(smy_pdhg* h, double tau, double sigma, double theta, long long iter) {
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  return guarded([&]() -> int {
    return h->precision == SMY_FP64 ? step_impl<double>(h, tau, sigma, theta, iter)
                                    : step_impl<float>(h, tau, sigma, theta, iter);
  });
}

int smy_pdhg_run(smy_pdhg* h, double tau, double sigma, double theta, long long iter0, int count) {
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  return guarded([&]() -> int {
    return h->precision == SMY_FP64 ? run_impl<double>(h, tau, sigma, theta, iter0, count)
                                    : run_impl<float>(h, tau, sigma, theta, iter0, count);
  });
}

int smy_pdhg_time_kernels(smy_pdhg* h, int reps, double* dual_ms, double* primal_ms) {
  if (h == nullptr || dual_ms == nullptr || primal_ms == nullptr || reps <= 0)
    return fail(SMY_INVALID, "bad argument");
  return guarded([&]() -> int {
    return h->precision == SMY_FP64 ? time_kernels_impl<double>(h, reps, dual_ms, primal_ms)
                                    : time_kernels_impl<float>(h, reps, dual_ms, primal_ms);
  });
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
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  return guarded([&]() -> int {
    const auto go = [&](auto& S) {
      using T = std::remove_pointer_t<decltype(S.cost)>;
      hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks_for(h->nt)), dim3(kBlock), 0, h->stream,
                         static_cast<long long>(h->nt), S.v[h->p], S.v[1 - h->p]);
      hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks_for(h->nt)), dim3(kBlock), 0, h->stream,
                         static_cast<long long>(h->nt), S.v[h->p], S.vbar[h->p]);
    };
    if (h->precision == SMY_FP64) go(h->d); else go(h->f);
    SMY_HIP(hipGetLastError());
    h->vbar_any = true;
    return SMY_OK;
  });
}

int smy_pdhg_download(smy_pdhg* h, double* v, double* v_new, double* y, double* y_new,
                      double* v_erg, double* y_erg) {
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  return guarded([&]() -> int {
    return h->precision == SMY_FP64 ? download_impl<double>(h, v, v_new, y, y_new, v_erg, y_erg)
                                    : download_impl<float>(h, v, v_new, y, y_new, v_erg, y_erg);
  });
}

int smy_pdhg_evaluate(smy_pdhg* h, int which, double out[7]) {
  if (h == nullptr || out == nullptr) return fail(SMY_INVALID, "null argument");
  return guarded([&]() -> int {
    return h->precision == SMY_FP64 ? evaluate_impl<double>(h, which, out)
                                    : evaluate_impl<float>(h, which, out);
  });
}

int smy_pdhg_evaluate_both(smy_pdhg* h, double out[14]) {
  if (h == nullptr || out == nullptr) return fail(SMY_INVALID, "null argument");
  return guarded([&]() -> int {
    return h->precision == SMY_FP64 ? evaluate_both_impl<double>(h, out)
                                    : evaluate_both_impl<float>(h, out);
  });
}

int smy_pdhg_restart_distance(smy_pdhg* h, int which, double* dv_sq, double* dy_sq) {
  if (h == nullptr || dv_sq == nullptr || dy_sq == nullptr) return fail(SMY_INVALID, "null argument");
  return guarded([&]() -> int {
    return h->precision == SMY_FP64 ? distance_impl<double>(h, which, dv_sq, dy_sq)
                                    : distance_impl<float>(h, which, dv_sq, dy_sq);
  });
}

int smy_pdhg_restart_to(smy_pdhg* h, int which) {
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  return guarded([&]() -> int {
    return h->precision == SMY_FP64 ? restart_to_impl<double>(h, which)
                                    : restart_to_impl<float>(h, which);
  });
}

int smy_pdhg_set_precision(smy_pdhg* h, int precision) {
  if (h == nullptr) return fail(SMY_INVALID, "null handle");
  if (precision != SMY_FP64 && precision != SMY_FP32) return fail(SMY_INVALID, "bad precision");
  if (precision == h->precision) return SMY_OK;
  return guarded([&]() -> int {
    if (precision == SMY_FP32) {
      if (!h->f.ready) {
        // FP32 values of both matrices and FP32 iterates, cost and bounds from the FP64 ones.
        for (DevMatrix* M : {&h->a, &h->at}) {
          if (M->val32 == nullptr) {
            M->val32 = h->alloc<float>(M->nnz);
            if (M->nnz > 0)
              hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(M->nnz)),
                                 dim3(kBlock), 0, h->stream, M->nnz, M->val64, M->val32);
          }
        }
        alloc_iter<float>(h, h->f, nullptr, nullptr, nullptr);
        for (auto [src, dst] : {std::pair{h->d.cost, h->f.cost}, std::pair{h->d.lower, h->f.lower},
                                std::pair{h->d.upper, h->f.upper}})
          hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(h->nt)), dim3(kBlock),
                             0, h->stream, static_cast<long long>(h->nt), src, dst);
      }
      convert_state<double, float>(h);
    } else {
      convert_state<float, double>(h);
    }
    SMY_HIP(hipStreamSynchronize(h->stream));
    h->precision = precision;
    return SMY_OK;
  });
}

int smy_pdhg_set_start(smy_pdhg* h, const double* v, const double* y) {
  if (h == nullptr || (v == nullptr && h->nt > 0) || (y == nullptr && h->m > 0))
    return fail(SMY_INVALID, "null argument");
  return guarded([&]() -> int {
    if (h->precision == SMY_FP64) set_start_impl<double>(h, v, y); else set_start_impl<float>(h, v, y);
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

// ---------------------------------------------------------------------------------------------
// Standalone SpMV (the reference's csr_spmv) and prepared plans.
// ---------------------------------------------------------------------------------------------
struct smy_spmv_plan {
  int rows = 0;
  int precision = SMY_FP64;
  const int* ptr = nullptr;
  const int* idx = nullptr;
  const void* val = nullptr;
  int2* meta = nullptr;
  int nblocks = 0;
  ~smy_spmv_plan() {
    if (meta) (void)hipFree(meta);
  }
};

namespace smy {
namespace {

int plan_run(smy_spmv_plan* p, const void* x, void* y, hipStream_t s) {
  DevMatrix M;
  M.rows = p->rows;
  M.ptr = const_cast<int*>(p->ptr);
  M.idx = const_cast<int*>(p->idx);
  M.meta = p->meta;
  M.nblocks = p->nblocks;
  if (p->precision == SMY_FP64) {
    SpmvOp<double, double> op{static_cast<const double*>(x), static_cast<double*>(y)};
    launch_rows(M, static_cast<const double*>(p->val), op, s);
  } else {
    SpmvOp<float, float> op{static_cast<const float*>(x), static_cast<float*>(y)};
    launch_rows(M, static_cast<const float*>(p->val), op, s);
  }
  SMY_HIP(hipGetLastError());
  return SMY_OK;
}

int plan_make(int rows, const int* ptr_host, const int* ptr, const int* idx, const void* val,
              int precision, smy_spmv_plan** out) {
  auto p = std::make_unique<smy_spmv_plan>();
  p->rows = rows;
  p->precision = precision;
  p->ptr = ptr;
  p->idx = idx;
  p->val = val;
  const std::vector<RowBlockMeta> meta = build_row_blocks(ptr_host, rows);
  p->nblocks = static_cast<int>(meta.size()) - 1;
  SMY_HIP(hipMalloc(&p->meta, meta.size() * sizeof(int2)));
  SMY_HIP(hipMemcpy(p->meta, meta.data(), meta.size() * sizeof(int2), hipMemcpyHostToDevice));
  *out = p.release();
  return SMY_OK;
}

}  // namespace
}  // namespace smy

extern "C" {

int smy_spmv_plan_create(int rows, const int* ptr_host, const int* ptr_dev, const int* idx_dev,
                         const void* val_dev, int precision, smy_spmv_plan** out) {
  if (out == nullptr || ptr_host == nullptr || rows < 0) return fail(SMY_INVALID, "bad argument");
  *out = nullptr;
  if (!smy_gpu_available()) return fail(SMY_NO_DEVICE, "no usable gfx950 GPU");
  return guarded([&]() -> int { return plan_make(rows, ptr_host, ptr_dev, idx_dev, val_dev, precision, out); });
}

int smy_spmv_plan_run(smy_spmv_plan* p, const void* x_dev, void* y_dev, void* stream) {
  if (p == nullptr) return fail(SMY_INVALID, "null plan");
  return guarded([&]() -> int { return plan_run(p, x_dev, y_dev, static_cast<hipStream_t>(stream)); });
}

void smy_spmv_plan_destroy(smy_spmv_plan* p) { delete p; }

static int spmv_once(int rows, const int* ptr, const int* idx, const void* val, const void* x,
                     void* y, void* stream, int precision) {
  if (rows < 0 || (rows > 0 && (ptr == nullptr || x == nullptr || y == nullptr)))
    return fail(SMY_INVALID, "bad argument");
  if (!smy_gpu_available()) return fail(SMY_NO_DEVICE, "no usable gfx950 GPU");
  if (rows == 0) return SMY_OK;
  return guarded([&]() -> int {
    const hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<int> ptr_host(static_cast<size_t>(rows) + 1);
    SMY_HIP(hipMemcpyAsync(ptr_host.data(), ptr, ptr_host.size() * sizeof(int),
                           hipMemcpyDeviceToHost, s));
    SMY_HIP(hipStreamSynchronize(s));
    smy_spmv_plan* raw = nullptr;
    const int rc = plan_make(rows, ptr_host.data(), ptr, idx, val, precision, &raw);
    std::unique_ptr<smy_spmv_plan> plan(raw);
    if (rc != SMY_OK) return rc;
    plan_run(plan.get(), x, y, s);
    SMY_HIP(hipStreamSynchronize(s));
    return SMY_OK;
  });
}

int smy_csr_spmv(int rows, const int* ptr, const int* idx, const double* val, const double* x,
                 double* y, void* stream, int threads_per_row) {
  (void)threads_per_row;
  return spmv_once(rows, ptr, idx, val, x, y, stream, SMY_FP64);
}

int smy_csr_spmv_f32(int rows, const int* ptr, const int* idx, const float* val, const float* x,
                     float* y, void* stream) {
  return spmv_once(rows, ptr, idx, val, x, y, stream, SMY_FP32);
}

}  // extern "C"
