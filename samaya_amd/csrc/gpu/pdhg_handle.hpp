// The device handle behind include/samaya_pdhg.h, shared by the two device sources: the stacked
// batch (pdhg_device.hip: any LPs, block-diagonal) and the shared-matrix batch (pdhg_shared.hip:
// one A and cost, per-instance column bounds, iterates stored instance-minor for SpMM products).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/samaya_pdhg.h"
#include "csr_kernels.hpp"
#include "detect_kernels.hpp"
#include "hip_common.hpp"
#include "row_blocks.hpp"
#include "rule_kernels.hpp"

namespace smy {

// The calling thread's last error message, in a fixed buffer: reporting a failure allocates
// nothing (the allocation-free SpMV entry point reports through it too).
inline thread_local char g_last_error[512] = {0};

inline int fail(int code, const char* what) {
  std::snprintf(g_last_error, sizeof g_last_error, "%s", what);
  return code;
}

inline int fail(int code, const std::string& what) { return fail(code, what.c_str()); }

template <typename F>
int guarded(F&& f) {
  try {
    g_last_error[0] = '\0';
    return f();
  } catch (const HipError& e) {
    return fail(SMY_ERROR, e.what());
  } catch (const std::exception& e) {
    return fail(SMY_ERROR, e.what());
  } catch (...) {
    return fail(SMY_ERROR, "unknown error");
  }
}

inline bool device_is_gfx950(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return false;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

inline int blocks_for(long long count) { return static_cast<int>((count + kBlock - 1) / kBlock); }

constexpr int kFrozen = -1;
constexpr int kMaxBatch = 16384;  // instances per handle (grid y of the per-instance reductions)

// How a product launch maps row blocks to workgroups (CsrRows::per_wg, CsrRows::xcd).
struct Tile {
  int per_wg = 1;
  bool xcd = false;
};

// Size bins of a stacked batch's instances, by their row blocks in a matrix (about 1,024
// nonzeros each): < 64, < 512, the rest. A block list holds the bins one after the other.
constexpr int kTileBins = SMY_TILE_BINS;
inline int tile_bin(int blocks) { return blocks < 64 ? 0 : blocks < 512 ? 1 : 2; }

struct DevMatrix {
  int rows = 0;
  long long nnz = 0;
  int* ptr = nullptr;
  int* idx = nullptr;
  double* val64 = nullptr;
  float* val32 = nullptr;
  int2* meta = nullptr;   // row-block schedule: nblocks + 1 (first row, first nonzero)
  int* binst = nullptr;   // instance of each block
  int nblocks = 0;
  int* list[2] = {nullptr, nullptr};  // blocks of the instances stepping in FP64 / FP32
  int nlist[2] = {0, 0};
  Tile tile[2];                 // by value precision (SMY_FP64, SMY_FP32): chosen by the tuner
  bool tuned[2] = {false, false};
  // Per size bin: where each bin's blocks start in list[q] (bin_start[q][kTileBins] = nlist[q]),
  // and, when `split`, the launches of list[q] go bin by bin, each with its own tile.
  int bin_start[2][kTileBins + 1] = {};
  Tile bin_tile[2][kTileBins];
  bool split[2] = {false, false};
  // One instance: launches pass no block-instance table, so a workgroup knows its instance (0)
  // without a dependent load and reads the instance's step state in parallel with its schedule.
  bool single = false;
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
  // Adaptive steps only (allocated at their first run): A x of both slots of v, and the spare
  // slot of the averages.
  T* ax[2] = {nullptr, nullptr};
  T* v_erg2 = nullptr;
  T* y_erg2 = nullptr;
};

struct Layout {
  int n = 0, m = 0;
  int ox = 0, oy = 0;    // first structural variable and first row in the stack
  int ba0 = 0, ba1 = 0;  // its blocks in the A schedule
  int bt0 = 0, bt1 = 0;  // its blocks in the A' schedule
  int prec = SMY_FP64;   // precision of its steps, or kFrozen
  int last_prec = SMY_FP64;  // precision of its state (also when frozen)
  bool has_original = false;
  // Every structural variable has two finite model-space bounds: the LP is bounded, so it has no
  // improving ray, and its detection skips the ray test (whose failure is then certain).
  bool boxed = false;
};

// Whether the first n entries of both bound vectors are finite.
inline bool all_boxed(int n, const double* lower, const double* upper) {
  for (int j = 0; j < n; ++j)
    if (!(lower[j] > -INFINITY && upper[j] < INFINITY)) return false;
  return true;
}

template <typename T>
constexpr int prec_of() {
  return std::is_same_v<T, double> ? SMY_FP64 : SMY_FP32;
}

// Returns the grid used (0: nothing launched). The tile is the matrix's for values of type V; the
// blocks of list[q] with values of precision q go bin by bin when the tuner split them.
template <typename V, typename Op>
int launch_rows(const DevMatrix& M, const V* val, const int* list, int nlist, const Op& op,
                hipStream_t s, const Tile* force = nullptr) {
  if (nlist == 0) return 0;
  constexpr int q = prec_of<V>();
  if (force == nullptr && M.split[q] && list == M.list[q] && nlist == M.nlist[q]) {
    int grid = 0;
    for (int b = 0; b < kTileBins; ++b) {
      const int first = M.bin_start[q][b], count = M.bin_start[q][b + 1] - first;
      if (count > 0) grid += launch_rows(M, val, list + first, count, op, s, &M.bin_tile[q][b]);
    }
    return grid;
  }
  const Tile t = force != nullptr ? *force : M.tile[q];
  CsrRows<V> A;
  A.rows = M.rows;
  A.nblocks = M.nblocks;
  A.ptr = M.ptr;
  A.idx = M.idx;
  A.val = val;
  A.meta = M.meta;
  A.binst = M.single ? nullptr : M.binst;
  A.list = list;
  A.count = nlist;
  A.per_wg = t.per_wg;
  A.xcd = t.xcd;
  const int grid = (nlist + t.per_wg - 1) / t.per_wg;
  if (t.per_wg == 1 && !t.xcd)
    hipLaunchKernelGGL((csr_rows_kernel<V, Op, false>), dim3(grid), dim3(kBlock), 0, s, A, op);
  else
    hipLaunchKernelGGL((csr_rows_kernel<V, Op, true>), dim3(grid), dim3(kBlock), 0, s, A, op);
  return grid;
}

struct SharedBatch;
void shared_free(SharedBatch* s);

}  // namespace smy

using namespace smy;

struct smy_pdhg {
  int K = 1;
  // Test hook (smy_pdhg_inject_fault): steps run so far by the step entry points, and the step
  // at which they start failing (-1: never).
  long long steps_run = 0;
  long long fault_at = -1;
  int n = 0, m = 0, nt = 0;
  std::vector<Layout> inst;
  DevMatrix a, at;  // A by rows, A' by rows
  Iter<double> d;
  Iter<float> f;
  int p = 0;                 // parity: v = v[p], v_new = v[1 - p] (same for y, vbar)
  double theta_vbar = 1.0;   // theta the current vbar was formed with
  bool vbar_any = true;      // vbar == v (valid for every theta): start, restart
  bool stepped = false;      // a step since the last advance / restart_to
  double* cs = nullptr;
  double* rs = nullptr;
  double* cost_o = nullptr;
  double* lower_o = nullptr;
  double* upper_o = nullptr;
  double* partial = nullptr;  // 14 x (a.nblocks + at.nblocks): block b of A', column a.nblocks + b
  int2* ranges = nullptr;     // per instance: its A blocks, then its A' blocks (partial columns)
  double* sums_dev = nullptr;   // K x 2 x 14
  double* sums_host = nullptr;  // pinned
  StepScalars* scalars = nullptr;
  StepScalars* scalars_host = nullptr;  // pinned staging
  int* flags = nullptr;                 // 2 x K per-instance flags (by precision)
  int* flags_host = nullptr;            // pinned staging
  hipEvent_t ev_staged = nullptr;       // the last copy out of the pinned staging buffers
  bool staged = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool timed = false;
  // Captured runs of kGraphSteps steps, by (precisions in use, parity); valid for one generation
  // of the block lists.
  static constexpr int kGraphSteps = 64;
  struct Captured {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int gen = -1;
  };
  Captured graphs[4][2];
  Captured agraphs[4][2];    // the same for runs of adaptive attempts
  Captured hgraphs[4][2];    // the same for runs of Halpern steps
  // Restarted Halpern steps instead of PDHG steps with averages (smy_pdhg_set_halpern; stacked
  // only), with reflection rho.
  bool halpern = false;
  double rho = 1.0;
  // The fixed-point residual sums formed with the last evaluation (valid until the next step).
  std::vector<double> fixed_cache;
  bool fixed_fresh = false;
  int list_gen = 0;
  bool lists_dirty = false;  // an instance changed precision or froze since the lists were built
  int* step_prec = nullptr;  // per instance: precision of its steps, or kFrozen (device copy)
  // Adaptive steps: per-instance rule state (device, pinned staging), and whether ax holds A x of
  // the current iterates.
  AdaptState* adapt = nullptr;
  AdaptState* adapt_host = nullptr;
  std::vector<int> ax_stale;  // per instance: A x must be formed again before its next attempt
  int* ax_flags = nullptr;    // device copy for the refresh launch
  double* decide_slices = nullptr;     // per instance and slice: the decision's 3 partial totals
  unsigned* decide_arrivals = nullptr;  // per instance: slices done (0 between attempts)
  int decide_nslices = 1;               // slices per instance of the decision kernel
  // Restart rule on the device (allocated by smy_pdhg_rule_init).
  RuleState* rule = nullptr;
  RuleState* rule_host = nullptr;    // pinned
  RuleInput* rule_in = nullptr;
  RuleInput* rule_in_host = nullptr;  // pinned staging
  double* rule_dist = nullptr;        // per instance and segment: ||dv||^2, ||dy||^2
  // FP32 steps accumulate their products in FP64 (smy_pdhg_set_fp64_accumulation; stacked only).
  bool acc64 = false;
  // Tile tuner: on by default; its scratch output (max(m, n) doubles, at first use).
  bool autotune = true;
  double* tune_out = nullptr;
  SharedBatch* shared = nullptr;  // the shared-matrix batch (then the stacked fields are unused)
  // Device-side infeasibility / unboundedness tests (allocated at their first use).
  double* det_partial = nullptr;   // detection slots x blocks
  double* det_scratch = nullptr;   // per structural variable and instance
  double* det_ray = nullptr;
  double* det_prm = nullptr;       // per instance: drop, y_max, d_max
  double* det_prm_host = nullptr;  // pinned
  double* det_out = nullptr;       // per instance: the detection slots
  double* det_out_host = nullptr;  // pinned
  int2* det_ranges = nullptr;      // per instance: its blocks in the detection partials
  // The Farkas test's dual when some instance tests a sign projection (flags 2, 3): a copy in
  // the layout of y_erg per precision (at first use), and whether the running detection reads it.
  double* det_y64 = nullptr;
  float* det_y32 = nullptr;
  bool det_projected = false;
  bool det_y_ready = false;  // det_y holds the duals of a detection with flags above 1
  // Whether the running detection takes the ray test: some instance with flag 1 is not boxed (the
  // ray slots of the others stay 0, so their test fails as it would).
  bool det_ray_test = true;
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
  void drop_graph(Captured& c) {
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c = Captured{};
  }

  ~smy_pdhg() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (shared) shared_free(shared);
    for (auto& row : graphs)
      for (auto& c : row) drop_graph(c);
    for (auto& row : agraphs)
      for (auto& c : row) drop_graph(c);
    for (auto& row : hgraphs)
      for (auto& c : row) drop_graph(c);
    for (void* ptr : allocations) (void)hipFree(ptr);
    if (sums_host) (void)hipHostFree(sums_host);
    if (adapt_host) (void)hipHostFree(adapt_host);
    if (rule_host) (void)hipHostFree(rule_host);
    if (rule_in_host) (void)hipHostFree(rule_in_host);
    if (det_prm_host) (void)hipHostFree(det_prm_host);
    if (det_out_host) (void)hipHostFree(det_out_host);
    if (scalars_host) (void)hipHostFree(scalars_host);
    if (flags_host) (void)hipHostFree(flags_host);
    if (ev_staged) (void)hipEventDestroy(ev_staged);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace smy {

// The pinned staging buffers may be rewritten once the copy that last read them has run.
inline void staging_wait(smy_pdhg* h) {
  if (h->staged) {
    SMY_HIP(hipEventSynchronize(h->ev_staged));
    h->staged = false;
  }
}

inline void staging_sent(smy_pdhg* h) {
  SMY_HIP(hipEventRecord(h->ev_staged, h->stream));
  h->staged = true;
}

// Average duration of one launch step over `reps` back-to-back repeats (after one untimed), from
// one HIP event pair around them on the handle's stream. Repeats keep the dispatch pipelined as a
// solve does; an event between every two launches would add its own serialization to each
// (measured: 5-8 us on a 90 us kernel).
template <typename Step>
double time_repeated(smy_pdhg* h, int reps, Step&& step) {
  step();
  SMY_HIP(hipEventRecord(h->ev_start, h->stream));
  for (int r = 0; r < reps; ++r) step();
  SMY_HIP(hipEventRecord(h->ev_stop, h->stream));
  SMY_HIP(hipEventSynchronize(h->ev_stop));
  float t = 0.0f;
  SMY_HIP(hipEventElapsedTime(&t, h->ev_start, h->ev_stop));
  return static_cast<double>(t) / reps;
}

// The shared-matrix batch (pdhg_shared.hip): the batch entry points of a handle whose `shared` is
// set go here.
int shared_create(const smy_csc* A, const smy_csc* At, int count, const double* cost,
                  const double* const* lower, const double* const* upper, int precision,
                  smy_pdhg** out);
int shared_set_original(smy_pdhg* h, int k, const double* cs, const double* rs, const double* cost,
                        const double* lower, const double* upper);
int shared_run(smy_pdhg* h, const double* tau, const double* sigma, double theta,
               const long long* k0, int count);
int shared_evaluate(smy_pdhg* h, double* out);
int shared_fixed_point(smy_pdhg* h, double* out);
int shared_distance(smy_pdhg* h, const int* which, double* out);
int shared_restart_to(smy_pdhg* h, const int* which);
int shared_set_precision(smy_pdhg* h, const int* precision);
int shared_freeze(smy_pdhg* h, int k);
int shared_download(smy_pdhg* h, int k, double* v, double* v_new, double* y, double* y_new,
                    double* v_erg, double* y_erg);
int shared_set_start(smy_pdhg* h, int k, const double* v, const double* y);
int shared_time_kernels(smy_pdhg* h, int reps, double* dual_ms, double* primal_ms);
int shared_detect(smy_pdhg* h, const int* flags, double* out);
int shared_download_ray(smy_pdhg* h, int k, double* ray);
int shared_download_farkas(smy_pdhg* h, int k, double* y);
int shared_rule_decide(smy_pdhg* h, const smy_pdhg_rule_in* in, long long iter1, smy_pdhg_rule_out* out);

}  // namespace smy
