// Standalone CSR products on gfx950: the drop-in for the reference's gpu::csr_spmv
// (src/gpu/spmv.hpp:16-21, spmv.cu:91-120) and its prepared plans.
//
// smy_csr_spmv / smy_csr_spmv_f32 take a CSR matrix already in device memory and build its
// row-block schedule on the device (sched_spmv.hpp), so no copy of the row offsets goes to the
// host; the schedule's workspace is allocated per call. smy_spmv_plan_* keep the host-built
// greedy schedule of row_blocks.hpp (the one the PDHG handle's step kernels use), from row
// offsets the caller also holds on the host.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../../../include/samaya_pdhg.h"
#include "pdhg_handle.hpp"
#include "sched_spmv.hpp"

namespace smy {
namespace {

using sched::sched_layout;
using sched::sched_spmv;

int spmv_alloc(int rows, const int* ptr, const int* idx, const void* val, const void* x, void* y,
               void* stream, int precision) {
  if (rows < 0 || (rows > 0 && (ptr == nullptr || idx == nullptr || val == nullptr || x == nullptr ||
                                y == nullptr)))
    return fail(SMY_INVALID, "bad argument");
  if (!smy_gpu_available()) return fail(SMY_NO_DEVICE, "no usable gfx950 GPU");
  if (rows == 0) return SMY_OK;
  return guarded([&]() -> int {
    const hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned char* work = nullptr;
    SMY_HIP(hipMalloc(&work, sched_layout(rows).bytes));
    std::unique_ptr<unsigned char, decltype(&hipFree)> hold(work, &hipFree);
    const hipError_t e =
        precision == SMY_FP64
            ? sched_spmv(rows, ptr, idx, static_cast<const double*>(val), static_cast<const double*>(x),
                         static_cast<double*>(y), work, s)
            : sched_spmv(rows, ptr, idx, static_cast<const float*>(val), static_cast<const float*>(x),
                         static_cast<float*>(y), work, s);
    SMY_HIP(e);
    SMY_HIP(hipStreamSynchronize(s));
    return SMY_OK;
  });
}

}  // namespace
}  // namespace smy

// ---------------------------------------------------------------------------------------------
// Prepared plans over the host-built greedy schedule.
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
    launch_rows(M, static_cast<const double*>(p->val), nullptr, M.nblocks, op, s);
  } else {
    SpmvOp<float, float> op{static_cast<const float*>(x), static_cast<float*>(y)};
    launch_rows(M, static_cast<const float*>(p->val), nullptr, M.nblocks, op, s);
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

int smy_csr_spmv(int rows, const int* ptr, const int* idx, const double* val, const double* x,
                 double* y, void* stream, int threads_per_row) {
  (void)threads_per_row;
  return spmv_alloc(rows, ptr, idx, val, x, y, stream, SMY_FP64);
}

int smy_csr_spmv_f32(int rows, const int* ptr, const int* idx, const float* val, const float* x,
                     float* y, void* stream) {
  return spmv_alloc(rows, ptr, idx, val, x, y, stream, SMY_FP32);
}

}  // extern "C"
