// smy_csr_spmv_noalloc: the standalone CSR product (smy_csr_spmv / smy_csr_spmv_f32) for callers
// that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). No counterpart in the reference.
//
// It runs the same kernels over the same device-built schedule (sched_spmv.hpp), so it gives the
// same bits, but its workspace (schedule, per-workgroup counts, block count) is the file-scope
// device array `scratch` below instead of a per-call allocation. Each GPU holds its own copy of
// the array until the process ends; host code takes its device address with
// hipGetSymbolAddress once per device, with that device current, and passes only that address
// to kernels. The workspace limits the matrix to kNoAllocMaxRows rows; the call rejects more.
// Every kernel writes the part of the workspace it reads before reading it, so nothing is
// cleared between calls. Calls that overlap on one GPU (on different streams or threads) share
// the array and race; calls on one stream are ordered by it.
//
// The entry point calls no hipMalloc, hipFree or other allocating routine, throws nothing, and
// reports failures through the fixed last-error buffer.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "../../../include/samaya_pdhg.h"
#include "pdhg_handle.hpp"
#include "sched_spmv.hpp"

namespace smy {
namespace {

constexpr int kNoAllocMaxRows = 1 << 20;
constexpr size_t kNoAllocBytes = sched::sched_layout(kNoAllocMaxRows).bytes;  // ~8 MiB

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::atomic<unsigned char*> g_scratch[kMaxDevices];
std::mutex g_scratch_mutex;

// The device address of `scratch` on `device` (the current device), looked up on the first call
// on that device.
unsigned char* scratch_address(int device, hipError_t* err) {
  *err = hipSuccess;
  unsigned char* p = g_scratch[device].load(std::memory_order_acquire);
  if (p != nullptr) return p;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  p = g_scratch[device].load(std::memory_order_relaxed);
  if (p == nullptr) {
    void* addr = nullptr;
    *err = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    if (*err != hipSuccess) return nullptr;
    p = static_cast<unsigned char*>(addr);
    g_scratch[device].store(p, std::memory_order_release);
  }
  return p;
}

}  // namespace
}  // namespace smy

extern "C" {

int smy_csr_spmv_noalloc(int rows, const int* ptr, const int* idx, const void* val,
                         const void* x, void* y, int precision, void* stream) {
  using namespace smy;
  if (rows < 0 || (rows > 0 && (ptr == nullptr || idx == nullptr || val == nullptr || x == nullptr ||
                                y == nullptr)))
    return fail(SMY_INVALID, "bad argument");
  if (precision != SMY_FP64 && precision != SMY_FP32) return fail(SMY_INVALID, "bad precision");
  if (rows > kNoAllocMaxRows)
    return fail(SMY_TOO_LARGE,
                "smy_csr_spmv_noalloc: more than 1048576 rows (the size of its fixed workspace)");
  g_last_error[0] = '\0';
  if (rows == 0) return SMY_OK;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices)
    return fail(SMY_NO_DEVICE, "no usable gfx950 GPU");
  hipError_t e = hipSuccess;
  unsigned char* work = scratch_address(device, &e);
  if (work == nullptr) return fail(SMY_ERROR, hipGetErrorString(e));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  e = precision == SMY_FP64
          ? sched::sched_spmv(rows, ptr, idx, static_cast<const double*>(val),
                              static_cast<const double*>(x), static_cast<double*>(y), work, s)
          : sched::sched_spmv(rows, ptr, idx, static_cast<const float*>(val),
                              static_cast<const float*>(x), static_cast<float*>(y), work, s);
  if (e != hipSuccess) return fail(SMY_ERROR, hipGetErrorString(e));
  return SMY_OK;
}

}  // extern "C"
