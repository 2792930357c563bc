// Measurement helper: the ceiling of the products' memory pattern on this GPU. A product streams
// values and column indices and gathers one vector entry per nonzero. Four kernels time that
// pattern without any row structure, over `n` synthetic entries and a vector of `len` entries:
//   0  stream only (values and indices, 4 + s bytes per entry);
//   1  stream + gathers at random positions in the whole vector;
//   2  stream + gathers at random positions inside a 64K-entry window (L2-resident);
//   3  stream + gathers at sorted positions (consecutive lanes, coalesced).
// The product kernels' times against these show what bounds them. Not part of a solve.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/samaya_pdhg.h"
#include "hip_common.hpp"

namespace smy {
namespace {

__device__ inline uint64_t mix64(uint64_t z) {  // splitmix64's finalizer (Steele et al. 2014)
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ void probe_fill(long long n, int len, int kind, T* __restrict__ val,
                           int* __restrict__ idx, T* __restrict__ x) {
  const long long end = n > len ? n : len;
  for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < end;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (k < n) {
      const uint64_t h = mix64(static_cast<uint64_t>(k));
      val[k] = static_cast<T>(1.0 + static_cast<double>(h & 7));
      const int window = len < 65536 ? len : 65536;
      idx[k] = kind == 2 ? static_cast<int>(h % static_cast<uint64_t>(window))
             : kind == 3 ? static_cast<int>((k * static_cast<long long>(len)) / n)
                         : static_cast<int>(h % static_cast<uint64_t>(len));
    }
    if (k < len) x[k] = static_cast<T>(static_cast<double>(k % 13));
  }
}

template <typename T, bool Gather>
__global__ void probe_sum(long long n, const T* __restrict__ val, const int* __restrict__ idx,
                          const T* __restrict__ x, T* __restrict__ out) {
  T s = T(0);
  for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    const T a = __builtin_nontemporal_load(val + k);
    const int c = __builtin_nontemporal_load(idx + k);
    s += Gather ? a * x[c] : a * static_cast<T>(c & 1);
  }
  out[blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x] = s;
}

template <typename T>
int probe(long long n, int len, double* us) {
  constexpr int kGrid = 4096, kReps = 20;
  T *val = nullptr, *x = nullptr, *out = nullptr;
  int* idx = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = SMY_OK;
  const auto fail_if = [&](hipError_t e) {
    if (e != hipSuccess && rc == SMY_OK) rc = SMY_ERROR;
    return e == hipSuccess;
  };
  if (fail_if(hipMalloc(&val, n * sizeof(T))) && fail_if(hipMalloc(&idx, n * sizeof(int))) &&
      fail_if(hipMalloc(&x, static_cast<size_t>(len) * sizeof(T))) &&
      fail_if(hipMalloc(&out, static_cast<size_t>(kGrid) * kBlock * sizeof(T))) &&
      fail_if(hipEventCreate(&a)) && fail_if(hipEventCreate(&b))) {
    for (int kind = 0; kind < 4 && rc == SMY_OK; ++kind) {
      hipLaunchKernelGGL((probe_fill<T>), dim3(kGrid), dim3(kBlock), 0, nullptr, n, len,
                         kind == 0 ? 1 : kind, val, idx, x);
      const auto go = [&] {
        if (kind == 0)
          hipLaunchKernelGGL((probe_sum<T, false>), dim3(kGrid), dim3(kBlock), 0, nullptr, n, val, idx,
                             x, out);
        else
          hipLaunchKernelGGL((probe_sum<T, true>), dim3(kGrid), dim3(kBlock), 0, nullptr, n, val, idx,
                             x, out);
      };
      go();  // warm
      float ms = 0.0f;
      if (fail_if(hipEventRecord(a, nullptr))) {
        for (int r = 0; r < kReps; ++r) go();
        if (fail_if(hipEventRecord(b, nullptr)) && fail_if(hipEventSynchronize(b)) &&
            fail_if(hipEventElapsedTime(&ms, a, b)))
          us[kind] = 1e3 * ms / kReps;
      }
      fail_if(hipGetLastError());
    }
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (val) (void)hipFree(val);
  if (idx) (void)hipFree(idx);
  if (x) (void)hipFree(x);
  if (out) (void)hipFree(out);
  return rc;
}

}  // namespace
}  // namespace smy

extern "C" int smy_probe_gather(long long n, int len, int precision, double* us) {
  if (n <= 0 || len <= 0 || us == nullptr || (precision != SMY_FP64 && precision != SMY_FP32))
    return SMY_INVALID;
  for (int k = 0; k < 4; ++k) us[k] = 0.0;
  return precision == SMY_FP64 ? smy::probe<double>(n, len, us) : smy::probe<float>(n, len, us);
}
