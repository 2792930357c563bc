// Shared HIP helpers for the gfx950 kernels: error checks that turn a HIP failure into an
// exception (caught at the C ABI and turned into an error code), and wavefront-64 reductions in
// a fixed order.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace smy {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] inline void hip_fail(hipError_t e, const char* call, const char* file, int line) {
  throw HipError(std::string("HIP error at ") + file + ":" + std::to_string(line) + " in " + call +
                 ": " + hipGetErrorString(e));
}

}  // namespace smy

#define SMY_HIP(call)                                                      \
  do {                                                                     \
    const hipError_t smy_e_ = (call);                                      \
    if (smy_e_ != hipSuccess) ::smy::hip_fail(smy_e_, #call, __FILE__, __LINE__); \
  } while (0)

namespace smy {

constexpr int kWave = 64;     // CDNA wavefront width.
constexpr int kBlock = 256;   // Threads per workgroup for every streaming kernel (4 wavefronts).

// Butterfly sum over aligned groups of G lanes (G a power of two <= 64): every lane of a group
// ends with the same total, summed in the same order on every run.
template <int G, typename T>
__device__ inline T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

template <int G, typename T>
__device__ inline T group_max(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, G));
  return v;
}

template <typename T>
__device__ inline T group_sum_dyn(T v, int g) {
  for (int off = g >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
}  // namespace smy
