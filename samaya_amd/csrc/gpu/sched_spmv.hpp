// Row-block schedule built on the device, and the product over it, for the standalone CSR
// products (spmv.hip, spmv_noalloc.hip).
//
// Row r starts a block when r == 0, r % 256 == 0, row r - 1 or row r is longer than 512
// nonzeros, or ptr[r] and ptr[r - 1] fall in different 512-nonzero buckets (ptr / 512). Rows of
// one block then start within one bucket and are at most 512 long, so a block holds at most
// 1,023 nonzeros and 256 rows: the LDS-staged block product of csr_kernels.hpp applies, and a
// longer row (its own block) takes its strided path. The rule needs no sequential scan: three
// small kernels build the schedule (flags counted per workgroup, one workgroup scans the counts,
// the entries written in row order), and the product runs over it with csr_rows_block. Every
// row's sum depends only on its block, so any two calls on the same matrix give the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include "csr_kernels.hpp"
#include "hip_common.hpp"
#include "row_blocks.hpp"

namespace smy {
namespace sched {
namespace {  // internal linkage: each source that includes this has its own kernels

constexpr int kSchedBucket = 512;  // a block's rows start within one bucket of this many nonzeros
constexpr int kSchedLong = 512;    // longer rows get a block of their own
constexpr int kScanThreads = 1024;
constexpr int kSpmvGrid = 8192;    // workgroups of the product (each strides over the blocks)

// Workspace of one device-scheduled product over `rows` rows, laid out in one buffer:
// the schedule (rows + 1 pairs), one count per workgroup of the flag kernels, the block count.
struct SchedLayout {
  size_t meta_off = 0, counts_off = 0, nblocks_off = 0, bytes = 0;
  int groups = 0;
};

constexpr size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

constexpr SchedLayout sched_layout(long long rows) {
  SchedLayout L;
  L.groups = static_cast<int>((rows + kBlock - 1) / kBlock);
  L.meta_off = 0;
  L.counts_off = align_up(static_cast<size_t>(rows + 1) * sizeof(int2), 256);
  L.nblocks_off = align_up(L.counts_off + static_cast<size_t>(L.groups) * sizeof(int), 256);
  L.bytes = L.nblocks_off + 256;
  return L;
}

__device__ inline bool starts_block(const int* __restrict__ ptr, int r) {
  if (r == 0 || r % kBlockRows == 0) return true;
  const int p0 = ptr[r - 1], p1 = ptr[r], p2 = ptr[r + 1];
  if (p1 - p0 > kSchedLong || p2 - p1 > kSchedLong) return true;
  return p1 / kSchedBucket != p0 / kSchedBucket;
}

__global__ __launch_bounds__(kBlock) void sched_count_kernel(int rows, const int* __restrict__ ptr,
                                                             int* __restrict__ counts) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  const int f = r < rows && starts_block(ptr, r) ? 1 : 0;
  const int total = __syncthreads_count(f);
  if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the per-workgroup counts in place, in chunks of 1,024, the
// total into *nblocks and the closing schedule entry (rows, nnz).
__global__ __launch_bounds__(kScanThreads) void sched_scan_kernel(int groups, int rows,
                                                                  const int* __restrict__ ptr,
                                                                  int* __restrict__ counts,
                                                                  int2* __restrict__ meta,
                                                                  int* __restrict__ nblocks) {
  __shared__ int part[kScanThreads];
  __shared__ int carry;
  const int t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < groups; base += kScanThreads) {
    const int g = base + t;
    const int v = g < groups ? counts[g] : 0;
    part[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {  // Hillis-Steele inclusive scan
      const int add = t >= off ? part[t - off] : 0;
      __syncthreads();
      part[t] += add;
      __syncthreads();
    }
    if (g < groups) counts[g] = carry + part[t] - v;
    __syncthreads();
    if (t == kScanThreads - 1) carry += part[t];
    __syncthreads();
  }
  if (t == 0) {
    *nblocks = carry;
    meta[carry] = make_int2(rows, ptr[rows]);
  }
}

__global__ __launch_bounds__(kBlock) void sched_write_kernel(int rows, const int* __restrict__ ptr,
                                                             const int* __restrict__ offsets,
                                                             int2* __restrict__ meta) {
  __shared__ int wave_total[kBlock / kWave];
  const int r = blockIdx.x * kBlock + threadIdx.x;
  const bool f = r < rows && starts_block(ptr, r);
  const unsigned long long mask = __ballot(f);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int below = __popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) wave_total[wave] = __popcll(mask);
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wave_total[w];
  if (f) meta[offsets[blockIdx.x] + before + below] = make_int2(r, ptr[r]);
}

// The product over a device-built schedule: each workgroup strides over the blocks.
template <typename V>
__global__ __launch_bounds__(kBlock) void sched_spmv_kernel(CsrRows<V> A, SpmvOp<V, V> op,
                                                            const int* __restrict__ nblocks) {
  const int count = *nblocks;
  for (int b = blockIdx.x; b < count; b += gridDim.x) {
    if (b != static_cast<int>(blockIdx.x)) __syncthreads();  // the last block's LDS reads are done
    csr_rows_block(A, op, b, 0);
  }
}

// Launches the schedule build and the product into `work` (sched_layout(rows).bytes). Returns a
// HIP error code instead of throwing, and allocates nothing.
template <typename V>
hipError_t sched_spmv(int rows, const int* ptr, const int* idx, const V* val, const V* x, V* y,
                      unsigned char* work, hipStream_t s) {
  const SchedLayout L = sched_layout(rows);
  int2* meta = reinterpret_cast<int2*>(work + L.meta_off);
  int* counts = reinterpret_cast<int*>(work + L.counts_off);
  int* nblocks = reinterpret_cast<int*>(work + L.nblocks_off);
  hipLaunchKernelGGL(sched_count_kernel, dim3(L.groups), dim3(kBlock), 0, s, rows, ptr, counts);
  hipLaunchKernelGGL(sched_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, L.groups, rows, ptr,
                     counts, meta, nblocks);
  hipLaunchKernelGGL(sched_write_kernel, dim3(L.groups), dim3(kBlock), 0, s, rows, ptr, counts,
                     meta);
  CsrRows<V> A;
  A.rows = rows;
  A.ptr = ptr;
  A.idx = idx;
  A.val = val;
  A.meta = meta;
  SpmvOp<V, V> op{x, y};
  const int grid = rows < kSpmvGrid ? rows : kSpmvGrid;  // at most one block per row
  hipLaunchKernelGGL((sched_spmv_kernel<V>), dim3(grid), dim3(kBlock), 0, s, A, op, nblocks);
  return hipGetLastError();
}

}  // namespace
}  // namespace sched
}  // namespace smy
