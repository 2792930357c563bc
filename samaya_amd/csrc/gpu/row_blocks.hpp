// Row-block schedule for the LDS-staged CSR kernels (host side).
//
// A matrix stored by rows is cut into blocks of consecutive rows. A block holds at most
// kBlockNnz nonzeros and at most kBlockRows rows, so its products fit in LDS and every thread of
// a 256-thread workgroup streams kBlockNnz / 256 of the block's values and indices with fully
// coalesced loads. A row longer than kLongRow gets a block of its own, which the kernel sums with
// a strided loop over the whole workgroup instead of LDS staging (so no lane group ever reduces
// more than kLongRow / g products serially). Each block is described by the pair (first row,
// first nonzero); the next pair closes it. The schedule depends only on the sparsity pattern, so
// every sum the kernels form is taken in the same order on every run.
#pragma once

#include <cstdint>
#include <vector>

namespace smy {

constexpr int kBlockNnz = 1024;  // Nonzeros staged in LDS per block.
constexpr int kBlockRows = 256;  // Rows per block (one epilogue thread per row).
constexpr int kLongRow = kBlockNnz;  // Longer rows get a block of their own.

struct RowBlockMeta {
  int row;  // first row of the block
  int nz;   // offset of its first nonzero
};

inline std::vector<RowBlockMeta> build_row_blocks(const int* ptr, int rows) {
  std::vector<RowBlockMeta> meta;
  int r = 0;
  while (r < rows) {
    meta.push_back({r, ptr[r]});
    if (ptr[r + 1] - ptr[r] > kLongRow) {
      ++r;
      continue;
    }
    const int first = r;
    long long nnz = 0;
    while (r < rows && r - first < kBlockRows) {
      const int len = ptr[r + 1] - ptr[r];
      if (len > kLongRow || nnz + len > kBlockNnz) break;
      nnz += len;
      ++r;
    }
  }
  meta.push_back({rows, ptr[rows]});
  return meta;
}

}  // namespace smy
