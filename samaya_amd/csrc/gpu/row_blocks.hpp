// Row-block schedule for the LDS-staged CSR kernels (host side).
//
// A matrix stored by rows is cut into blocks of consecutive rows; one 256-thread workgroup
// processes one block. A block holds at most kBlockNnz nonzeros and at most kBlockRows rows, so
// its products fit in LDS and every thread of the workgroup streams its share of the block's
// values and indices with fully coalesced loads. A row longer than kBlockNnz gets a block of its
// own, which the kernel processes with a strided loop over the row instead of LDS staging.
// The schedule depends only on the sparsity pattern, so every sum the kernels form is taken in
// the same order on every run.
#pragma once

#include <cstdint>
#include <vector>

namespace smy {

constexpr int kBlockNnz = 1024;  // Nonzeros staged in LDS per workgroup.
constexpr int kBlockRows = 256;  // Rows per workgroup (one epilogue thread per row).

inline std::vector<int> build_row_blocks(const int* ptr, int rows, int cap = kBlockNnz,
                                         int max_rows = kBlockRows) {
  std::vector<int> starts{0};
  int r = 0;
  while (r < rows) {
    if (ptr[r + 1] - ptr[r] > cap) {
      ++r;
      starts.push_back(r);
      continue;
    }
    const int first = r;
    long long nnz = 0;
    while (r < rows && r - first < max_rows) {
      const int len = ptr[r + 1] - ptr[r];
      if (len > cap || nnz + len > cap) break;
      nnz += len;
      ++r;
    }
    starts.push_back(r);
  }
  return starts;
}

}  // namespace smy
