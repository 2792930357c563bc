#include "lp_data.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>

namespace smy::host {

Csc csc_from_triplets(int rows, int cols, std::vector<Triplet> entries) {
  // src/linalg/sparse_matrix.cpp:15-46.
  for (const Triplet& t : entries)
    if (t.row < 0 || t.row >= rows || t.col < 0 || t.col >= cols)
      throw std::invalid_argument("triplet (" + std::to_string(t.row) + ", " +
                                  std::to_string(t.col) + ") out of range");
  std::sort(entries.begin(), entries.end(), [](const Triplet& a, const Triplet& b) {
    return a.col != b.col ? a.col < b.col : a.row < b.row;
  });
  Csc out;
  out.rows = rows;
  out.cols = cols;
  out.start.assign(static_cast<size_t>(cols) + 1, 0);
  out.index.reserve(entries.size());
  out.value.reserve(entries.size());
  size_t k = 0;
  for (int j = 0; j < cols; ++j) {
    while (k < entries.size() && entries[k].col == j) {
      const int i = entries[k].row;
      double sum = 0.0;
      for (; k < entries.size() && entries[k].col == j && entries[k].row == i; ++k)
        sum += entries[k].value;
      if (sum != 0.0) {
        out.index.push_back(i);
        out.value.push_back(sum);
      }
    }
    out.start[static_cast<size_t>(j) + 1] = static_cast<int64_t>(out.index.size());
  }
  return out;
}

Csc csc_transpose(const Csc& a) {
  // src/linalg/sparse_matrix.cpp:48-67: count, prefix sum, scatter by increasing column.
  Csc t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.start.assign(static_cast<size_t>(a.rows) + 1, 0);
  t.index.resize(a.index.size());
  t.value.resize(a.value.size());
  for (int i : a.index) ++t.start[static_cast<size_t>(i) + 1];
  std::partial_sum(t.start.begin(), t.start.end(), t.start.begin());
  std::vector<int64_t> next(t.start.begin(), t.start.end() - 1);
  for (int j = 0; j < a.cols; ++j)
    for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p) {
      const int64_t q = next[a.index[p]]++;
      t.index[q] = j;
      t.value[q] = a.value[p];
    }
  return t;
}

Csc csc_scaled(const Csc& a, const std::vector<double>& row, const std::vector<double>& col) {
  // The entries keep their (sorted) positions; a product that underflows to zero is dropped, as
  // the reference's rebuild through from_triplets drops it (pdlp.cpp:190-209, scaling.cpp:78-90).
  Csc out;
  out.rows = a.rows;
  out.cols = a.cols;
  out.start.assign(static_cast<size_t>(a.cols) + 1, 0);
  out.index.reserve(a.index.size());
  out.value.reserve(a.value.size());
  for (int j = 0; j < a.cols; ++j) {
    for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p) {
      const double v = a.value[p] * row[a.index[p]] * col[j];
      if (v != 0.0) {
        out.index.push_back(a.index[p]);
        out.value.push_back(v);
      }
    }
    out.start[static_cast<size_t>(j) + 1] = static_cast<int64_t>(out.index.size());
  }
  return out;
}

namespace {

double nearest_power_of_two(double v) {
  if (!(v > 0.0) || !std::isfinite(v)) return 1.0;
  return std::ldexp(1.0, static_cast<int>(std::lround(std::log2(v))));
}

double clamp_factor(double v) { return std::clamp(v, 0x1p-40, 0x1p40); }

double spread(const Csc& a, const Scaling& s) {
  double lo = kInf, hi = 0.0;
  for (int j = 0; j < a.cols; ++j)
    for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p) {
      const double v = std::fabs(a.value[p]) * s.row[a.index[p]] * s.col[j];
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
  return hi > 0.0 ? hi / lo : 1.0;
}

}  // namespace

Scaling compute_scaling(const Csc& a) {
  // src/linalg/scaling.cpp:130-200 with its default options: up to 10 geometric passes (stop when
  // a pass gains less than 10% or makes things worse), then equilibration, then powers of two.
  Scaling s;
  s.row.assign(static_cast<size_t>(a.rows), 1.0);
  s.col.assign(static_cast<size_t>(a.cols), 1.0);
  if (a.nnz() == 0) return s;
  std::vector<double> row_lo(static_cast<size_t>(a.rows)), row_hi(static_cast<size_t>(a.rows));
  const auto scan_rows = [&] {
    std::fill(row_lo.begin(), row_lo.end(), kInf);
    std::fill(row_hi.begin(), row_hi.end(), 0.0);
    for (int j = 0; j < a.cols; ++j)
      for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p) {
        const double v = std::fabs(a.value[p]) * s.col[j];
        row_lo[a.index[p]] = std::min(row_lo[a.index[p]], v);
        row_hi[a.index[p]] = std::max(row_hi[a.index[p]], v);
      }
  };
  double ratio = spread(a, s);
  for (int pass = 0; pass < 10 && ratio > 1.0; ++pass) {
    const Scaling keep = s;
    scan_rows();
    for (int i = 0; i < a.rows; ++i)
      if (row_hi[i] > 0.0) s.row[i] = clamp_factor(1.0 / std::sqrt(row_lo[i] * row_hi[i]));
    for (int j = 0; j < a.cols; ++j) {
      double lo = kInf, hi = 0.0;
      for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p) {
        const double v = std::fabs(a.value[p]) * s.row[a.index[p]];
        lo = std::min(lo, v);
        hi = std::max(hi, v);
      }
      if (hi > 0.0) s.col[j] = clamp_factor(1.0 / std::sqrt(lo * hi));
    }
    const double next = spread(a, s);
    if (next > ratio) {
      s = keep;
      break;
    }
    const bool little = next > 0.9 * ratio;
    ratio = next;
    if (little) break;
  }
  scan_rows();
  for (int i = 0; i < a.rows; ++i)
    if (row_hi[i] > 0.0) s.row[i] = clamp_factor(1.0 / row_hi[i]);
  for (int j = 0; j < a.cols; ++j) {
    double hi = 0.0;
    for (int64_t p = a.start[j]; p < a.start[j + 1]; ++p)
      hi = std::max(hi, std::fabs(a.value[p]) * s.row[a.index[p]]);
    if (hi > 0.0) s.col[j] = clamp_factor(1.0 / hi);
  }
  for (double& f : s.row) f = nearest_power_of_two(f);
  for (double& f : s.col) f = nearest_power_of_two(f);
  return s;
}

LpProblem make_problem(const Model& model, const Scaling& s) {
  // src/lp/lp_solver.cpp:28-51: maximization negates the cost; logical n + i carries row i's
  // bounds with cost 0.
  const double sense = model.maximize ? -1.0 : 1.0;
  LpProblem lp;
  lp.m = model.m;
  lp.n = model.n;
  lp.A = csc_scaled(model.A, s.row, s.col);
  lp.At = csc_transpose(lp.A);
  const size_t nt = static_cast<size_t>(model.n + model.m);
  lp.cost.assign(nt, 0.0);
  lp.lower.resize(nt);
  lp.upper.resize(nt);
  for (int j = 0; j < model.n; ++j) {
    lp.cost[j] = sense * model.obj[j] * s.col[j];
    lp.lower[j] = model.col_lower[j] / s.col[j];
    lp.upper[j] = model.col_upper[j] / s.col[j];
  }
  for (int i = 0; i < model.m; ++i) {
    lp.lower[model.n + i] = model.row_lower[i] * s.row[i];
    lp.upper[model.n + i] = model.row_upper[i] * s.row[i];
  }
  return lp;
}

}  // namespace smy::host
