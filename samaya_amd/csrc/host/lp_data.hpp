// Host-side LP data for the PDLP path: a CSC matrix, the model, the computational form and the
// power-of-two scaling that produces it.
//
// Mirrors samaya's SparseMatrix (include/samaya/sparse_matrix.hpp:36-69), Model
// (include/samaya/model.hpp:41-80, LP fields only), LpProblem (src/lp/simplex.hpp:18-26),
// Scaling / compute_scaling (src/linalg/scaling.hpp:18-46) and make_problem
// (src/lp/lp_solver.cpp:28-51), with the same meaning of every field.
#pragma once

#include <cstdint>
#include <limits>
#include <vector>

namespace smy::host {

constexpr double kInf = std::numeric_limits<double>::infinity();

struct Csc {
  int rows = 0;
  int cols = 0;
  std::vector<int64_t> start{0};
  std::vector<int> index;
  std::vector<double> value;
  int64_t nnz() const { return static_cast<int64_t>(value.size()); }
};

struct Triplet {
  int row;
  int col;
  double value;
};

// CSC from unordered triplets: sorted by (column, row), duplicates summed, exact zeros dropped.
Csc csc_from_triplets(int rows, int cols, std::vector<Triplet> entries);
// CSC of the transpose (the CSR of the matrix); indices stay sorted.
Csc csc_transpose(const Csc& a);
// diag(row) * a * diag(col), same pattern (entries that round to zero are dropped).
Csc csc_scaled(const Csc& a, const std::vector<double>& row, const std::vector<double>& col);

struct Model {
  int m = 0;
  int n = 0;
  bool maximize = false;
  double offset = 0.0;
  Csc A;  // m x n
  std::vector<double> obj, col_lower, col_upper, row_lower, row_upper;
};

// min cost'v  s.t.  [A -I] v = 0,  lower <= v <= upper,  v = (x, r): n structural then m logicals.
struct LpProblem {
  int m = 0;
  int n = 0;
  Csc A;
  Csc At;
  std::vector<double> cost, lower, upper;  // n + m
};

struct Scaling {
  std::vector<double> col;
  std::vector<double> row;
};

// Iterated geometric-mean scaling then one equilibration pass, rounded to powers of two.
Scaling compute_scaling(const Csc& A);
LpProblem make_problem(const Model& model, const Scaling& s);

}  // namespace smy::host
