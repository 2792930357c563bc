// MPS input and plain-text output of the LP path: the drop-in for samaya's read_mps
// (include/samaya/io.hpp, src/io/mps_reader.cpp) and its CLI's --solution writer
// (apps/cli/main.cpp:137-198).
//
// The reader parses the COLUMNS section (nearly all of a large model) on several threads: the
// section is cut at line boundaries, each part resolves its row names in the shared row table and
// keeps its entries by column run, and one pass then numbers the columns in file order. Its
// conventions are the reference's (free MPS first, then fixed MPS; the first N row is the
// objective unless OBJNAME names another, other N rows are dropped; an RHS on the objective sets
// the offset to minus its value; columns default to [0, +inf); UP < 0 on a column whose lower
// bound was not set makes the lower -inf; |value| >= 1e30 is infinite; duplicate entries are
// summed, zero entries dropped). Quadratic sections (QUADOBJ, QMATRIX, QCMATRIX) belong to the QP
// path and are rejected.
#pragma once

#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "lp_data.hpp"

namespace smy::host {

class MpsError : public std::runtime_error {
 public:
  MpsError(const std::string& message, long long line)
      : std::runtime_error(line > 0 ? "line " + std::to_string(line) + ": " + message : message),
        line_(line) {}
  long long line() const { return line_; }

 private:
  long long line_;
};

struct MpsModel {
  Model model;
  std::string name;
  std::vector<std::string> row_names, col_names;
  std::vector<char> col_integer;  // integrality markers and BV / LI / UI bounds (1 = integer)
};

// threads <= 0: the machine's hardware threads (at most 16).
MpsModel read_mps_text(std::string_view text, int threads = 0);
MpsModel read_mps_file(const std::string& path, int threads = 0);

// Free MPS of a model: the names in `names` (R0.. / C0.. where it has none, or where a name holds a
// blank), integer markers from its col_integer, numbers in the shortest form that reads back
// exactly and infinities as +-1e30. Reading it back gives the same model, except that a ranged
// row's far end is recomputed as rhs +- |range| (written from the end that makes it exact where
// one does).
std::string write_mps(const Model& model, const MpsModel* names = nullptr);

// The reference CLI's --solution file: status, verified, objective, then one line per column
// (name value [reduced_cost]) and per row (name activity [dual]), and a certificate or ray when
// given (vectors may be empty).
std::string solution_text(const std::string& status, bool verified, double objective,
                          const std::vector<std::string>& col_names,
                          const std::vector<std::string>& row_names,
                          const std::vector<double>& col_value, const std::vector<double>& col_dual,
                          const std::vector<double>& row_activity,
                          const std::vector<double>& row_dual,
                          const std::vector<double>& certificate,
                          const std::vector<double>& ray);

}  // namespace smy::host
