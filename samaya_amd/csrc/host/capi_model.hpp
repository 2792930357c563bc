// The model handle behind the C ABI (smy_model of include/samaya_pdlp.h): the LP, and the names
// and integrality an MPS file gives it (empty for models built from arrays or generators).
#pragma once

#include <string>
#include <vector>

#include "lp_data.hpp"

struct smy_model {
  smy::host::Model model;
  std::string name;
  std::vector<std::string> row_names, col_names;
  std::vector<char> col_integer;  // n entries (1 = integer) or empty
};
