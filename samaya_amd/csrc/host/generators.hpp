// Deterministic synthetic LP generators (see generators.cpp for the random-number transforms).
#pragma once

#include <cstdint>

#include "lp_data.hpp"

namespace smy::host {

// m rows, n columns, exactly per_col nonzeros per column at distinct uniform rows, values N(0,1);
// feasible by construction around x0 ~ U[0,5]: even rows equalities, odd rows <= with slack
// U[0,1]; costs U[-1,1]; 0 <= x <= 10.
Model generate_sparse_uniform(uint64_t seed, int m, int n, int per_col);

// The same rows and objective with column counts drawn from a power law P(k) ~ k^-a on
// [2, kmax] whose mean is `mean` (column 0 gets kmax), values N(0,1) x 10^U[-2,2].
Model generate_sparse_powerlaw(uint64_t seed, int m, int n, int kmax, double mean);

}  // namespace smy::host
