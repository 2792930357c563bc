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

// Multi-period blending LP with a block-staircase structure: `periods` periods (0: drawn from
// U{4..48}); in each, `units` blending units hold a tank of each of `streams` streams (use, buy,
// inventory per tank and period, inventories linking each period to the next through the tank
// balances), unit balances, capacities with weights U[0.5, 1.5], and `qualities` quality rows per
// unit (0: drawn from U{1..30}) with stream qualities U[0.5, 1.5] against a limit U[0.9, 1.3] on
// the unit's output. Feasible by construction around x0 ~ U[0, 5]; costs U[-1, 1]; 0 <= x <= 10.
// With 20 units and 30 streams: 660..1240 rows per period, 2.6k..60k rows.
Model generate_blending(uint64_t seed, int periods, int units, int streams, int qualities);

// Multi-dimensional knapsack LP: maximize profit'x (profits U{1..100}) subject to W x <= 0.4 x
// (row sums of W), 0 <= x <= 1, with per_col weights U{1..100} per column at distinct uniform rows.
Model generate_knapsack(uint64_t seed, int m, int n, int per_col);

// Branch-and-bound fixings of a bound vector: a fraction U[frac_min, frac_max] of the n columns
// (distinct, uniform) fixed at their lower or their upper bound (even odds); the other columns
// keep lower / upper.
void generate_fixings(uint64_t seed, int n, double frac_min, double frac_max, const double* lower,
                      const double* upper, double* out_lower, double* out_upper);

}  // namespace smy::host
