// Deterministic synthetic LP families for tests and benchmarks.
//
// Random numbers come from std::mt19937_64 (whose output the C++ standard fixes bit for bit)
// through this file's own transforms: u01 = (x >> 11) 2^-53, integers below n by rejection,
// normals by Box-Muller (one per draw, the cosine branch). std::*_distribution is not used
// because its output differs between standard libraries.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <unordered_set>
#include <vector>

#include "generators.hpp"

namespace smy::host {
namespace {

class Rng {
 public:
  explicit Rng(uint64_t seed) : g_(seed) {}
  double u01() { return static_cast<double>(g_() >> 11) * 0x1p-53; }
  double uniform(double a, double b) { return a + (b - a) * u01(); }
  double normal() {
    const double u1 = 1.0 - u01();  // (0, 1]
    const double u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  uint64_t below(uint64_t n) {
    const uint64_t limit = UINT64_MAX - UINT64_MAX % n;
    uint64_t x = g_();
    while (x >= limit) x = g_();
    return x % n;
  }

 private:
  std::mt19937_64 g_;
};

// k distinct integers of [0, m), sorted (Floyd's sampling).
void distinct_sorted(Rng& rng, int m, int k, std::vector<int>& out) {
  out.clear();
  if (k <= 64) {
    for (int j = m - k; j < m; ++j) {
      const int t = static_cast<int>(rng.below(static_cast<uint64_t>(j) + 1));
      out.push_back(std::find(out.begin(), out.end(), t) == out.end() ? t : j);
    }
  } else {
    std::unordered_set<int> seen;
    seen.reserve(static_cast<size_t>(k) * 2);
    for (int j = m - k; j < m; ++j) {
      const int t = static_cast<int>(rng.below(static_cast<uint64_t>(j) + 1));
      const int pick = seen.count(t) ? j : t;
      seen.insert(pick);
      out.push_back(pick);
    }
  }
  std::sort(out.begin(), out.end());
}

// Feasible-by-construction rows and objective: b = A x0 with x0 ~ U[0, 5]; even rows are
// equalities, odd rows are <= b + U[0, 1]; costs U[-1, 1]; columns boxed in [0, 10].
void finish_random_lp(Rng& rng, Model& md) {
  const int m = md.m, n = md.n;
  std::vector<double> x0(static_cast<size_t>(n));
  for (double& t : x0) t = rng.uniform(0.0, 5.0);
  std::vector<double> b(static_cast<size_t>(m), 0.0);
  for (int j = 0; j < n; ++j)
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) b[md.A.index[p]] += md.A.value[p] * x0[j];
  md.row_lower.resize(static_cast<size_t>(m));
  md.row_upper.resize(static_cast<size_t>(m));
  for (int i = 0; i < m; ++i) {
    if (i % 2 == 0) {
      md.row_lower[i] = md.row_upper[i] = b[i];
    } else {
      md.row_lower[i] = -kInf;
      md.row_upper[i] = b[i] + rng.uniform(0.0, 1.0);
    }
  }
  md.obj.resize(static_cast<size_t>(n));
  for (double& c : md.obj) c = rng.uniform(-1.0, 1.0);
  md.col_lower.assign(static_cast<size_t>(n), 0.0);
  md.col_upper.assign(static_cast<size_t>(n), 10.0);
}

// Exponent of a power law P(k) ~ k^-a on [kmin, kmax] whose mean is `mean` (bisection).
double power_law_exponent(int kmin, int kmax, double mean) {
  const auto mean_of = [&](double a) {
    double s0 = 0.0, s1 = 0.0;
    for (int k = kmin; k <= kmax; ++k) {
      const double w = std::pow(static_cast<double>(k), -a);
      s0 += w;
      s1 += w * k;
    }
    return s1 / s0;
  };
  double lo = 0.0, hi = 6.0;  // mean decreases with the exponent
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mean_of(mid) > mean) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace

Model generate_sparse_uniform(uint64_t seed, int m, int n, int per_col) {
  if (m <= 0 || n <= 0 || per_col < 0 || per_col > m) throw std::invalid_argument("bad sizes");
  Rng rng(seed);
  Model md;
  md.m = m;
  md.n = n;
  md.A.rows = m;
  md.A.cols = n;
  md.A.start.assign(static_cast<size_t>(n) + 1, 0);
  md.A.index.reserve(static_cast<size_t>(n) * per_col);
  md.A.value.reserve(static_cast<size_t>(n) * per_col);
  std::vector<int> rows;
  for (int j = 0; j < n; ++j) {
    distinct_sorted(rng, m, per_col, rows);
    for (int i : rows) {
      md.A.index.push_back(i);
      md.A.value.push_back(rng.normal());
    }
    md.A.start[j + 1] = static_cast<int64_t>(md.A.index.size());
  }
  finish_random_lp(rng, md);
  return md;
}

Model generate_sparse_powerlaw(uint64_t seed, int m, int n, int kmax, double mean) {
  if (m <= 0 || n <= 0 || kmax < 2 || kmax > m) throw std::invalid_argument("bad sizes");
  Rng rng(seed);
  const double a = power_law_exponent(2, kmax, mean);
  std::vector<double> cdf(static_cast<size_t>(kmax - 1));
  double total = 0.0;
  for (int k = 2; k <= kmax; ++k) total += std::pow(static_cast<double>(k), -a);
  double run = 0.0;
  for (int k = 2; k <= kmax; ++k) {
    run += std::pow(static_cast<double>(k), -a) / total;
    cdf[k - 2] = run;
  }
  std::vector<int> count(static_cast<size_t>(n));
  for (int j = 0; j < n; ++j) {
    const double u = rng.u01();
    const size_t pos = static_cast<size_t>(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    count[j] = 2 + static_cast<int>(std::min(pos, cdf.size() - 1));
  }
  count[0] = kmax;  // the largest column is realized
  Model md;
  md.m = m;
  md.n = n;
  md.A.rows = m;
  md.A.cols = n;
  md.A.start.assign(static_cast<size_t>(n) + 1, 0);
  long long nnz = 0;
  for (int k : count) nnz += k;
  md.A.index.reserve(static_cast<size_t>(nnz));
  md.A.value.reserve(static_cast<size_t>(nnz));
  std::vector<int> rows;
  for (int j = 0; j < n; ++j) {
    distinct_sorted(rng, m, count[j], rows);
    for (int i : rows) {
      md.A.index.push_back(i);
      md.A.value.push_back(rng.normal() * std::pow(10.0, rng.uniform(-2.0, 2.0)));
    }
    md.A.start[j + 1] = static_cast<int64_t>(md.A.index.size());
  }
  finish_random_lp(rng, md);
  return md;
}

Model generate_blending(uint64_t seed, int periods, int units, int streams, int qualities) {
  if (periods < 0 || units <= 0 || streams <= 0 || qualities < 0) throw std::invalid_argument("bad sizes");
  Rng rng(seed);
  const int T = periods > 0 ? periods : 4 + static_cast<int>(rng.below(45));     // U{4..48}
  const int Q = qualities > 0 ? qualities : 1 + static_cast<int>(rng.below(30));  // U{1..30}
  const int U = units, S = streams;
  // Stream qualities and weights, each unit's quality limits.
  std::vector<double> qual(static_cast<size_t>(S) * Q), weight(static_cast<size_t>(S));
  for (double& t : qual) t = rng.uniform(0.5, 1.5);
  for (double& t : weight) t = rng.uniform(0.5, 1.5);
  std::vector<double> spec(static_cast<size_t>(U) * Q);
  for (double& t : spec) t = rng.uniform(0.9, 1.3);

  // Columns of period t: use(t,u,s), buy(t,u,s), inv(t,u,s) (unit u's tank of stream s), then
  // out(t,u). Rows of period t: tank balances (U S, equality: inv(t-1,u,s) + buy(t,u,s) -
  // use(t,u,s) - inv(t,u,s)), unit balances (U, equality: sum_s use(t,u,s) - out(t,u)),
  // capacities (U, <=: sum_s w_s use(t,u,s)), quality rows (U Q, <=:
  // sum_s q_sk use(t,u,s) - spec_uk out(t,u)).
  const int US = U * S;
  const int cols_per = 3 * US + U;
  const int rows_per = US + 2 * U + U * Q;
  Model md;
  md.m = T * rows_per;
  md.n = T * cols_per;
  const auto use = [&](int t, int u, int s) { return t * cols_per + u * S + s; };
  const auto buy = [&](int t, int u, int s) { return t * cols_per + US + u * S + s; };
  const auto inv = [&](int t, int u, int s) { return t * cols_per + 2 * US + u * S + s; };
  const auto out = [&](int t, int u) { return t * cols_per + 3 * US + u; };
  std::vector<Triplet> e;
  std::vector<char> equality(static_cast<size_t>(md.m), 0);
  for (int t = 0; t < T; ++t) {
    const int r0 = t * rows_per;
    for (int u = 0; u < U; ++u) {
      for (int s = 0; s < S; ++s) {
        const int r = r0 + u * S + s;
        equality[r] = 1;
        if (t > 0) e.push_back({r, inv(t - 1, u, s), 1.0});
        e.push_back({r, buy(t, u, s), 1.0});
        e.push_back({r, use(t, u, s), -1.0});
        e.push_back({r, inv(t, u, s), -1.0});
      }
      const int rb = r0 + US + u, rc = r0 + US + U + u;
      equality[rb] = 1;
      for (int s = 0; s < S; ++s) {
        e.push_back({rb, use(t, u, s), 1.0});
        e.push_back({rc, use(t, u, s), weight[s]});
      }
      e.push_back({rb, out(t, u), -1.0});
      for (int q = 0; q < Q; ++q) {
        const int rq = r0 + US + 2 * U + u * Q + q;
        for (int s = 0; s < S; ++s) e.push_back({rq, use(t, u, s), qual[static_cast<size_t>(s) * Q + q]});
        e.push_back({rq, out(t, u), -spec[static_cast<size_t>(u) * Q + q]});
      }
    }
  }
  md.A = csc_from_triplets(md.m, md.n, std::move(e));
  // Feasible by construction around x0 ~ U[0, 5]: equalities at A x0, the other rows <= A x0 +
  // U[0, 1]; costs U[-1, 1]; 0 <= x <= 10.
  std::vector<double> x0(static_cast<size_t>(md.n));
  for (double& t : x0) t = rng.uniform(0.0, 5.0);
  std::vector<double> b(static_cast<size_t>(md.m), 0.0);
  for (int j = 0; j < md.n; ++j)
    for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) b[md.A.index[p]] += md.A.value[p] * x0[j];
  md.row_lower.resize(static_cast<size_t>(md.m));
  md.row_upper.resize(static_cast<size_t>(md.m));
  for (int i = 0; i < md.m; ++i) {
    if (equality[i]) {
      md.row_lower[i] = md.row_upper[i] = b[i];
    } else {
      md.row_lower[i] = -kInf;
      md.row_upper[i] = b[i] + rng.uniform(0.0, 1.0);
    }
  }
  md.obj.resize(static_cast<size_t>(md.n));
  for (double& c : md.obj) c = rng.uniform(-1.0, 1.0);
  md.col_lower.assign(static_cast<size_t>(md.n), 0.0);
  md.col_upper.assign(static_cast<size_t>(md.n), 10.0);
  return md;
}

Model generate_knapsack(uint64_t seed, int m, int n, int per_col) {
  if (m <= 0 || n <= 0 || per_col <= 0 || per_col > m) throw std::invalid_argument("bad sizes");
  Rng rng(seed);
  Model md;
  md.m = m;
  md.n = n;
  md.maximize = true;
  md.A.rows = m;
  md.A.cols = n;
  md.A.start.assign(static_cast<size_t>(n) + 1, 0);
  md.A.index.reserve(static_cast<size_t>(n) * per_col);
  md.A.value.reserve(static_cast<size_t>(n) * per_col);
  std::vector<int> rows;
  std::vector<double> row_sum(static_cast<size_t>(m), 0.0);
  for (int j = 0; j < n; ++j) {
    distinct_sorted(rng, m, per_col, rows);
    for (int i : rows) {
      const double w = static_cast<double>(1 + rng.below(100));  // U{1..100}
      md.A.index.push_back(i);
      md.A.value.push_back(w);
      row_sum[i] += w;
    }
    md.A.start[j + 1] = static_cast<int64_t>(md.A.index.size());
  }
  md.obj.resize(static_cast<size_t>(n));
  for (double& p : md.obj) p = static_cast<double>(1 + rng.below(100));  // profits U{1..100}
  md.row_lower.assign(static_cast<size_t>(m), -kInf);
  md.row_upper.resize(static_cast<size_t>(m));
  for (int i = 0; i < m; ++i) md.row_upper[i] = 0.4 * row_sum[i];
  md.col_lower.assign(static_cast<size_t>(n), 0.0);
  md.col_upper.assign(static_cast<size_t>(n), 1.0);
  return md;
}

void generate_fixings(uint64_t seed, int n, double frac_min, double frac_max, const double* lower,
                      const double* upper, double* out_lower, double* out_upper) {
  if (n < 0 || frac_min < 0.0 || frac_max < frac_min || frac_max > 1.0)
    throw std::invalid_argument("bad fixing fractions");
  Rng rng(seed);
  std::copy(lower, lower + n, out_lower);
  std::copy(upper, upper + n, out_upper);
  const double f = rng.uniform(frac_min, frac_max);
  const int count = std::min(n, static_cast<int>(std::lround(f * n)));
  std::vector<int> cols;
  distinct_sorted(rng, n, count, cols);
  for (int j : cols) {
    const double v = rng.below(2) == 0 ? lower[j] : upper[j];
    out_lower[j] = out_upper[j] = v;
  }
}

}  // namespace smy::host
