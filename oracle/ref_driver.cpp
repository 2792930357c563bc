// Driver around the reference's own PDLP sources, compiled where they lie under the reference
// tree by oracle/Makefile into oracle/_ref/libsamaya_ref.so. Test infrastructure only: it is
// used to pin the oracle (pdlp_oracle.cpp) and to write golden fixtures under tests/golden/.
//
// The functions below only marshal plain arrays into the reference's types and call its code:
// samaya::compute_scaling (src/linalg/scaling.cpp:130), samaya::make_problem
// (src/lp/lp_solver.cpp:28), samaya::solve_pdlp (src/lp/pdlp.cpp:228), the verifier
// (src/verify/verify.cpp:251,269,345), the MPS reader read_mps_from_string
// (src/io/mps_reader.cpp:516-527), and the test generators random_lp
// (tests/lp_generators.hpp:33) and ReferenceLp (tests/reference_lp.hpp:20). The PDLP result
// mapping mirrors solve_pdlp_lp (src/core/solver.cpp:405-437), which is not compiled here
// because solver.cpp pulls in every other method of the library.

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "core/log.hpp"
#include "core/xreal.hpp"
#include "linalg/scaling.hpp"
#include "lp/lp_solver.hpp"
#include "lp/pdlp.hpp"
#include "lp_generators.hpp"
#include "reference_lp.hpp"
#include "samaya/io.hpp"
#include "samaya/model.hpp"
#include "samaya/verify.hpp"

extern "C" {

struct ref_model {
  int m, n;
  int maximize;
  double offset;
  const int64_t* start;
  const int* index;
  const double* value;
  const double* obj;
  const double* col_lower;
  const double* col_upper;
  const double* row_lower;
  const double* row_upper;
};

// Caller-owned buffers a generated model is written into (capacities in the *_cap fields).
struct ref_model_buf {
  int m, n, maximize;
  double offset;
  long long nnz;
  int64_t* start;  // n + 1
  int* index;      // nnz
  double* value;   // nnz
  double* obj;
  double* col_lower;
  double* col_upper;
  double* row_lower;
  double* row_upper;
  int rows_cap, cols_cap;
  long long nnz_cap;
};

struct ref_lp_out {
  int status;  // SimplexStatus (src/lp/simplex.hpp:28-35)
  long long iterations;
  double objective;
  int verified;
  double max_primal_violation, max_dual_violation;
  double primal_residual, dual_residual, gap;
  double* col_value;  // n
  double* row_dual;   // m
  double* col_dual;   // n
  double* ray;        // m or n, may be null
  double* pdlp_x;     // n + m, scaled space, may be null
  double* pdlp_y;     // m, scaled space, may be null
  char message[256];
};

}  // extern "C"

namespace {

samaya::Model to_model(const ref_model* p) {
  samaya::Model md;
  md.sense = p->maximize ? samaya::ObjSense::kMaximize : samaya::ObjSense::kMinimize;
  md.obj_offset = p->offset;
  md.obj.assign(p->obj, p->obj + p->n);
  md.col_lower.assign(p->col_lower, p->col_lower + p->n);
  md.col_upper.assign(p->col_upper, p->col_upper + p->n);
  md.col_type.assign(static_cast<size_t>(p->n), samaya::VarType::kContinuous);
  md.row_lower.assign(p->row_lower, p->row_lower + p->m);
  md.row_upper.assign(p->row_upper, p->row_upper + p->m);
  std::vector<samaya::Triplet> t;
  for (int j = 0; j < p->n; ++j)
    for (int64_t q = p->start[j]; q < p->start[j + 1]; ++q) t.push_back({p->index[q], j, p->value[q]});
  md.A = samaya::SparseMatrix::from_triplets(p->m, p->n, std::move(t));
  return md;
}

void put(const std::vector<double>& v, double* dst) {
  if (dst != nullptr && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(double));
}

}  // namespace

extern "C" {

void* ref_rng_new(unsigned seed) { return new std::mt19937(seed); }
void ref_rng_free(void* rng) { delete static_cast<std::mt19937*>(rng); }

// std::uniform_real_distribution<double>(a, b) drawn from the caller's generator, as the
// reference's tests draw their random equilibration scales (tests/test_gpu.cpp:522-526).
double ref_uniform(void* rng, double a, double b) {
  return std::uniform_real_distribution<double>(a, b)(*static_cast<std::mt19937*>(rng));
}

// One draw of tests/lp_generators.hpp:33 random_lp from the caller's generator state. Returns 0,
// or -1 when a buffer is too small.
int ref_random_lp(void* rng, int family, int max_rows, int max_cols, ref_model_buf* out) {
  const samaya::Model md = samaya::test::random_lp(static_cast<samaya::test::LpFamily>(family),
                                                   max_rows, max_cols,
                                                   *static_cast<std::mt19937*>(rng));
  const int m = md.num_rows(), n = md.num_cols();
  if (m > out->rows_cap || n > out->cols_cap || md.A.nnz() > out->nnz_cap) return -1;
  out->m = m;
  out->n = n;
  out->maximize = md.sense == samaya::ObjSense::kMaximize ? 1 : 0;
  out->offset = md.obj_offset;
  out->nnz = md.A.nnz();
  for (int j = 0; j <= n; ++j) out->start[j] = md.A.col_start()[j];
  for (int64_t q = 0; q < md.A.nnz(); ++q) {
    out->index[q] = md.A.row_index()[q];
    out->value[q] = md.A.values()[q];
  }
  put(md.obj, out->obj);
  put(md.col_lower, out->col_lower);
  put(md.col_upper, out->col_upper);
  put(md.row_lower, out->row_lower);
  put(md.row_upper, out->row_upper);
  return 0;
}

// The dense long-double tableau simplex of the reference's tests (tests/reference_lp.hpp).
// Returns 0 optimal, 1 infeasible, 2 unbounded.
int ref_reference_lp(const ref_model* model, double* objective) {
  const samaya::Model md = to_model(model);
  const samaya::test::ReferenceResult r = samaya::test::ReferenceLp(md).solve();
  *objective = r.objective;
  return static_cast<int>(r.status);
}

// The LP-by-PDLP path (solver.cpp:392-437): scaling, make_problem, solve_pdlp (CPU), mapping, and
// the verifier at default tolerances. `log_level` goes to the reference's Logger (stdout).
int ref_solve_lp(const ref_model* model, double tol, long long max_iterations, int log_level,
                 ref_lp_out* out) {
  const samaya::Model md = to_model(model);
  const samaya::Scaling scaled = samaya::compute_scaling(md.A);
  const samaya::LpProblem lp = samaya::make_problem(md, scaled);
  samaya::PdlpOptions opt;
  opt.tol = tol;
  opt.max_iterations = max_iterations;
  const samaya::Logger log(log_level);
  const samaya::PdlpResult r = samaya::solve_pdlp(lp, opt, log);
  const int n = md.num_cols(), m = md.num_rows();
  const double sense = md.sense == samaya::ObjSense::kMaximize ? -1.0 : 1.0;
  out->status = static_cast<int>(r.status);
  out->iterations = r.iterations;
  out->primal_residual = r.primal_residual;
  out->dual_residual = r.dual_residual;
  out->gap = r.gap;
  put(r.x, out->pdlp_x);
  put(r.y, out->pdlp_y);
  std::vector<double> x(r.x.begin(), r.x.begin() + n), y = r.y, d(static_cast<size_t>(n), 0.0);
  scaled.unscale_cols(x);
  scaled.unscale_row_duals(y);
  for (double& v : y) v *= -sense;
  md.A.multiply_transpose(y, d);
  samaya::Xreal objective = md.obj_offset;
  for (int j = 0; j < n; ++j) {
    d[j] = md.obj[j] - d[j];
    objective += samaya::Xreal(md.obj[j]) * x[j];
  }
  out->objective = static_cast<double>(objective);
  put(x, out->col_value);
  put(y, out->row_dual);
  put(d, out->col_dual);
  samaya::VerifyReport rep;
  bool answered = true;
  if (r.status == samaya::SimplexStatus::kOptimal) {
    rep = samaya::verify_lp_optimality(md, x, y);
  } else if (r.status == samaya::SimplexStatus::kInfeasible && !r.dual_ray.empty()) {
    std::vector<double> ray = r.dual_ray;
    scaled.unscale_row_duals(ray);
    for (double& v : ray) v *= sense;
    put(ray, out->ray);
    rep = samaya::verify_infeasibility(md, ray);
  } else if (r.status == samaya::SimplexStatus::kUnbounded && !r.primal_ray.empty()) {
    std::vector<double> ray = r.primal_ray;
    scaled.unscale_cols(ray);
    put(ray, out->ray);
    rep = samaya::verify_unbounded_ray(md, ray);
  } else {
    answered = false;
  }
  out->verified = answered && rep.ok ? 1 : 0;
  out->max_primal_violation = std::max(rep.max_bound_violation, rep.max_row_violation);
  out->max_dual_violation = rep.max_dual_violation;
  std::snprintf(out->message, sizeof out->message, "%s", rep.message.c_str());
  (void)m;
  return 0;
}

// The reference verifier on model-space x and y (verify.cpp:251). out3: primal violation, dual
// violation, objective. Returns 1 when the point verifies.
int ref_verify_lp(const ref_model* model, const double* x, const double* y, double* out3) {
  const samaya::Model md = to_model(model);
  const samaya::VerifyReport r = samaya::verify_lp_optimality(
      md, std::span<const double>(x, static_cast<size_t>(md.num_cols())),
      std::span<const double>(y, static_cast<size_t>(md.num_rows())));
  out3[0] = std::max(r.max_bound_violation, r.max_row_violation);
  out3[1] = r.max_dual_violation;
  out3[2] = r.objective;
  return r.ok ? 1 : 0;
}

// compute_scaling + make_problem (scaling.cpp:130, lp_solver.cpp:28). a_value receives the
// scaled A in CSC order.
int ref_make_problem(const ref_model* model, double* col_scale, double* row_scale, double* cost,
                     double* lower, double* upper, double* a_value) {
  const samaya::Model md = to_model(model);
  const samaya::Scaling s = samaya::compute_scaling(md.A);
  const samaya::LpProblem lp = samaya::make_problem(md, s);
  put(s.col, col_scale);
  put(s.row, row_scale);
  put(lp.cost, cost);
  put(lp.lower, lower);
  put(lp.upper, upper);
  for (int64_t q = 0; q < lp.A.nnz(); ++q) a_value[q] = lp.A.values()[q];
  return static_cast<int>(lp.A.nnz());
}

// read_mps_from_string on `len` bytes of text: a handle to the parsed model, or null with the
// ParseError's message in err and its line in *line. *seconds: the parse's wall time.
void* ref_mps_read(const char* text, long long len, char* err, int err_cap, long long* line,
                   double* seconds) {
  *line = 0;
  if (err_cap > 0) err[0] = '\0';
  const auto t0 = std::chrono::steady_clock::now();
  try {
    auto* md = new samaya::Model(samaya::read_mps_from_string(std::string_view(text, static_cast<size_t>(len))));
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return md;
  } catch (const samaya::ParseError& e) {
    std::snprintf(err, static_cast<size_t>(err_cap), "%s", e.what());
    *line = e.line();
  } catch (const std::exception& e) {
    std::snprintf(err, static_cast<size_t>(err_cap), "%s", e.what());
  }
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return nullptr;
}

// out: m, n, nnz, maximize, then the bytes of the row names, the column names (each followed by
// '\n') and the model name.
void ref_mps_dims(const void* h, long long* out) {
  const auto& md = *static_cast<const samaya::Model*>(h);
  long long rb = 0, cb = 0;
  for (const std::string& s : md.row_names) rb += static_cast<long long>(s.size()) + 1;
  for (const std::string& s : md.col_names) cb += static_cast<long long>(s.size()) + 1;
  out[0] = md.num_rows();
  out[1] = md.num_cols();
  out[2] = md.A.nnz();
  out[3] = md.sense == samaya::ObjSense::kMaximize ? 1 : 0;
  out[4] = rb;
  out[5] = cb;
  out[6] = static_cast<long long>(md.name.size());
}

// Copies a parsed model out: arrays into buf (sized by ref_mps_dims), col_type (n bytes, 1 =
// integer), the names joined with '\n' terminators, the model name (not terminated). Returns the
// number of quadratic entries (0 for an LP).
long long ref_mps_copy(const void* h, ref_model_buf* buf, char* col_type, char* row_names,
                       char* col_names, char* name) {
  const auto& md = *static_cast<const samaya::Model*>(h);
  const int m = md.num_rows(), n = md.num_cols();
  buf->m = m;
  buf->n = n;
  buf->maximize = md.sense == samaya::ObjSense::kMaximize ? 1 : 0;
  buf->offset = md.obj_offset;
  buf->nnz = md.A.nnz();
  const auto start = md.A.col_start();
  const auto index = md.A.row_index();
  for (int j = 0; j <= n; ++j) buf->start[j] = start[j];
  for (long long q = 0; q < buf->nnz; ++q) {
    buf->index[q] = index[q];
    buf->value[q] = md.A.values()[q];
  }
  put(md.obj, buf->obj);
  put(md.col_lower, buf->col_lower);
  put(md.col_upper, buf->col_upper);
  put(md.row_lower, buf->row_lower);
  put(md.row_upper, buf->row_upper);
  for (int j = 0; j < n; ++j) col_type[j] = md.col_type[j] == samaya::VarType::kInteger ? 1 : 0;
  for (const std::string& s : md.row_names) {
    std::memcpy(row_names, s.data(), s.size());
    row_names += s.size();
    *row_names++ = '\n';
  }
  for (const std::string& s : md.col_names) {
    std::memcpy(col_names, s.data(), s.size());
    col_names += s.size();
    *col_names++ = '\n';
  }
  std::memcpy(name, md.name.data(), md.name.size());
  return md.Q.nnz() + static_cast<long long>(md.row_quadratic.size());
}

void ref_mps_free(void* h) { delete static_cast<samaya::Model*>(h); }

}  // extern "C"
