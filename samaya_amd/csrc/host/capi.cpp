// C ABI of include/samaya_pdlp.h over the C++ host solver.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <stdexcept>

#include "../../../include/samaya_pdhg.h"
#include "../../../include/samaya_mip.h"
#include "../../../include/samaya_pdlp.h"
#include "capi_model.hpp"
#include "generators.hpp"
#include "mip.hpp"
#include "pdlp.hpp"

using namespace smy::host;

namespace {

void copy_text(char* dst, size_t len, const std::string& s) {
  if (dst == nullptr || len == 0) return;
  const size_t k = std::min(s.size(), len - 1);
  std::memcpy(dst, s.data(), k);
  dst[k] = '\0';
}

template <typename T>
void put(const std::vector<T>& v, T* dst) {
  if (dst != nullptr && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
}

PdlpOptions to_options(const smy_pdlp_params* p) {
  smy_pdlp_params d;
  smy_pdlp_params_default(&d);
  if (p == nullptr) p = &d;
  PdlpOptions o;
  o.tol = p->tol;
  o.max_iterations = p->max_iterations;
  o.time_limit = p->time_limit;
  o.use_gpu = p->use_gpu != 0;
  o.precision = p->precision == SMY_PRECISION_MIXED ? Precision::kMixed : Precision::kFp64;
  o.fp32_switch_tol = p->fp32_switch_tol;
  o.use_graph = p->use_graph != 0;
  o.detect = p->detect != 0;
  o.log_level = p->log_level;
  o.algorithm = p->algorithm == 1 ? Algorithm::kEnhanced : Algorithm::kReference;
  o.share_matrix = p->share_matrix != 0;
  o.device_detect = p->device_detect != 0;
  o.start_omega = p->start_omega;
  o.adaptive_steps = p->adaptive_steps != 0;
  o.fp64_accumulation = p->fp64_accumulation != 0;
  o.inject_fault_step = p->inject_fault_step;
  o.device_scaling = p->device_scaling != 0;
  o.column_order = static_cast<ColumnOrder>(p->column_order >= 1 && p->column_order <= 3 ? p->column_order : 0);
  o.fp32_stall_level = p->fp32_stall_level;
  o.device_restart = p->device_restart != 0;
  o.halpern = p->halpern != 0;
  o.reflection = p->reflection;
  o.pid_primal_weight = p->pid_primal_weight != 0;
  o.farkas_projection = p->farkas_projection < 0 ? o.algorithm != Algorithm::kReference
                                                 : p->farkas_projection != 0;
  return o;
}

void fill_info(const PdlpResult& r, smy_pdlp_info* info) {
  if (info == nullptr) return;
  info->status = static_cast<int>(r.status);
  info->iterations = r.iterations;
  info->fp32_iterations = r.fp32_iterations;
  info->primal_residual = r.primal_residual;
  info->dual_residual = r.dual_residual;
  info->gap = r.gap;
  info->operator_norm = r.operator_norm;
  info->omega = r.omega;
  info->rejected_steps = r.rejected_steps;
  info->setup_seconds = r.setup_seconds;
  info->solve_seconds = r.solve_seconds;
  info->kernel_ms = r.kernel_ms;
  info->restarts = static_cast<int>(r.restarts.size());
  info->evaluations = static_cast<int>(r.evals.size());
  copy_text(info->message, sizeof info->message, r.message);
}

bool valid_csc(int rows, int cols, const int64_t* start, const int32_t* index) {
  if (rows < 0 || cols < 0 || start == nullptr) return false;
  if (start[0] != 0) return false;
  for (int j = 0; j < cols; ++j) {
    if (start[j + 1] < start[j]) return false;
    for (int64_t p = start[j]; p < start[j + 1]; ++p) {
      if (index[p] < 0 || index[p] >= rows) return false;
      if (p > start[j] && index[p] <= index[p - 1]) return false;  // sorted, unique
    }
  }
  return true;
}

LpSolveOptions to_lp_options(const smy_pdlp_params* params) {
  smy_pdlp_params d;
  smy_pdlp_params_default(&d);
  if (params == nullptr) params = &d;
  LpSolveOptions o;
  o.pdlp = to_options(params);
  o.verify = params->verify != 0;
  o.polish_rounds = params->polish_rounds;
  return o;
}

void fill_result(const LpSolveResult& r, smy_lp_result* result) {
  result->status = static_cast<int>(r.status);
  result->verified = r.verified ? 1 : 0;
  result->objective = r.objective;
  result->max_primal_violation = r.max_primal_violation;
  result->max_dual_violation = r.max_dual_violation;
  result->iterations = r.iterations;
  result->fp32_iterations = r.fp32_iterations;
  result->rounds = r.rounds;
  fill_info(r.pdlp, &result->pdlp);
  copy_text(result->message, sizeof result->message, r.message);
  result->dual_bound = r.dual_bound;
}

}  // namespace

extern "C" {

void smy_pdlp_params_default(smy_pdlp_params* p) {
  if (p == nullptr) return;
  p->tol = 1e-8;
  p->max_iterations = -1;
  p->time_limit = std::numeric_limits<double>::infinity();
  p->use_gpu = 1;
  p->precision = SMY_PRECISION_FP64;
  p->fp32_switch_tol = 1e-5;
  p->use_graph = 1;
  p->detect = 1;
  p->log_level = 0;
  p->verify = 1;
  p->polish_rounds = 0;
  p->algorithm = SMY_ALGORITHM_REFERENCE;
  p->share_matrix = 1;
  p->device_detect = 1;
  p->start_omega = 0.0;
  p->adaptive_steps = 0;
  p->fp64_accumulation = 0;
  p->inject_fault_step = -1;
  p->device_scaling = 1;
  p->column_order = 0;
  p->fp32_stall_level = 1e-4;
  p->device_restart = 1;
  p->halpern = 0;
  p->reflection = 1.0;
  p->pid_primal_weight = 0;
  p->farkas_projection = -1;
}

int smy_model_create(int m, int n, int maximize, double offset, const int64_t* start,
                     const int32_t* index, const double* value, const double* obj,
                     const double* col_lower, const double* col_upper, const double* row_lower,
                     const double* row_upper, smy_model** out) {
  if (out == nullptr) return SMY_INVALID;
  *out = nullptr;
  if (!valid_csc(m, n, start, index) || obj == nullptr || col_lower == nullptr ||
      col_upper == nullptr || (m > 0 && (row_lower == nullptr || row_upper == nullptr)))
    return SMY_INVALID;
  try {
    auto md = std::make_unique<smy_model>();
    Model& M = md->model;
    M.m = m;
    M.n = n;
    M.maximize = maximize != 0;
    M.offset = offset;
    M.A.rows = m;
    M.A.cols = n;
    M.A.start.assign(start, start + n + 1);
    M.A.index.assign(index, index + start[n]);
    M.A.value.assign(value, value + start[n]);
    M.obj.assign(obj, obj + n);
    M.col_lower.assign(col_lower, col_lower + n);
    M.col_upper.assign(col_upper, col_upper + n);
    M.row_lower.assign(row_lower, row_lower + m);
    M.row_upper.assign(row_upper, row_upper + m);
    *out = md.release();
    return SMY_OK;
  } catch (...) {
    return SMY_ERROR;
  }
}

int smy_model_generate(int kind, uint64_t seed, long long p0, long long p1, long long p2,
                       double p3, smy_model** out) {
  if (out == nullptr) return SMY_INVALID;
  *out = nullptr;
  try {
    auto md = std::make_unique<smy_model>();
    switch (kind) {
      case SMY_GEN_SPARSE_UNIFORM:
        md->model = generate_sparse_uniform(seed, static_cast<int>(p0), static_cast<int>(p1),
                                            static_cast<int>(p2));
        break;
      case SMY_GEN_SPARSE_POWERLAW:
        md->model = generate_sparse_powerlaw(seed, static_cast<int>(p0), static_cast<int>(p1),
                                             static_cast<int>(p2), p3);
        break;
      case SMY_GEN_KNAPSACK:
        md->model = generate_knapsack(seed, static_cast<int>(p0), static_cast<int>(p1),
                                      static_cast<int>(p2));
        break;
      case SMY_GEN_BLENDING:
        md->model = generate_blending(seed, static_cast<int>(p0), static_cast<int>(p1),
                                      static_cast<int>(p2), static_cast<int>(p3));
        break;
      default:
        return SMY_INVALID;
    }
    *out = md.release();
    return SMY_OK;
  } catch (const std::invalid_argument&) {
    return SMY_INVALID;
  } catch (...) {
    return SMY_ERROR;
  }
}

void smy_model_free(smy_model* model) { delete model; }
int smy_model_num_rows(const smy_model* model) { return model ? model->model.m : 0; }
int smy_model_num_cols(const smy_model* model) { return model ? model->model.n : 0; }
long long smy_model_num_nonzeros(const smy_model* model) {
  return model ? static_cast<long long>(model->model.A.nnz()) : 0;
}

int smy_model_export(const smy_model* model, int64_t* start, int32_t* index, double* value,
                     double* obj, double* col_lower, double* col_upper, double* row_lower,
                     double* row_upper, int* maximize, double* offset) {
  if (model == nullptr) return SMY_INVALID;
  const Model& M = model->model;
  put(M.A.start, start);
  put(M.A.index, index);
  put(M.A.value, value);
  put(M.obj, obj);
  put(M.col_lower, col_lower);
  put(M.col_upper, col_upper);
  put(M.row_lower, row_lower);
  put(M.row_upper, row_upper);
  if (maximize) *maximize = M.maximize ? 1 : 0;
  if (offset) *offset = M.offset;
  return SMY_OK;
}

int smy_compute_scaling(const smy_model* model, double* col_scale, double* row_scale) {
  if (model == nullptr) return SMY_INVALID;
  try {
    const Scaling s = compute_scaling(model->model.A);
    put(s.col, col_scale);
    put(s.row, row_scale);
    return SMY_OK;
  } catch (...) {
    return SMY_ERROR;
  }
}

int smy_lp_dual_bound(const smy_model* model, const double* row_dual, double* bound) {
  if (model == nullptr || bound == nullptr || (model->model.m > 0 && row_dual == nullptr)) return SMY_INVALID;
  try {
    *bound = safe_dual_bound(model->model, row_dual);
    return SMY_OK;
  } catch (...) {
    return SMY_ERROR;
  }
}

int smy_pdlp_equilibrate(const smy_csc* A, const smy_csc* At, int algorithm, int passes,
                         int pock_chambolle, double* col_scale, double* row_scale, double out[3]) {
  if (A == nullptr || At == nullptr || out == nullptr || passes < 0) return SMY_INVALID;
  try {
    const auto own = [](const smy_csc* c) {
      Csc m;
      m.rows = c->rows;
      m.cols = c->cols;
      m.start.assign(c->start, c->start + c->cols + 1);
      m.index.assign(c->index, c->index + c->start[c->cols]);
      m.value.assign(c->value, c->value + c->start[c->cols]);
      return m;
    };
    const Csc a = own(A), at = own(At);
    std::vector<double> rs, cs;
    equilibrate_host(a, at, algorithm == SMY_ALGORITHM_ENHANCED ? Algorithm::kEnhanced : Algorithm::kReference,
                     passes, pock_chambolle != 0, rs, cs, out);
    put(cs, col_scale);
    put(rs, row_scale);
    return SMY_OK;
  } catch (...) {
    return SMY_ERROR;
  }
}

int smy_make_problem(const smy_model* model, double* cost, double* lower, double* upper,
                     double* a_value) {
  if (model == nullptr) return SMY_INVALID;
  try {
    const LpProblem lp = make_problem(model->model, compute_scaling(model->model.A));
    if (lp.A.nnz() != model->model.A.nnz()) return SMY_ERROR;  // a scaled entry underflowed
    put(lp.cost, cost);
    put(lp.lower, lower);
    put(lp.upper, upper);
    put(lp.A.value, a_value);
    return SMY_OK;
  } catch (...) {
    return SMY_ERROR;
  }
}

int smy_solve_pdlp(int m, int n, const int64_t* a_start, const int32_t* a_index,
                   const double* a_value, const double* cost, const double* lower,
                   const double* upper, const smy_pdlp_params* params, const double* start_x,
                   const double* start_y, smy_pdlp_info* info, double* x, double* y,
                   double* dual_ray, double* primal_ray, long long* restart_iter,
                   int* restart_to_average, double* restart_kkt, double* restart_omega,
                   int restart_cap, long long* eval_iter, double* eval_measures, int eval_cap) {
  if (!valid_csc(m, n, a_start, a_index) || cost == nullptr || lower == nullptr || upper == nullptr)
    return SMY_INVALID;
  try {
    LpProblem lp;
    lp.m = m;
    lp.n = n;
    lp.A.rows = m;
    lp.A.cols = n;
    lp.A.start.assign(a_start, a_start + n + 1);
    lp.A.index.assign(a_index, a_index + a_start[n]);
    lp.A.value.assign(a_value, a_value + a_start[n]);
    lp.At = csc_transpose(lp.A);
    const size_t nt = static_cast<size_t>(n + m);
    lp.cost.assign(cost, cost + nt);
    lp.lower.assign(lower, lower + nt);
    lp.upper.assign(upper, upper + nt);
    PdlpOptions o = to_options(params);
    if (start_x != nullptr) o.start_x.assign(start_x, start_x + nt);
    if (start_y != nullptr) o.start_y.assign(start_y, start_y + m);
    const PdlpResult r = solve_pdlp(lp, o);
    fill_info(r, info);
    put(r.x, x);
    put(r.y, y);
    put(r.dual_ray, dual_ray);
    put(r.primal_ray, primal_ray);
    for (int k = 0; k < static_cast<int>(r.restarts.size()) && k < restart_cap; ++k) {
      if (restart_iter) restart_iter[k] = r.restarts[k].iter;
      if (restart_to_average) restart_to_average[k] = r.restarts[k].to_average;
      if (restart_kkt) restart_kkt[k] = r.restarts[k].kkt;
      if (restart_omega) restart_omega[k] = r.restarts[k].omega;
    }
    for (int k = 0; k < static_cast<int>(r.evals.size()) && k < eval_cap; ++k) {
      if (eval_iter) eval_iter[k] = r.evals[k].iter;
      if (eval_measures)
        for (int q = 0; q < 3; ++q) {
          eval_measures[6 * k + q] = r.evals[k].current[q];
          eval_measures[6 * k + 3 + q] = r.evals[k].average[q];
        }
    }
    return r.status == Status::kNumericalError && !r.message.empty() ? SMY_ERROR : SMY_OK;
  } catch (const std::exception& e) {
    if (info) {
      info->status = static_cast<int>(Status::kNumericalError);
      copy_text(info->message, sizeof info->message, e.what());
    }
    return SMY_ERROR;
  }
}

int smy_solve_lp(const smy_model* model, const smy_pdlp_params* params, smy_lp_result* result,
                 double* col_value, double* row_dual, double* col_dual, double* row_activity,
                 double* certificate) {
  if (model == nullptr || result == nullptr) return SMY_INVALID;
  std::memset(result, 0, sizeof *result);
  try {
    const LpSolveResult r = solve_lp(model->model, to_lp_options(params));
    fill_result(r, result);
    put(r.col_value, col_value);
    put(r.row_dual, row_dual);
    put(r.col_dual, col_dual);
    put(r.row_activity, row_activity);
    if (!r.infeasibility_certificate.empty()) put(r.infeasibility_certificate, certificate);
    if (!r.unbounded_ray.empty()) put(r.unbounded_ray, certificate);
    return SMY_OK;
  } catch (const std::exception& e) {
    result->status = static_cast<int>(Status::kNumericalError);
    copy_text(result->message, sizeof result->message, e.what());
    return SMY_ERROR;
  }
}

int smy_solve_lp_batch(const smy_model* const* models, int count, const smy_pdlp_params* params,
                       smy_lp_result* results, double* const* col_values,
                       double* const* row_duals, double* times) {
  if (models == nullptr || results == nullptr || count <= 0) return SMY_INVALID;
  for (int i = 0; i < count; ++i)
    if (models[i] == nullptr) return SMY_INVALID;
  std::memset(results, 0, sizeof(smy_lp_result) * static_cast<size_t>(count));
  try {
    std::vector<const Model*> ms;
    for (int i = 0; i < count; ++i) ms.push_back(&models[i]->model);
    PdlpBatchInfo info;
    const std::vector<LpSolveResult> rs = solve_lp_batch(ms, to_lp_options(params), &info);
    for (int i = 0; i < count; ++i) {
      fill_result(rs[i], &results[i]);
      if (col_values != nullptr) put(rs[i].col_value, col_values[i]);
      if (row_duals != nullptr) put(rs[i].row_dual, row_duals[i]);
    }
    if (times != nullptr) {
      times[0] = info.setup_seconds;
      times[1] = info.solve_seconds;
      times[2] = info.kernel_ms;
    }
    return SMY_OK;
  } catch (const std::exception& e) {
    for (int i = 0; i < count; ++i) {
      results[i].status = static_cast<int>(Status::kNumericalError);
      copy_text(results[i].message, sizeof results[i].message, e.what());
    }
    return SMY_ERROR;
  }
}

int smy_solve_lp_bounds_batch(const smy_model* model, int count, const double* const* col_lower,
                              const double* const* col_upper, const smy_pdlp_params* params,
                              const double* start_col_value, const double* start_row_dual,
                              smy_lp_result* results, double* const* col_values,
                              double* const* row_duals, double* times) {
  if (model == nullptr || results == nullptr || count <= 0 || col_lower == nullptr ||
      col_upper == nullptr)
    return SMY_INVALID;
  for (int i = 0; i < count; ++i)
    if (model->model.n > 0 && (col_lower[i] == nullptr || col_upper[i] == nullptr)) return SMY_INVALID;
  std::memset(results, 0, sizeof(smy_lp_result) * static_cast<size_t>(count));
  try {
    std::vector<const double*> lo(col_lower, col_lower + count), up(col_upper, col_upper + count);
    PdlpBatchInfo info;
    const std::vector<LpSolveResult> rs = solve_lp_bounds_batch(
        model->model, lo, up, to_lp_options(params), start_col_value, start_row_dual, &info);
    for (int i = 0; i < count; ++i) {
      fill_result(rs[i], &results[i]);
      if (col_values != nullptr) put(rs[i].col_value, col_values[i]);
      if (row_duals != nullptr) put(rs[i].row_dual, row_duals[i]);
    }
    if (times != nullptr) {
      times[0] = info.setup_seconds;
      times[1] = info.solve_seconds;
      times[2] = info.kernel_ms;
    }
    return SMY_OK;
  } catch (const std::exception& e) {
    for (int i = 0; i < count; ++i) {
      results[i].status = static_cast<int>(Status::kNumericalError);
      copy_text(results[i].message, sizeof results[i].message, e.what());
    }
    return SMY_ERROR;
  }
}

void smy_mip_params_default(smy_mip_params* params) {
  if (params == nullptr) return;
  smy_pdlp_params_default(&params->lp);
  const MipOptions d;
  params->gap_tol = d.gap_tol;
  params->wave = d.wave;
  params->max_nodes = d.max_nodes;
  params->time_limit = d.time_limit;
  params->integrality_tol = d.integrality_tol;
  params->feasibility_tol = d.feasibility_tol;
  params->log_level = d.log_level;
}

int smy_mip_solve(const smy_model* model, const unsigned char* integer, const smy_mip_params* params,
                  smy_mip_result* result, double* col_value) {
  if (model == nullptr || result == nullptr) return SMY_INVALID;
  std::memset(result, 0, sizeof *result);
  smy_mip_params d;
  smy_mip_params_default(&d);
  if (params == nullptr) params = &d;
  const int n = model->model.n;
  std::vector<char> flags(static_cast<size_t>(n), 0);
  if (integer != nullptr) {
    for (int j = 0; j < n; ++j) flags[j] = integer[j] != 0;
  } else if (static_cast<int>(model->col_integer.size()) == n) {
    flags.assign(model->col_integer.begin(), model->col_integer.end());
  }
  if (std::none_of(flags.begin(), flags.end(), [](char f) { return f != 0; })) {
    copy_text(result->message, sizeof result->message, "no integer columns");
    return SMY_INVALID;
  }
  if (!(params->gap_tol >= 0.0) || params->wave <= 0 || params->max_nodes <= 0) return SMY_INVALID;
  try {
    MipOptions o;
    o.lp = to_lp_options(&params->lp);
    o.gap_tol = params->gap_tol;
    o.wave = params->wave;
    o.max_nodes = params->max_nodes;
    o.time_limit = params->time_limit;
    o.integrality_tol = params->integrality_tol;
    o.feasibility_tol = params->feasibility_tol;
    o.log_level = params->log_level;
    const MipResult r = solve_mip(model->model, flags, o);
    result->status = static_cast<int>(r.status);
    result->objective = r.objective;
    result->bound = r.bound;
    result->gap = r.gap;
    result->nodes = r.nodes;
    result->waves = r.waves;
    result->pruned_infeasible = r.pruned_infeasible;
    result->pruned_bound = r.pruned_bound;
    result->integral = r.integral;
    result->unresolved = r.unresolved;
    result->inexact = r.inexact;
    result->fixed_solves = r.fixed_solves;
    result->lp_iterations = r.lp_iterations;
    result->seconds = r.seconds;
    if (col_value != nullptr && !r.col_value.empty()) put(r.col_value, col_value);
    return SMY_OK;
  } catch (const std::exception& e) {
    result->status = static_cast<int>(Status::kNumericalError);
    copy_text(result->message, sizeof result->message, e.what());
    return SMY_ERROR;
  }
}

int smy_mip_propagate(const smy_model* model, const unsigned char* integer, double* col_lower,
                      double* col_upper, double* row_lower, double* row_upper, int* crossed) {
  if (model == nullptr || crossed == nullptr) return SMY_INVALID;
  const int n = model->model.n;
  std::vector<char> flags(static_cast<size_t>(n), 0);
  if (integer != nullptr) {
    for (int j = 0; j < n; ++j) flags[j] = integer[j] != 0;
  } else if (static_cast<int>(model->col_integer.size()) == n) {
    flags.assign(model->col_integer.begin(), model->col_integer.end());
  }
  try {
    Model md = model->model;
    *crossed = propagate_bounds(md, flags) ? 0 : 1;
    if (col_lower != nullptr) put(md.col_lower, col_lower);
    if (col_upper != nullptr) put(md.col_upper, col_upper);
    if (row_lower != nullptr) put(md.row_lower, row_lower);
    if (row_upper != nullptr) put(md.row_upper, row_upper);
    return SMY_OK;
  } catch (const std::exception&) {
    return SMY_ERROR;
  }
}

int smy_generate_fixings(uint64_t seed, int n, double frac_min, double frac_max,
                         const double* lower, const double* upper, double* out_lower,
                         double* out_upper) {
  if (n < 0 || (n > 0 && (lower == nullptr || upper == nullptr || out_lower == nullptr ||
                          out_upper == nullptr)))
    return SMY_INVALID;
  try {
    generate_fixings(seed, n, frac_min, frac_max, lower, upper, out_lower, out_upper);
    return SMY_OK;
  } catch (const std::invalid_argument&) {
    return SMY_INVALID;
  }
}

int smy_solve(const smy_model* model, double* objective) {
  // samaya::Status values (include/samaya_c.h:14-28).
  if (objective != nullptr) *objective = std::numeric_limits<double>::quiet_NaN();
  if (model == nullptr) return 9;  // SAMAYA_INVALID_MODEL
  smy_lp_result r;
  if (smy_solve_lp(model, nullptr, &r, nullptr, nullptr, nullptr, nullptr, nullptr) != SMY_OK)
    return 8;
  if (r.status == SMY_STATUS_OPTIMAL && !r.verified) return 8;
  switch (r.status) {
    case SMY_STATUS_OPTIMAL:
      if (objective != nullptr) *objective = r.objective;
      return 1;
    case SMY_STATUS_INFEASIBLE: return 2;
    case SMY_STATUS_UNBOUNDED: return 3;
    case SMY_STATUS_TIME_LIMIT: return 5;
    case SMY_STATUS_ITERATION_LIMIT: return 6;
    default: return 8;
  }
}

int smy_verify_lp(const smy_model* model, const double* x, const double* y, double primal_tol,
                  double dual_tol, double* out3, char* message, size_t message_len) {
  if (model == nullptr || x == nullptr || (y == nullptr && model->model.m > 0)) return -1;
  const VerifyReport r = verify_lp_optimality(model->model, x, y, primal_tol, dual_tol);
  if (out3 != nullptr) {
    out3[0] = std::max(r.max_bound_violation, r.max_row_violation);
    out3[1] = r.max_dual_violation;
    out3[2] = r.objective;
  }
  copy_text(message, message_len, r.message);
  return r.ok ? 1 : 0;
}

int smy_verify_infeasibility(const smy_model* model, const double* y, char* message,
                             size_t message_len) {
  if (model == nullptr || y == nullptr) return -1;
  const VerifyReport r = verify_infeasibility(model->model, y);
  copy_text(message, message_len, r.message);
  return r.ok ? 1 : 0;
}

int smy_verify_unbounded_ray(const smy_model* model, const double* ray, char* message,
                             size_t message_len) {
  if (model == nullptr || ray == nullptr) return -1;
  const VerifyReport r = verify_unbounded_ray(model->model, ray);
  copy_text(message, message_len, r.message);
  return r.ok ? 1 : 0;
}

int smy_gpu_info(int* built_with_gpu_kernels, char* device, size_t device_len) {
  if (built_with_gpu_kernels) *built_with_gpu_kernels = 1;
  if (device != nullptr && device_len > 0) {
    device[0] = '\0';
    if (smy_gpu_available()) smy_gpu_device_name(device, device_len);
  }
  return SMY_OK;
}

}  // extern "C"
