// C ABI of include/samaya_io.h over the MPS reader and writers of mps.hpp.

#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/samaya_io.h"
#include "capi_model.hpp"
#include "mps.hpp"
#include "pdlp.hpp"

using namespace smy::host;

namespace {

void message(char* buf, size_t len, const std::string& s) {
  if (buf == nullptr || len == 0) return;
  const size_t k = std::min(s.size(), len - 1);
  std::memcpy(buf, s.data(), k);
  buf[k] = '\0';
}

smy_model* adopt(MpsModel&& mps) {
  auto md = std::make_unique<smy_model>();
  md->model = std::move(mps.model);
  md->name = std::move(mps.name);
  md->row_names = std::move(mps.row_names);
  md->col_names = std::move(mps.col_names);
  md->col_integer = std::move(mps.col_integer);
  return md.release();
}

MpsModel names_of(const smy_model* model) {
  MpsModel m;
  m.name = model->name;
  m.row_names = model->row_names;
  m.col_names = model->col_names;
  m.col_integer = model->col_integer;
  return m;
}

long long give(const std::string& s, char* buf, size_t cap) {
  const size_t need = s.size() + 1;
  if (buf != nullptr && cap >= need) std::memcpy(buf, s.c_str(), need);
  return static_cast<long long>(need);
}

// The status words of samaya::to_string(Status) (src/core/status.cpp:5-22) for the statuses an
// LP answer by PDLP maps to (src/core/solver.cpp:441-466).
const char* status_word(int status) {
  switch (status) {
    case SMY_STATUS_OPTIMAL: return "optimal";
    case SMY_STATUS_INFEASIBLE: return "infeasible";
    case SMY_STATUS_UNBOUNDED: return "unbounded";
    case SMY_STATUS_ITERATION_LIMIT: return "iteration_limit";
    case SMY_STATUS_TIME_LIMIT: return "time_limit";
    case SMY_STATUS_NUMERICAL_ERROR: return "numerical_error";
    default: return "not_solved";
  }
}

std::string solution(const smy_model* model, const smy_lp_result* r, const double* col_value,
                     const double* col_dual, const double* row_activity, const double* row_dual,
                     const double* certificate) {
  const Model& md = model->model;
  const size_t n = static_cast<size_t>(md.n), m = static_cast<size_t>(md.m);
  const auto vec = [](const double* p, size_t k) {
    return p != nullptr ? std::vector<double>(p, p + k) : std::vector<double>();
  };
  const bool optimal = r->status == SMY_STATUS_OPTIMAL;
  // Only answered statuses carry a point (src/core/solver.cpp:445-475).
  const bool answered = optimal || r->status == SMY_STATUS_INFEASIBLE || r->status == SMY_STATUS_UNBOUNDED;
  if (!answered) col_value = nullptr;
  const bool duals = optimal && col_dual != nullptr && row_dual != nullptr;
  std::vector<double> cv = vec(col_value, n), ra, cd, rd, cert, ray;
  if (col_value != nullptr) {
    ra = row_activity != nullptr ? vec(row_activity, m) : std::vector<double>(m, 0.0);
    if (row_activity == nullptr) {  // activity from the column values
      for (size_t j = 0; j < n; ++j)
        for (int64_t p = md.A.start[j]; p < md.A.start[j + 1]; ++p) ra[md.A.index[p]] += md.A.value[p] * cv[j];
    }
  }
  if (duals && col_value != nullptr) {
    cd = vec(col_dual, n);
    rd = vec(row_dual, m);
  }
  if (r->status == SMY_STATUS_INFEASIBLE) cert = vec(certificate, m);
  if (r->status == SMY_STATUS_UNBOUNDED) ray = vec(certificate, n);
  const double objective = optimal ? r->objective : std::nan("");
  return solution_text(status_word(r->status), r->verified != 0, objective, model->col_names,
                       model->row_names, cv, cd, ra, rd, cert, ray);
}

}  // namespace

extern "C" {

int smy_read_mps_text(const char* text, size_t len, int threads, smy_model** out, char* errbuf,
                      size_t errlen, long long* error_line) {
  if (out == nullptr || (text == nullptr && len > 0)) return SMY_INVALID;
  *out = nullptr;
  if (error_line != nullptr) *error_line = 0;
  message(errbuf, errlen, "");
  try {
    *out = adopt(read_mps_text(std::string_view(text == nullptr ? "" : text, len), threads));
    return SMY_OK;
  } catch (const MpsError& e) {
    message(errbuf, errlen, e.what());
    if (error_line != nullptr) *error_line = e.line();
    return SMY_INVALID;
  } catch (const std::exception& e) {
    message(errbuf, errlen, e.what());
    return SMY_ERROR;
  } catch (...) {
    return SMY_ERROR;
  }
}

int smy_read_mps(const char* path, smy_model** out, char* errbuf, size_t errlen) {
  if (out == nullptr || path == nullptr) return SMY_INVALID;
  *out = nullptr;
  message(errbuf, errlen, "");
  try {
    *out = adopt(read_mps_file(path, 0));
    return SMY_OK;
  } catch (const MpsError& e) {
    message(errbuf, errlen, e.what());
    return SMY_INVALID;
  } catch (const std::exception& e) {
    message(errbuf, errlen, e.what());
    return SMY_ERROR;
  } catch (...) {
    return SMY_ERROR;
  }
}

int smy_model_name(const smy_model* model, char* buf, size_t len) {
  if (model == nullptr) return SMY_INVALID;
  message(buf, len, model->name);
  return SMY_OK;
}

long long smy_model_names(const smy_model* model, int which, char* buf, size_t cap) {
  if (model == nullptr || (which != 0 && which != 1)) return -SMY_INVALID;
  const std::vector<std::string>& v = which == 0 ? model->row_names : model->col_names;
  size_t need = 1;
  for (const std::string& s : v) need += s.size() + 1;
  if (buf != nullptr && cap >= need) {
    char* p = buf;
    for (const std::string& s : v) {
      std::memcpy(p, s.data(), s.size());
      p += s.size();
      *p++ = '\n';
    }
    *p = '\0';
  }
  return static_cast<long long>(need);
}

int smy_model_set_names(smy_model* model, int which, const char* names, size_t len) {
  if (model == nullptr || (which != 0 && which != 1) || (names == nullptr && len > 0)) return SMY_INVALID;
  try {
    std::vector<std::string> v;
    size_t pos = 0;
    while (pos < len) {
      const void* hit = std::memchr(names + pos, '\n', len - pos);
      if (hit == nullptr) return SMY_INVALID;  // every name ends with '\n'
      const size_t end = static_cast<size_t>(static_cast<const char*>(hit) - names);
      v.emplace_back(names + pos, end - pos);
      pos = end + 1;
    }
    const size_t want = static_cast<size_t>(which == 0 ? model->model.m : model->model.n);
    if (!v.empty() && v.size() != want) return SMY_INVALID;
    (which == 0 ? model->row_names : model->col_names) = std::move(v);
    return SMY_OK;
  } catch (...) {
    return SMY_ERROR;
  }
}

int smy_model_set_integrality(smy_model* model, const char* marks) {
  if (model == nullptr) return SMY_INVALID;
  try {
    if (marks == nullptr)
      model->col_integer.clear();
    else
      model->col_integer.assign(marks, marks + model->model.n);
    return SMY_OK;
  } catch (...) {
    return SMY_ERROR;
  }
}

long long smy_model_integrality(const smy_model* model, char* out) {
  if (model == nullptr) return -SMY_INVALID;
  const size_t n = static_cast<size_t>(model->model.n);
  long long count = 0;
  for (size_t j = 0; j < n; ++j) {
    const char v = j < model->col_integer.size() ? model->col_integer[j] : 0;
    if (out != nullptr) out[j] = v;
    count += v != 0;
  }
  return count;
}

long long smy_write_mps_text(const smy_model* model, char* buf, size_t cap) {
  if (model == nullptr) return -SMY_INVALID;
  try {
    const MpsModel names = names_of(model);
    return give(write_mps(model->model, &names), buf, cap);
  } catch (...) {
    return -SMY_ERROR;
  }
}

int smy_write_mps(const smy_model* model, const char* path, char* errbuf, size_t errlen) {
  if (model == nullptr || path == nullptr) return SMY_INVALID;
  message(errbuf, errlen, "");
  try {
    const MpsModel names = names_of(model);
    const std::string text = write_mps(model->model, &names);
    std::ofstream out(path, std::ios::binary);
    if (!out) {
      message(errbuf, errlen, std::string("cannot open '") + path + "' for writing");
      return SMY_ERROR;
    }
    out.write(text.data(), static_cast<std::streamsize>(text.size()));
    if (!out) {
      message(errbuf, errlen, std::string("cannot write '") + path + "'");
      return SMY_ERROR;
    }
    return SMY_OK;
  } catch (const std::exception& e) {
    message(errbuf, errlen, e.what());
    return SMY_ERROR;
  } catch (...) {
    return SMY_ERROR;
  }
}

long long smy_solution_text(const smy_model* model, const smy_lp_result* result,
                            const double* col_value, const double* col_dual,
                            const double* row_activity, const double* row_dual,
                            const double* certificate, char* buf, size_t cap) {
  if (model == nullptr || result == nullptr) return -SMY_INVALID;
  try {
    return give(solution(model, result, col_value, col_dual, row_activity, row_dual, certificate), buf, cap);
  } catch (...) {
    return -SMY_ERROR;
  }
}

int smy_write_solution(const smy_model* model, const char* path, const smy_lp_result* result,
                       const double* col_value, const double* col_dual,
                       const double* row_activity, const double* row_dual,
                       const double* certificate, char* errbuf, size_t errlen) {
  if (model == nullptr || path == nullptr || result == nullptr) return SMY_INVALID;
  message(errbuf, errlen, "");
  try {
    const std::string text = solution(model, result, col_value, col_dual, row_activity, row_dual, certificate);
    std::ofstream out(path, std::ios::binary);
    if (!out) {
      message(errbuf, errlen, std::string("cannot open '") + path + "' for writing");
      return SMY_ERROR;
    }
    out.write(text.data(), static_cast<std::streamsize>(text.size()));
    return out ? SMY_OK : SMY_ERROR;
  } catch (const std::exception& e) {
    message(errbuf, errlen, e.what());
    return SMY_ERROR;
  } catch (...) {
    return SMY_ERROR;
  }
}

}  // extern "C"
