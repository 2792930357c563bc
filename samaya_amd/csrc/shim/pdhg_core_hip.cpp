// samaya's gpu::PdhgDevice (src/gpu/pdhg_core.hpp:24-79) over this library's C ABI
// (include/samaya_pdhg.h). It replaces samaya's src/gpu/pdhg_core.cu in a samaya build that links
// libsamaya_amd.so: samaya's host loop (src/lp/pdlp.cpp) stays unchanged and calls the gfx950
// kernels through it. Compiled by the host compiler against samaya's own headers
// (-I<samaya>/src -I<samaya>/include) and this repository's include/; not part of this library.
//
// Semantics kept from pdhg_core.cu:
//   create returns null when there is no usable GPU or nnz does not fit 32-bit offsets
//   (pdhg_core.cu:246-247, here SMY_NO_DEVICE / SMY_TOO_LARGE) and throws std::runtime_error on
//   any other failure, as every method does (SAMAYA_CUDA_CHECK, src/gpu/device.hpp:83-91). So
//   samaya's pdlp.cpp:343-360 (create fails: CPU iterations) and :231-242 (a failure during the
//   solve: CPU re-solve) keep their meaning.
//   download resizes the caller's vectors (n + m or m entries) before copying.
#include <stdexcept>
#include <string>
#include <vector>

#include "gpu/pdhg_core.hpp"
#include "samaya_pdhg.h"

namespace samaya {
namespace gpu {
namespace {

void check(int rc, const char* what) {
  if (rc != SMY_OK) throw std::runtime_error(std::string(what) + ": " + smy_last_error());
}

smy_csc view(const SparseMatrix& m) {
  return smy_csc{static_cast<int32_t>(m.rows()), static_cast<int32_t>(m.cols()),
                 m.col_start().data(), m.row_index().data(), m.values().data()};
}

int which_of(PdhgDevice::Iterate w) {
  return w == PdhgDevice::Iterate::kCurrent ? SMY_CURRENT : SMY_AVERAGE;
}

}  // namespace

struct PdhgDevice::Impl {
  smy_pdhg* h = nullptr;
  int n = 0;
  int m = 0;
  ~Impl() { smy_pdhg_destroy(h); }
};

PdhgDevice::PdhgDevice(std::unique_ptr<Impl> impl) : impl_(std::move(impl)) {}
PdhgDevice::~PdhgDevice() = default;

std::unique_ptr<PdhgDevice> PdhgDevice::create(const SparseMatrix& A, const SparseMatrix& At,
                                               const std::vector<double>& cost,
                                               const std::vector<double>& lower,
                                               const std::vector<double>& upper) {
  const smy_csc a = view(A), at = view(At);
  auto impl = std::make_unique<Impl>();
  const int rc =
      smy_pdhg_create(&a, &at, cost.data(), lower.data(), upper.data(), SMY_FP64, &impl->h);
  if (rc == SMY_NO_DEVICE || rc == SMY_TOO_LARGE) return nullptr;
  check(rc, "PdhgDevice::create");
  impl->n = static_cast<int>(A.cols());
  impl->m = static_cast<int>(A.rows());
  return std::unique_ptr<PdhgDevice>(new PdhgDevice(std::move(impl)));
}

void PdhgDevice::step(double tau, double sigma, double theta, long long iter) {
  check(smy_pdhg_step(impl_->h, tau, sigma, theta, iter), "PdhgDevice::step");
}

void PdhgDevice::advance() { check(smy_pdhg_advance(impl_->h), "PdhgDevice::advance"); }

void PdhgDevice::restart() { check(smy_pdhg_restart(impl_->h), "PdhgDevice::restart"); }

void PdhgDevice::download(std::vector<double>* v, std::vector<double>* v_new,
                          std::vector<double>* y, std::vector<double>* y_new,
                          std::vector<double>* v_erg, std::vector<double>* y_erg) const {
  const size_t nt = static_cast<size_t>(impl_->n) + impl_->m, m = impl_->m;
  const auto sized = [](std::vector<double>* out, size_t k) -> double* {
    if (out == nullptr) return nullptr;
    out->resize(k);
    return out->data();
  };
  check(smy_pdhg_download(impl_->h, sized(v, nt), sized(v_new, nt), sized(y, m), sized(y_new, m),
                          sized(v_erg, nt), sized(y_erg, m)),
        "PdhgDevice::download");
}

void PdhgDevice::set_original(const std::vector<double>& col_scale,
                              const std::vector<double>& row_scale,
                              const std::vector<double>& cost, const std::vector<double>& lower,
                              const std::vector<double>& upper) {
  check(smy_pdhg_set_original(impl_->h, col_scale.data(), row_scale.data(), cost.data(),
                              lower.data(), upper.data()),
        "PdhgDevice::set_original");
}

PdhgDevice::Sums PdhgDevice::evaluate(Iterate which) {
  double o[7];
  check(smy_pdhg_evaluate(impl_->h, which_of(which), o), "PdhgDevice::evaluate");
  Sums s;
  s.primal_sq = o[0];
  s.dual_sq = o[1];
  s.primal_obj = o[2];
  s.dual_obj = o[3];
  s.x_sq = o[4];
  s.r_sq = o[5];
  s.y_sq = o[6];
  return s;
}

void PdhgDevice::restart_distance(Iterate which, double& dv_sq, double& dy_sq) {
  check(smy_pdhg_restart_distance(impl_->h, which_of(which), &dv_sq, &dy_sq),
        "PdhgDevice::restart_distance");
}

void PdhgDevice::restart_to(Iterate which) {
  check(smy_pdhg_restart_to(impl_->h, which_of(which)), "PdhgDevice::restart_to");
}

}  // namespace gpu
}  // namespace samaya
