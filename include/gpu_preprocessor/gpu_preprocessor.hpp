// gpu_preprocessor.hpp -- header-only C++ front end of the MI355X presolve
// engine with the names of the reference's public API
// (cpp/include/gpu_preprocessor/gpu_preprocessor.hpp:9-39):
//   gpu_preprocessor::{Parameters, DeviceLp, Summary, Solution,
//                      preprocess, postprocess, free_summary}
// All work goes through the C ABI of gpupre.h (libgpupre.so).  The device LP
// types are layout-identical to the ABI structs (checked below), so a caller
// that fills a reference-style LPInfoGpu hands it over without copies.
//
// Contract (same as the reference):
//  * preprocess(): the input LP is caller-owned device memory and is not
//    modified; keep it and the returned Summary alive for postprocess().  On a
//    run that changes nothing the reduced LP borrows the input arrays
//    (presolve.owns_reduced_lp == false).  Summary owns device memory: it is
//    move-only.
//  * postprocess(): x, y, z are caller-owned device arrays in reduced order
//    (reduced_cols, reduced_rows, reduced_cols); the result is host vectors in
//    original order.  Dual convention c - A^T y - z = 0.
//  * free_summary(): releases the reduced LP (when owned) and the recovery
//    record; safe to call repeatedly.
//  * Errors: std::invalid_argument, std::length_error, std::runtime_error as in
//    the reference; infeasible / unbounded are flags, not errors.
//  * Folding (enable_folding) and the structure-specific Startup rules are not
//    part of this engine: the fields are accepted and ignored.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../gpupre.h"

namespace gpu_preprocessor {

using NnzOffset = gp_offset;

namespace detail {
inline void raise(int code) {
  if (code == GP_OK) return;
  const std::string msg = gp_last_error() ? gp_last_error() : "gpu_preprocessor error";
  if (code == GP_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  if (code == GP_ERR_LENGTH) throw std::length_error(msg);
  throw std::runtime_error(msg);
}
}  // namespace detail

namespace presolve {

// presolve_structs.hpp:160-167
struct DeviceCsrMatrix {
  std::int32_t rows = 0;
  std::int32_t cols = 0;
  NnzOffset nnz = 0;
  NnzOffset* rowPtr = nullptr;
  std::int32_t* colVal = nullptr;
  double* nzVal = nullptr;
};

// presolve_structs.hpp:169-177 (AT is a plain CSR transpose here: the
// reference's AT slack fields have no counterpart)
struct LPInfoGpu {
  DeviceCsrMatrix A;
  DeviceCsrMatrix AT;
  double* c = nullptr;
  double* AL = nullptr;
  double* AU = nullptr;
  double* l = nullptr;
  double* u = nullptr;
  double obj_constant = 0.0;
};

static_assert(sizeof(DeviceCsrMatrix) == sizeof(gp_csr), "DeviceCsrMatrix layout");
static_assert(offsetof(DeviceCsrMatrix, rowPtr) == offsetof(gp_csr, row_ptr), "DeviceCsrMatrix layout");
static_assert(offsetof(DeviceCsrMatrix, nzVal) == offsetof(gp_csr, vals), "DeviceCsrMatrix layout");
static_assert(sizeof(LPInfoGpu) == sizeof(gp_device_lp), "LPInfoGpu layout");
static_assert(offsetof(LPInfoGpu, AT) == offsetof(gp_device_lp, AT), "LPInfoGpu layout");
static_assert(offsetof(LPInfoGpu, u) == offsetof(gp_device_lp, u), "LPInfoGpu layout");
static_assert(offsetof(LPInfoGpu, obj_constant) == offsetof(gp_device_lp, obj_constant), "LPInfoGpu layout");

// PresolveParams (presolve_structs.hpp:28-158): the ABI struct with the
// reference defaults filled in by the library.
struct PresolveParams : gp_params {
  PresolveParams() : gp_params() { gp_default_params(this); }
};

// GpuPresolveSummary (gpu_presolve.hpp:14-30), a view of the owning record.
struct GpuPresolveSummary {
  std::int32_t original_rows = 0;
  std::int32_t original_cols = 0;
  std::int32_t reduced_rows = 0;
  std::int32_t reduced_cols = 0;
  NnzOffset reduced_nnz = 0;
  std::int32_t iterations = 0;
  bool stopped_at_iteration_limit = false;
  bool has_change = false;
  double elapsed_seconds = 0.0;
  double obj_constant_delta = 0.0;
  bool has_infeasible = false;
  bool has_unbounded = false;
  const void* record = nullptr;  // opaque, device resident (gp_record_copy exports it)
  LPInfoGpu reduced_lp;
  bool owns_reduced_lp = false;
};

}  // namespace presolve

namespace folding {

// FoldingPipelineSummary (folding_structs.hpp:94-109).  Move-only: it owns
// the reduced LP and the recovery record until free_summary().
struct FoldingPipelineSummary {
  double preprocessing_seconds = 0.0;
  double workspace_setup_seconds = 0.0;
  double workspace_release_seconds = 0.0;
  bool presolve_enabled = true;
  bool folding_applied = false;
  std::int32_t original_rows = 0;
  std::int32_t original_cols = 0;
  std::int32_t folded_rows = 0;
  std::int32_t folded_cols = 0;
  presolve::GpuPresolveSummary presolve;

  FoldingPipelineSummary() : raw_() {}
  FoldingPipelineSummary(const FoldingPipelineSummary&) = delete;
  FoldingPipelineSummary& operator=(const FoldingPipelineSummary&) = delete;
  FoldingPipelineSummary(FoldingPipelineSummary&& o) noexcept { take(o); }
  FoldingPipelineSummary& operator=(FoldingPipelineSummary&& o) noexcept {
    if (this != &o) {
      gp_free_summary(&raw_);
      take(o);
    }
    return *this;
  }
  ~FoldingPipelineSummary() { gp_free_summary(&raw_); }

  const gp_summary& raw() const { return raw_; }
  gp_summary* raw_ptr() { return &raw_; }

  // Refresh the reference-style view from the ABI summary.
  void sync() {
    preprocessing_seconds = raw_.preprocessing_seconds;
    workspace_setup_seconds = raw_.workspace_setup_seconds;
    workspace_release_seconds = raw_.workspace_release_seconds;
    presolve_enabled = raw_.presolve_enabled != 0;
    folding_applied = raw_.folding_applied != 0;
    original_rows = folded_rows = raw_.original_rows;
    original_cols = folded_cols = raw_.original_cols;
    presolve.original_rows = raw_.original_rows;
    presolve.original_cols = raw_.original_cols;
    presolve.reduced_rows = raw_.reduced_rows;
    presolve.reduced_cols = raw_.reduced_cols;
    presolve.reduced_nnz = raw_.reduced_nnz;
    presolve.iterations = raw_.iterations;
    presolve.stopped_at_iteration_limit = raw_.stopped_at_iteration_limit != 0;
    presolve.has_change = raw_.has_change != 0;
    presolve.elapsed_seconds = raw_.elapsed_seconds;
    presolve.obj_constant_delta = raw_.obj_constant_delta;
    presolve.has_infeasible = raw_.has_infeasible != 0;
    presolve.has_unbounded = raw_.has_unbounded != 0;
    presolve.record = raw_.record;
    presolve.reduced_lp = *reinterpret_cast<const presolve::LPInfoGpu*>(&raw_.reduced_lp);
    presolve.owns_reduced_lp = raw_.owns_reduced_lp != 0;
  }

 private:
  gp_summary raw_;
  void take(FoldingPipelineSummary& o) {
    preprocessing_seconds = o.preprocessing_seconds;
    workspace_setup_seconds = o.workspace_setup_seconds;
    workspace_release_seconds = o.workspace_release_seconds;
    presolve_enabled = o.presolve_enabled;
    folding_applied = o.folding_applied;
    original_rows = o.original_rows;
    original_cols = o.original_cols;
    folded_rows = o.folded_rows;
    folded_cols = o.folded_cols;
    presolve = o.presolve;
    raw_ = o.raw_;
    o.raw_ = gp_summary();
    o.presolve = presolve::GpuPresolveSummary();
  }
};

// UnfoldedSolutionHost (folding_structs.hpp:111-115)
struct UnfoldedSolutionHost {
  std::vector<double> x;
  std::vector<double> y;
  std::vector<double> z;
};

}  // namespace folding

using Parameters = presolve::PresolveParams;
using DeviceLp = presolve::LPInfoGpu;
using Summary = folding::FoldingPipelineSummary;
using Solution = folding::UnfoldedSolutionHost;

// gpu_preprocessor.hpp:22-24
inline Summary preprocess(const DeviceLp& input, const Parameters& params = {}) {
  Summary s;
  detail::raise(gp_preprocess(reinterpret_cast<const gp_device_lp*>(&input), &params, s.raw_ptr()));
  s.sync();
  return s;
}

// gpu_preprocessor.hpp:26-32
inline Solution postprocess(double* x, double* y, double* z, const Summary& summary, const DeviceLp& original) {
  Solution out;
  out.x.resize(static_cast<std::size_t>(summary.raw().original_cols));
  out.y.resize(static_cast<std::size_t>(summary.raw().original_rows));
  out.z.resize(static_cast<std::size_t>(summary.raw().original_cols));
  detail::raise(gp_postprocess(x, y, z, &summary.raw(), reinterpret_cast<const gp_device_lp*>(&original),
                               out.x.data(), out.y.data(), out.z.data()));
  return out;
}

// gpu_preprocessor.hpp:35-39
inline void free_summary(Summary& summary) {
  gp_free_summary(summary.raw_ptr());
  summary = Summary{};
}

}  // namespace gpu_preprocessor
