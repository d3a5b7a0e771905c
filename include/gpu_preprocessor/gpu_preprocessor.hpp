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
//  * Folding (enable_folding, on by default as in the reference) runs first;
//    presolve then runs on the folded LP and postprocess unfolds.  The
//    structure-specific Startup rules are not part of this engine: the fields
//    are accepted, the reduction is the one with them off, and
//    Summary::unsupported_stages reports a request for them.
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

// PostsolveTape (presolve_structs.hpp:317-332), host copy.
struct PostsolveTape {
  std::vector<std::int32_t> types;
  std::vector<NnzOffset> index_starts{0};
  std::vector<NnzOffset> value_starts{0};
  std::vector<std::uint8_t> dual_modes;
  std::vector<std::int32_t> indices;
  std::vector<double> vals;
};

// PresolveRecordGpu (presolve_structs.hpp:338-368): dimensions and objective
// constants always; the maps, fixed columns and tape stay on the device until
// export_recovery_record() copies them into the vectors.
struct PresolveRecordGpu {
  const gp_summary* device = nullptr;  // the owning summary's device record
  std::int32_t m0 = 0;
  std::int32_t n0 = 0;
  std::int32_t m1 = 0;
  std::int32_t n1 = 0;
  std::vector<std::int32_t> row_red2org;
  std::vector<std::int32_t> col_red2org;
  std::vector<std::int32_t> fixed_idx;
  std::vector<double> fixed_val;
  double obj_constant_old = 0.0;
  double obj_constant_new = 0.0;
  // postsolve of these reductions restores the primal point only
  bool has_primal_only_piecewise_linear_reformulation_variable_pair_reduction = false;
  bool has_primal_only_column_domination_reduction = false;
  bool has_primal_only_projected_auxiliary_reduction = false;
  PostsolveTape tape;
};

// export_recovery_record (device_recovery.hpp): device record -> host vectors.
inline void export_recovery_record(PresolveRecordGpu& record) {
  if (!record.device) return;
  gp_record_sizes sz{};
  detail::raise(gp_record_get_sizes(record.device, &sz));
  record.row_red2org.resize(static_cast<std::size_t>(sz.m1));
  record.col_red2org.resize(static_cast<std::size_t>(sz.n1));
  record.fixed_idx.resize(static_cast<std::size_t>(sz.fixed_count));
  record.fixed_val.resize(static_cast<std::size_t>(sz.fixed_count));
  PostsolveTape& t = record.tape;
  t.types.resize(static_cast<std::size_t>(sz.record_count));
  t.index_starts.resize(static_cast<std::size_t>(sz.record_count + 1));
  t.value_starts.resize(static_cast<std::size_t>(sz.record_count + 1));
  t.dual_modes.resize(static_cast<std::size_t>(sz.record_count));
  t.indices.resize(static_cast<std::size_t>(sz.index_count));
  t.vals.resize(static_cast<std::size_t>(sz.value_count));
  detail::raise(gp_record_copy(record.device, record.row_red2org.data(), record.col_red2org.data(),
                               record.fixed_idx.data(), record.fixed_val.data(), t.types.data(),
                               t.index_starts.data(), t.value_starts.data(), t.dual_modes.data(), t.indices.data(),
                               t.vals.data()));
}

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
  PresolveRecordGpu record;
  LPInfoGpu reduced_lp;
  bool owns_reduced_lp = false;
};

}  // namespace presolve

namespace folding {

// Folding profile and run summary (folding_structs.hpp:15-92), filled from
// the summary's folding stage: total time, rounds, refinement outcome and
// class counts (the finer per-phase timers stay zero), and, when applied, the
// map and the folded LP, both owned by the summary.
struct FoldingRefinementProfile {
  double total_seconds = 0.0;
  double init_color_seconds = 0.0;
  double spmv_at_seconds = 0.0;
  double refine_col_seconds = 0.0;
  double spmv_a_seconds = 0.0;
  double refine_row_seconds = 0.0;
  std::int32_t rounds = 0;
  bool refinement_ok = false;
  std::int32_t final_row_colors = 0;
  std::int32_t final_col_colors = 0;
};

struct FoldingReduceProfile {
  double total_seconds = 0.0;
  double count_color_seconds = 0.0;
  double emit_triplet_seconds = 0.0;
  double sort_reduce_seconds = 0.0;
  double copy_unique_to_host_seconds = 0.0;
  double host_filter_rebuild_seconds = 0.0;
  double host_csr_build_seconds = 0.0;
  double copy_reduced_matrix_to_device_seconds = 0.0;
  double transpose_seconds = 0.0;
  double sum_cost_seconds = 0.0;
  double row_bounds_seconds = 0.0;
  double col_bounds_seconds = 0.0;
  NnzOffset unique_entries = 0;
  NnzOffset kept_entries = 0;
};

struct FoldingMapProfile {
  double total_seconds = 0.0;
  double copy_color_ids_seconds = 0.0;
  double count_color_seconds = 0.0;
  double scale_seconds = 0.0;
};

struct FoldingProfile {
  double total_seconds = 0.0;
  FoldingRefinementProfile refinement;
  FoldingReduceProfile reduce;
  FoldingMapProfile map;
};

struct FoldingMapDevice {
  std::int32_t original_rows = 0;
  std::int32_t original_cols = 0;
  std::int32_t reduced_rows = 0;
  std::int32_t reduced_cols = 0;
  std::int32_t* row_color_id = nullptr;
  std::int32_t* col_color_id = nullptr;
  double* row_scale = nullptr;
  double* col_scale = nullptr;
  bool owns_buffers = false;
};

struct FoldingRunSummary {
  bool applied = false;
  std::int32_t original_rows = 0;
  std::int32_t original_cols = 0;
  std::int32_t folded_rows = 0;
  std::int32_t folded_cols = 0;
  FoldingProfile profile;
  FoldingMapDevice map;
  presolve::LPInfoGpu folded_lp;
  bool owns_folded_lp = false;
};

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
  FoldingRunSummary folding;
  presolve::GpuPresolveSummary presolve;
  // GP_STAGE_* bits of requested stages this engine does not run (no
  // reference counterpart: the reference runs them).
  std::int32_t unsupported_stages = 0;

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
    original_rows = raw_.original_rows;
    original_cols = raw_.original_cols;
    // presolve ran on the folded LP when folding applied
    folded_rows = folding_applied ? raw_.folding.folded_rows : raw_.original_rows;
    folded_cols = folding_applied ? raw_.folding.folded_cols : raw_.original_cols;
    presolve.original_rows = folded_rows;
    presolve.original_cols = folded_cols;
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
    presolve.record.device = raw_.record ? &raw_ : nullptr;
    presolve.record.m0 = folded_rows;
    presolve.record.n0 = folded_cols;
    presolve.record.m1 = raw_.reduced_rows;
    presolve.record.n1 = raw_.reduced_cols;
    presolve.record.obj_constant_new = raw_.reduced_lp.obj_constant;
    presolve.record.has_primal_only_piecewise_linear_reformulation_variable_pair_reduction =
        (raw_.primal_only_reductions & GP_PRIMAL_ONLY_VARIABLE_PAIR) != 0;
    presolve.record.has_primal_only_column_domination_reduction =
        (raw_.primal_only_reductions & GP_PRIMAL_ONLY_COLUMN_DOMINATION) != 0;
    presolve.record.has_primal_only_projected_auxiliary_reduction =
        (raw_.primal_only_reductions & GP_PRIMAL_ONLY_PROJECTED_AUXILIARY) != 0;
    presolve.reduced_lp = *reinterpret_cast<const presolve::LPInfoGpu*>(&raw_.reduced_lp);
    presolve.owns_reduced_lp = raw_.owns_reduced_lp != 0;
    folding = FoldingRunSummary{};
    const gp_fold_result& f = raw_.folding;
    folding.applied = f.applied != 0;
    folding.original_rows = raw_.original_rows;
    folding.original_cols = raw_.original_cols;
    folding.folded_rows = folded_rows;
    folding.folded_cols = folded_cols;
    folding.profile.total_seconds = f.seconds;
    folding.profile.refinement.rounds = f.rounds;
    folding.profile.refinement.refinement_ok = f.refinement_ok != 0;
    folding.profile.refinement.final_row_colors = f.final_row_colors;
    folding.profile.refinement.final_col_colors = f.final_col_colors;
    if (folding.applied) {
      // device arrays owned by the summary's record (freed by free_summary)
      folding.map = FoldingMapDevice{f.original_rows, f.original_cols, f.folded_rows, f.folded_cols,
                                     f.row_color_id,  f.col_color_id,  f.row_scale,   f.col_scale, false};
      folding.folded_lp = *reinterpret_cast<const presolve::LPInfoGpu*>(&f.folded_lp);
    }
    unsupported_stages = raw_.unsupported_stages;
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
    folding = o.folding;
    presolve = std::move(o.presolve);
    unsupported_stages = o.unsupported_stages;
    raw_ = o.raw_;
    // the record view follows the summary it belongs to
    if (presolve.record.device) presolve.record.device = &raw_;
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
  s.presolve.record.obj_constant_old = input.obj_constant;
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

namespace tools {

// An MPS file in host memory (mps_reader::MpsModel, mpsreader.hpp:14-17): its
// name, dimensions, vectors and matrix entries in file order.  Move-only.
class MpsModel {
 public:
  MpsModel() = default;
  MpsModel(const MpsModel&) = delete;
  MpsModel& operator=(const MpsModel&) = delete;
  MpsModel(MpsModel&& o) noexcept : raw_(std::exchange(o.raw_, gp_mps_model{})) {}
  MpsModel& operator=(MpsModel&& o) noexcept {
    if (this != &o) {
      gp_mps_free(&raw_);
      raw_ = std::exchange(o.raw_, gp_mps_model{});
    }
    return *this;
  }
  ~MpsModel() { gp_mps_free(&raw_); }

  std::string name() const { return raw_.name ? raw_.name : ""; }
  int num_rows() const noexcept { return raw_.rows; }
  int num_cols() const noexcept { return raw_.cols; }
  double obj_constant() const noexcept { return raw_.obj_constant; }
  const gp_mps_model& raw() const noexcept { return raw_; }
  gp_mps_model* raw_ptr() noexcept { return &raw_; }

 private:
  gp_mps_model raw_{};
};

// read_mps_file (gpu_preprocessor_tool_common.hpp:143-154) and read_mps on a
// text already in memory (mpsreader.hpp:19).
inline MpsModel read_mps_file(const std::string& path) {
  MpsModel model;
  detail::raise(gp_mps_read(path.c_str(), model.raw_ptr()));
  return model;
}
inline MpsModel read_mps_text(const std::string& text) {
  MpsModel model;
  detail::raise(gp_mps_parse(text.data(), text.size(), model.raw_ptr()));
  return model;
}

// DeviceLpOwner (gpu_preprocessor_tool_common.hpp:77-106): a device LP that
// frees its arrays on destruction.  Move-only.
struct DeviceLpOwner {
  DeviceLp lp;
  DeviceLpOwner() = default;
  DeviceLpOwner(const DeviceLpOwner&) = delete;
  DeviceLpOwner& operator=(const DeviceLpOwner&) = delete;
  DeviceLpOwner(DeviceLpOwner&& o) noexcept : lp(std::exchange(o.lp, DeviceLp{})) {}
  DeviceLpOwner& operator=(DeviceLpOwner&& o) noexcept {
    if (this != &o) {
      reset();
      lp = std::exchange(o.lp, DeviceLp{});
    }
    return *this;
  }
  ~DeviceLpOwner() { reset(); }
  void reset() noexcept {
    gp_lp_free(reinterpret_cast<gp_device_lp*>(&lp));
    lp = DeviceLp{};
  }
};

// upload_lp (gpu_preprocessor_tool_common.hpp:108-141): the model in HBM,
// duplicates summed and A, A^T assembled on the device.
inline DeviceLpOwner upload_lp(const MpsModel& model) {
  DeviceLpOwner owner;
  detail::raise(gp_mps_upload(&model.raw(), reinterpret_cast<gp_device_lp*>(&owner.lp)));
  return owner;
}

}  // namespace tools

}  // namespace gpu_preprocessor
