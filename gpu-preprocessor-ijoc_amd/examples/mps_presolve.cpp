// mps_presolve.cpp -- an MPS file through the drop-in API: read it
// (tools::read_mps_file), put it in HBM (tools::upload_lp: duplicates summed
// and A, A^T built on the device), presolve it (preprocess) and print the
// outcome as key/value lines, the way the reference's MPS tool reports a run
// (cpp/tools/gpu_preprocessor_mps.cpp:29-66).
//
//   mps_presolve <model.mps> [max_iters]
//
// Exit status: 0 presolved, 1 infeasible or unbounded, 2 error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/gpu_preprocessor/gpu_preprocessor.hpp"

namespace gpp = gpu_preprocessor;

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: mps_presolve <model.mps> [max_iters]\n";
    return 2;
  }
  try {
    const auto t_parse = std::chrono::steady_clock::now();
    gpp::tools::MpsModel model = gpp::tools::read_mps_file(argv[1]);
    const double parse_seconds = seconds_since(t_parse);
    const auto t_upload = std::chrono::steady_clock::now();
    gpp::tools::DeviceLpOwner device = gpp::tools::upload_lp(model);
    const double upload_seconds = seconds_since(t_upload);

    gpp::Parameters params;
    params.enable_folding = false;
    params.enable_structure_specific_rules = false;
    if (argc > 2) params.max_iters = std::atoi(argv[2]);
    gpp::Summary summary = gpp::preprocess(device.lp, params);
    const auto& ps = summary.presolve;
    std::vector<std::int32_t> fixed = ps.record.fixed_idx;
    std::sort(fixed.begin(), fixed.end());
    fixed.erase(std::unique(fixed.begin(), fixed.end()), fixed.end());

    std::cout << std::setprecision(12);
    std::cout << "model " << model.name() << '\n';
    std::cout << "status " << (ps.has_infeasible ? "infeasible" : ps.has_unbounded ? "unbounded" : "ok") << '\n';
    std::cout << "changed " << ps.has_change << '\n';
    std::cout << "original_rows " << device.lp.A.rows << '\n';
    std::cout << "original_cols " << device.lp.A.cols << '\n';
    std::cout << "original_nnz " << device.lp.A.nnz << '\n';
    std::cout << "reduced_rows " << ps.reduced_rows << '\n';
    std::cout << "reduced_cols " << ps.reduced_cols << '\n';
    std::cout << "reduced_nnz " << ps.reduced_nnz << '\n';
    std::cout << "fixed_columns " << fixed.size() << '\n';
    std::cout << "objective_shift " << ps.obj_constant_delta << '\n';
    std::cout << "iterations " << ps.iterations << '\n';
    std::cout << "parse_seconds " << parse_seconds << '\n';
    std::cout << "upload_seconds " << upload_seconds << '\n';
    std::cout << "presolve_seconds " << ps.elapsed_seconds << '\n';
    std::cout << "preprocessing_seconds " << summary.preprocessing_seconds << '\n';
    const int rc = ps.has_infeasible || ps.has_unbounded ? 1 : 0;
    gpp::free_summary(summary);
    return rc;
  } catch (const std::exception& e) {
    std::cerr << e.what() << '\n';
    return 2;
  }
}
