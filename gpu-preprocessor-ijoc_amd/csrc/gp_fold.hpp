// gp_fold.hpp -- LP folding by colour refinement (the preprocess stage that
// runs before presolve; gp_fold.hip).
#pragma once

#include <memory>

#include "gp_common.hpp"

namespace gp {

// A folded LP and the map back to the original one.  Rows (columns) of one
// class share a 1-based class id; the folded LP has one row (column) per
// class.  Unfolding copies x, and scales y and z by 1 / class size.
struct FoldState {
  int32_t m = 0, n = 0;    // original dimensions
  int32_t mf = 0, nf = 0;  // classes
  bool applied = false;
  bool refinement_ok = true;
  int32_t rounds = 0;
  double seconds = 0.0;
  gp_device_lp lp{};  // folded LP (pool-owned) when applied
  DVec<int32_t> row_class, col_class;
  DVec<double> row_scale, col_scale;
  FoldState() = default;
  FoldState(const FoldState&) = delete;
  FoldState& operator=(const FoldState&) = delete;
  ~FoldState();
  // the public view (the state pointer left null: the owner frees it)
  void fill(gp_fold_result& r) const;
};

// run_folding (cpp/src/folding/folding.cu:69-136): never null; applied only
// when the refinement reduces one dimension below 80% and changes the LP.
std::unique_ptr<FoldState> fold_lp(const gp_device_lp& lp, double tolerance);

// x[j] = x_red[class(j)], z[j] = z_red[class(j)] / |class(j)|,
// y[i] = y_red[class(i)] / |class(i)|; device in, device out (any may be null).
void unfold(const FoldState& f, const double* x_red, const double* y_red, const double* z_red, double* x, double* y,
            double* z);

}  // namespace gp
