// gp_rules.hip -- HIP kernels of the cleanup and propagation rules:
//   fixed variables      rule_infeasible_and_fixed_variables.cu:65-348
//   empty columns/rows   empty_rows_and_columns/{columns.cu:64-250, rows.cu:25-120}
//   singleton rows       rule_singleton_rows.cu:76-278
//   redundant rows       rule_infeasible_and_redundant_rows.cu:46-332
//   bound strengthening  rule_bound_strengthening.cu:349-1557
// Row sums (row shifts, error budgets, activities) use the fixed blocked
// Neumaier order of RowAcc in CSR order, and candidate bounds use total-order
// atomic max/min, so every result is independent of scheduling.
#include "gp_tape.hpp"

namespace gp {

namespace {
inline bool tape_on(const gp_plan& plan, const gp_params& p) { return plan.tape && p.record_postsolve_tape; }

// ---------------------------------------------------------------- fixed variables
__global__ void k_fixed_candidates(int32_t n, const uint8_t* keep_col, const double* l, const double* u,
                                   const double* c, double bound_tol, uint8_t* fixed, double* val,
                                   double* radius, int32_t* flags) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint8_t f = 0;
  double v = 0.0, r = 0.0;
  if (keep_col[j]) {
    const double lj = l[j], uj = u[j];
    if (isfinite(lj) && isfinite(uj) && lj <= uj && uj - lj <= bound_tol) {
      v = lj + 0.5 * (uj - lj);
      r = cmax(v - lj, uj - v);
      f = 1;
    }
  }
  fixed[j] = f;
  val[j] = v;
  radius[j] = r;
  if (f) wave_flag(flags + 1, 1);
  if (r > 0.0) wave_flag(flags + 2, 1);
}

__global__ void k_row_error(gp_csr A, const uint8_t* keep_row, const uint8_t* fixed, const double* radius,
                            double* row_error) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  RowAcc acc;
  if (keep_row[i]) {
    const gp_offset p0 = A.row_ptr[i];
    for (gp_offset p = p0; p < A.row_ptr[i + 1]; ++p) {
      acc.pos(p - p0);
      const int32_t j = A.col_idx[p];
      if (fixed[j] && radius[j] > 0.0) acc.add(fabs(A.vals[p]) * radius[j]);
    }
  }
  row_error[i] = keep_row[i] ? acc.value() : 0.0;
}

__global__ void k_fixed_filter(gp_csr AT, const uint8_t* keep_row, const double* l, const double* u,
                               const double* row_error, double budget, uint8_t* fixed, int32_t* count) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= AT.rows || !fixed[j]) return;
  if (l[j] != u[j]) {
    for (gp_offset p = AT.row_ptr[j]; p < AT.row_ptr[j + 1]; ++p) {
      const int32_t r = AT.col_idx[p];
      if (keep_row[r] && !(row_error[r] <= budget)) {
        fixed[j] = 0;
        return;
      }
    }
  }
  wave_count(count);
}

// row_shift[i] = sum over entries with fixed[col] of a * val[col] (CSR order)
__global__ void k_fix_columns(int32_t n, const uint8_t* fixed, const double* val, double* l, double* u,
                              uint8_t* keep_col) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n || !fixed[j]) return;
  l[j] = val[j];
  u[j] = val[j];
  keep_col[j] = 0;
}
}  // namespace

void rule_fixed_variables(gp_plan& plan, const gp_device_lp& lp, const gp_params& p) {
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t n = lp.A.cols, m = lp.A.rows;
  if (n == 0) return;
  // l, u, c and keep_col per column; AL, AU per row
  TimedScope ts(GP_KT_CLEANUP, 25.0 * n + 16.0 * m);
  DVec<uint8_t> fixed(n);
  DVec<double> val(n), radius(n), row_error(m ? m : 1);
  Flags flags;
  hipLaunchKernelGGL(k_fixed_candidates, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, plan.keep_col_mask, plan.new_l,
                     plan.new_u, plan.new_c, p.bound_tol, fixed.data(), val.data(), radius.data(), flags.d.data());
  GP_LAUNCH_CHECK("fixed candidates");
  const int32_t* cf = flags.read();
  if (!cf[1]) return;  // no candidate
  // With every radius zero the error budget of every row is zero and the
  // filter keeps all candidates; the budget pass is skipped.
  if (cf[2]) {
    if (m > 0) {
      hipLaunchKernelGGL(k_row_error, dim3(grid_for(m)), dim3(kBlock), 0, 0, lp.A, plan.keep_row_mask, fixed.data(),
                         radius.data(), row_error.data());
      GP_LAUNCH_CHECK("fixed row error");
    }
    hipLaunchKernelGGL(k_fixed_filter, dim3(grid_for(n)), dim3(kBlock), 0, 0, lp.AT, plan.keep_row_mask, plan.new_l,
                       plan.new_u, row_error.data(), 0.1 * p.feasibility_tol, fixed.data(), flags.d.data());
    GP_LAUNCH_CHECK("fixed filter");
    if (flags.read()[0] == 0) return;
  }
  const double obj = fixed_variable_records(tape_on(plan, p) && p.record_postsolve_tape_cpu ? plan_tape(plan) : nullptr,
                                            fixed.data(), val.data(), plan.keep_row_mask, plan.new_c, lp.AT, n, true);
  if (m > 0) {
    shift_rows_by_columns(lp.A, &lp.AT, plan.keep_row_mask, fixed.data(), val.data(), plan.new_AL,
                          plan.new_AU);
    GP_LAUNCH_CHECK("fixed row shift");
  }
  hipLaunchKernelGGL(k_fix_columns, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, fixed.data(), val.data(), plan.new_l,
                     plan.new_u, plan.keep_col_mask);
  GP_LAUNCH_CHECK("fixed apply");
  plan.obj_constant_delta += obj;
  plan.has_col_action = 1;
  plan.has_change = 1;
}

// ---------------------------------------------------------------- empty columns
namespace {
__global__ void k_empty_cols(int32_t n, const uint8_t* keep_col, const uint8_t* empty, const double* c,
                             const double* l, const double* u, double bound_tol, double zero_tol, uint8_t* fixed,
                             double* val, int32_t* flags, BatchMarks bm) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  fixed[j] = 0;
  val[j] = 0.0;
  if (!(keep_col[j] && empty[j])) return;
  const double cj = c[j], lj = l[j], uj = u[j];
  if (lj > uj + bound_tol) {
    wave_flag(flags + 0, 1);
    bm.col(j, 1);
    return;
  }
  if ((cj > zero_tol && !isfinite(lj)) || (cj < -zero_tol && !isfinite(uj))) {
    wave_flag(flags + 1, 1);
    bm.col(j, 2);
    return;
  }
  double v = 0.0;
  if (cj > zero_tol) v = lj;
  else if (cj < -zero_tol) v = uj;
  else if (isfinite(lj)) v = lj;
  else if (isfinite(uj)) v = uj;
  fixed[j] = 1;
  val[j] = v;
  wave_flag(flags + 2, 1);
}
}  // namespace

void rule_empty_columns(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t n = lp.A.cols;
  if (n == 0) return;
  // col_nnz, keep_col, c, l, u per column
  TimedScope ts(GP_KT_CLEANUP, 29.0 * n);
  DVec<uint8_t> fixed(n);
  DVec<double> val(n);
  Flags flags;
  hipLaunchKernelGGL(k_empty_cols, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, plan.keep_col_mask, s.empty_col_mask,
                     plan.new_c, plan.new_l, plan.new_u, p.bound_tol, p.zero_tol, fixed.data(), val.data(),
                     flags.d.data(), batch_marks());
  GP_LAUNCH_CHECK("empty columns");
  const int32_t* f = flags.read();
  if (f[0]) {
    plan.has_infeasible = 1;
  } else if (f[1]) {
    plan.has_unbounded = 1;
  } else if (f[2]) {
    const double obj = fixed_variable_records(tape_on(plan, p) && p.record_postsolve_tape_cpu ? plan_tape(plan) : nullptr,
                                              fixed.data(), val.data(), plan.keep_row_mask, plan.new_c, lp.AT, n,
                                              false);
    hipLaunchKernelGGL(k_fix_columns, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, fixed.data(), val.data(), plan.new_l,
                       plan.new_u, plan.keep_col_mask);
    GP_LAUNCH_CHECK("empty columns apply");
    plan.obj_constant_delta += obj;
    plan.has_col_action = 1;
    plan.has_change = 1;
  }
}

// ---------------------------------------------------------------- empty rows
namespace {
__global__ void k_empty_rows(int32_t m, const uint8_t* keep_row, const double* AL, const double* AU,
                             const int32_t* row_nnz, double tol, int32_t* flags, BatchMarks bm) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m || !(keep_row[i] && row_nnz[i] == 0)) return;
  if (AL[i] <= tol && AU[i] >= -tol) {
    wave_flag(flags + 1, 1);
  } else {
    wave_flag(flags + 0, 1);
    bm.row(i, 1);
  }
}
__global__ void k_empty_rows_apply(int32_t m, uint8_t* keep_row, const double* AL, const double* AU,
                                   const int32_t* row_nnz, double tol) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m || !(keep_row[i] && row_nnz[i] == 0)) return;
  if (AL[i] <= tol && AU[i] >= -tol) keep_row[i] = 0;
}
}  // namespace

void rule_empty_rows(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows;
  if (m == 0) return;
  // row_nnz, keep_row, AL, AU per row
  TimedScope ts(GP_KT_CLEANUP, 21.0 * m);
  Flags flags;
  hipLaunchKernelGGL(k_empty_rows, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, plan.keep_row_mask, plan.new_AL,
                     plan.new_AU, s.row_nnz, p.feasibility_tol, flags.d.data(), batch_marks());
  GP_LAUNCH_CHECK("empty rows");
  const int32_t* f = flags.read();
  if (f[0]) {
    plan.has_infeasible = 1;
    return;
  }
  if (f[1]) {
    hipLaunchKernelGGL(k_empty_rows_apply, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, plan.keep_row_mask, plan.new_AL,
                       plan.new_AU, s.row_nnz, p.feasibility_tol);
    GP_LAUNCH_CHECK("empty rows apply");
    plan.has_row_action = 1;
    plan.has_row_deletion = 1;
    plan.has_change = 1;
  }
}

// ---------------------------------------------------------------- singleton rows
namespace {
// Candidate bounds exist only for the columns of live singleton rows: the
// seed pass marks each such column and copies its current bounds, the next
// pass tightens them.  Every other column keeps its bounds, as the full
// copy-and-tighten of the restatement leaves them.
__global__ void k_singleton_seed(int32_t m, int32_t n, const uint8_t* keep_row, const uint8_t* singleton,
                                 const int32_t* scol, const double* sval, double zero_tol, const double* l,
                                 const double* u, double* cl, double* cu, uint8_t* touched, uint8_t* remove) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint8_t r = 0;
  if (keep_row[i] && singleton[i]) {
    const int32_t j = scol[i];
    if (j >= 0 && j < n && fabs(sval[i]) > zero_tol) {
      // rows sharing a column store the same values
      touched[j] = 1;
      cl[j] = l[j];
      cu[j] = u[j];
      r = 1;
    }
  }
  remove[i] = r;
}
__global__ void k_singleton_rows(int32_t m, const int32_t* scol, const double* sval, const double* AL,
                                 const double* AU, const uint8_t* remove, double* cl, double* cu) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m || !remove[i]) return;
  const int32_t j = scol[i];
  const double a = sval[i];
  const double lo = AL[i], hi = AU[i];
  const double il = a > 0.0 ? lo / a : hi / a;
  const double iu = a > 0.0 ? hi / a : lo / a;
  atomic_total_max(cl + j, il);
  atomic_total_min(cu + j, iu);
}
__global__ void k_singleton_check(int32_t m, int32_t n, const double* cl, const double* cu, const double* l,
                                  const double* u, const uint8_t* touched, const uint8_t* remove, double tol,
                                  int32_t* flags, BatchMarks bm) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n) {
    const double lj = l[j], uj = u[j];
    const bool t = touched[j] != 0;
    const double lo = t ? cl[j] : lj, hi = t ? cu[j] : uj;
    if (lo > hi + tol) {
      wave_flag(flags + 0, 1);
      bm.col(j, 1);
    }
    if (lo != lj || hi != uj) wave_flag(flags + 2, 1);
  }
  if (j < m && remove[j]) wave_flag(flags + 1, 1);
}
__global__ void k_singleton_commit(int32_t m, int32_t n, const double* cl, const double* cu, const uint8_t* touched,
                                   const uint8_t* remove, uint8_t* keep_row, double* l, double* u) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < m && remove[j]) keep_row[j] = 0;
  if (j < n && touched[j]) {
    l[j] = cl[j];
    u[j] = cu[j];
  }
}
__global__ void k_deleted_row_counts(int32_t m, const uint8_t* sel, int64_t ni, int64_t nv, int64_t* r, int64_t* ci,
                                     int64_t* cv) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  r[i] = sel[i] ? 1 : 0;
  ci[i] = sel[i] ? ni : 0;
  cv[i] = sel[i] ? nv : 0;
}
// DeletedRow records: {row[, col]} / {AL, AU[, a]}
__global__ void k_deleted_row_pack(TapeView out, int32_t m, const uint8_t* sel, const int32_t* scol,
                                   const double* AL, const double* AU, const double* sval, const int64_t* r,
                                   const int64_t* ci, const int64_t* cv, int with_support) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m || !sel[i]) return;
  const int64_t rec = out.rbase + r[i];
  int64_t ip = out.ibase + ci[i], vp = out.vbase + cv[i];
  out.types[rec] = GP_RED_DELETED_ROW;
  out.modes[rec] = GP_DUAL_MINIMAL;
  out.idx[ip++] = static_cast<int32_t>(i);
  if (with_support) out.idx[ip++] = scol[i];
  out.vals[vp++] = AL[i];
  out.vals[vp++] = AU[i];
  if (with_support) out.vals[vp++] = sval[i];
  out.istarts[rec + 1] = static_cast<gp_offset>(ip);
  out.vstarts[rec + 1] = static_cast<gp_offset>(vp);
}
void append_deleted_rows(DeviceTape* t, int32_t m, const uint8_t* sel, const int32_t* scol, const double* AL,
                         const double* AU, const double* sval, bool with_support) {
  AppendCounts cnt(m);
  hipLaunchKernelGGL(k_deleted_row_counts, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, sel, with_support ? 2 : 1,
                     with_support ? 3 : 2, cnt.r.data(), cnt.i.data(), cnt.v.data());
  GP_LAUNCH_CHECK("deleted row counts");
  cnt.scan(m);
  if (cnt.R == 0) return;
  t->reserve_more(cnt.R, cnt.I, cnt.V);
  TapeView v = tape_view(*t);
  hipLaunchKernelGGL(k_deleted_row_pack, dim3(grid_for(m)), dim3(kBlock), 0, 0, v, m, sel, scol, AL, AU, sval,
                     cnt.r.data(), cnt.i.data(), cnt.v.data(), with_support ? 1 : 0);
  GP_LAUNCH_CHECK("deleted row pack");
  t->commit(cnt.R, cnt.I, cnt.V);
}
}  // namespace

void rule_singleton_rows(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0) return;
  // singleton mask, column and value, keep_row, AL, AU per row; l, u and a mark per column
  TimedScope ts(GP_KT_CLEANUP, 30.0 * m + 17.0 * n);
  DVec<double> cl(n ? n : 1), cu(n ? n : 1);
  DVec<uint8_t> remove(m), touched(n ? n : 1);
  fill_u8(touched.data(), 0, n);
  hipLaunchKernelGGL(k_singleton_seed, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, n, plan.keep_row_mask,
                     s.singleton_row_mask, s.singleton_row_col, s.singleton_row_val, p.zero_tol, plan.new_l,
                     plan.new_u, cl.data(), cu.data(), touched.data(), remove.data());
  hipLaunchKernelGGL(k_singleton_rows, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, s.singleton_row_col,
                     s.singleton_row_val, plan.new_AL, plan.new_AU, remove.data(), cl.data(), cu.data());
  GP_LAUNCH_CHECK("singleton rows");
  Flags flags;
  const int32_t mn = m > n ? m : n;
  hipLaunchKernelGGL(k_singleton_check, dim3(grid_for(mn)), dim3(kBlock), 0, 0, m, n, cl.data(), cu.data(),
                     plan.new_l, plan.new_u, touched.data(), remove.data(), p.bound_tol, flags.d.data(),
                     batch_marks());
  GP_LAUNCH_CHECK("singleton rows check");
  const int32_t* f = flags.read();
  if (f[0]) {
    plan.has_infeasible = 1;
  } else if (f[1] || f[2]) {
    if (tape_on(plan, p) && p.record_postsolve_tape_cpu && f[1])
      append_deleted_rows(plan_tape(plan), m, remove.data(), s.singleton_row_col, plan.new_AL, plan.new_AU,
                          s.singleton_row_val, true);
    hipLaunchKernelGGL(k_singleton_commit, dim3(grid_for(mn)), dim3(kBlock), 0, 0, m, n, cl.data(), cu.data(),
                       touched.data(), remove.data(), plan.keep_row_mask, plan.new_l, plan.new_u);
    GP_LAUNCH_CHECK("singleton rows commit");
    plan.has_row_action = 1;
    plan.has_row_deletion = plan.has_row_deletion || f[1] != 0;
    plan.has_change = 1;
  }
}

// ---------------------------------------------------------------- redundant rows
namespace {
struct Act {
  RowAcc fin;
  int32_t ninf = 0, pinf = 0, nan = 0;
  __device__ void add(double t) {
    if (isfinite(t)) fin.add(t);
    else if (t != t) ++nan;
    else if (t < 0.0) ++ninf;
    else ++pinf;
  }
  __device__ double value() const {
    if (nan || (ninf && pinf)) return __longlong_as_double(0x7ff8000000000000LL);
    if (ninf) return -INFINITY;
    if (pinf) return INFINITY;
    return fin.value();
  }
};

// Row fold of the redundancy test: full activity range of each live row
// with two or more live entries that is not an equality.  The entries' (l, u)
// pairs come from the entry-bounds store where their column is unchanged
// since the store's last refresh (normally bound strengthening's sweep right
// before, on the same plan), else from the column arrays.
struct RedundantFold {
  const uint8_t* keep_row;
  const int32_t* row_nnz;
  const double* AL;
  const double* AU;
  double tol;
  uint8_t* full;
  uint8_t* dl;
  uint8_t* du;
  int32_t* flags;
  BatchMarks bm;
  struct State {
    Act lo, hi;
  };
  __device__ bool active(int64_t i) const {
    if (!keep_row[i] || row_nnz[i] <= 1) return false;
    const double al = AL[i], au = AU[i];
    return !(isfinite(al) && isfinite(au) && fabs(au - al) <= tol);
  }
  __device__ bool use(double) const { return true; }
  __device__ void init(State& s) const {
    s.lo = Act{};
    s.hi = Act{};
  }
  __device__ void add(State& s, int64_t k, double lo, double hi, bool skip) const {
    s.lo.fin.pos(k);
    s.hi.fin.pos(k);
    if (skip) return;
    s.lo.add(lo);
    s.hi.add(hi);
  }
  __device__ void finish(int64_t i, const State& s, bool active) const {
    full[i] = dl[i] = du[i] = 0;
    if (!active) return;
    const double al = AL[i], au = AU[i];
    const bool lf = isfinite(al), uf = isfinite(au);
    const double rmin = s.lo.value(), rmax = s.hi.value();
    if ((lf && rmax < al - tol) || (uf && rmin > au + tol)) {
      wave_flag(flags + 0, 1);
      bm.row(i, 1);
      return;
    }
    const bool li = !lf || rmin >= al - tol;
    const bool ui = !uf || rmax <= au + tol;
    if (li && ui) {
      full[i] = 1;
      wave_flag(flags + 1, 1);
    } else if (lf && li) {
      dl[i] = 1;
      wave_flag(flags + 2, 1);
    } else if (uf && ui) {
      du[i] = 1;
      wave_flag(flags + 3, 1);
    }
  }
};
__global__ void k_redundant_apply(int32_t m, const uint8_t* full, const uint8_t* dl, const uint8_t* du,
                                  uint8_t* keep_row, double* AL, double* AU) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (full[i]) keep_row[i] = 0;
  if (dl[i]) AL[i] = -INFINITY;
  if (du[i]) AU[i] = INFINITY;
}
__global__ void k_row_side_counts(int32_t m, const uint8_t* dl, const uint8_t* du, int64_t* r, int64_t* ci,
                                  int64_t* cv) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t k = (dl[i] ? 1 : 0) + (du[i] ? 1 : 0);
  r[i] = k;
  ci[i] = k;
  cv[i] = 4 * k;
}
__global__ void k_row_side_pack(TapeView out, int32_t m, const uint8_t* dl, const uint8_t* du, const double* AL,
                                const double* AU, const int64_t* r, const int64_t* ci, const int64_t* cv) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m || !(dl[i] || du[i])) return;
  int64_t rec = out.rbase + r[i], ip = out.ibase + ci[i], vp = out.vbase + cv[i];
  const double al = AL[i], au = AU[i];
  if (dl[i]) {
    out.types[rec] = GP_RED_ROW_LOWER_BOUND_CHANGE;
    out.modes[rec] = GP_DUAL_MINIMAL;
    out.idx[ip++] = static_cast<int32_t>(i);
    out.vals[vp++] = al;
    out.vals[vp++] = au;
    out.vals[vp++] = -INFINITY;
    out.vals[vp++] = au;
    out.istarts[rec + 1] = static_cast<gp_offset>(ip);
    out.vstarts[rec + 1] = static_cast<gp_offset>(vp);
    ++rec;
  }
  if (du[i]) {
    out.types[rec] = GP_RED_ROW_UPPER_BOUND_CHANGE;
    out.modes[rec] = GP_DUAL_MINIMAL;
    out.idx[ip++] = static_cast<int32_t>(i);
    out.vals[vp++] = al;
    out.vals[vp++] = au;
    out.vals[vp++] = al;
    out.vals[vp++] = INFINITY;
    out.istarts[rec + 1] = static_cast<gp_offset>(ip);
    out.vstarts[rec + 1] = static_cast<gp_offset>(vp);
  }
}
}  // namespace

void rule_redundant_rows(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows;
  if (m == 0) return;
  // outside the fold: the three decision flags and AL, AU per row
  TimedScope outer(GP_KT_BOUNDS, 19.0 * m);
  DVec<uint8_t> full(m), dl(m), du(m);
  Flags flags;
  const PackedMask live_cols(s.live_masks ? static_cast<const uint8_t*>(plan.keep_col_mask) : nullptr, lp.A.cols);
  const PackedBounds lu(plan.new_l, plan.new_u, lp.A.cols);
  {
  // A, (l, u) and a mask bit per column, keep_row, row_nnz, AL, AU, 3 flags out per row
  TimedScope ts(GP_KT_REDUNDANT, 4.0 * (m + 1) + 12.0 * lp.A.nnz + 16.0 * lp.A.cols + lp.A.cols / 8.0 + 24.0 * m);
  // the entry-bounds store (refreshed by bound strengthening's sweep right before) for unchanged columns
  DVec<uint32_t> current;
  const StoreView store = store_view(s.live_masks ? plan_store(plan) : nullptr, lp.A, lu.data(), plan.keep_col_mask,
                                     p.zero_tol, false, current);
  launch_row_fold(lp.A, lu.data(), live_cols.view(),
                  RedundantFold{plan.keep_row_mask, s.row_nnz, plan.new_AL, plan.new_AU, p.bound_tol, full.data(),
                                dl.data(), du.data(), flags.d.data(), batch_marks()},
                  store);
  GP_LAUNCH_CHECK("redundant rows");
  }
  const int32_t* f = flags.read();
  if (f[0]) {
    plan.has_infeasible = 1;
    return;
  }
  if (!(f[1] || f[2] || f[3])) return;
  if (tape_on(plan, p) && p.record_postsolve_tape_cpu) {
    DeviceTape* t = plan_tape(plan);
    if (f[1]) append_deleted_rows(t, m, full.data(), nullptr, plan.new_AL, plan.new_AU, nullptr, false);
    AppendCounts cnt(m);
    hipLaunchKernelGGL(k_row_side_counts, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, dl.data(), du.data(),
                       cnt.r.data(), cnt.i.data(), cnt.v.data());
    GP_LAUNCH_CHECK("row side counts");
    cnt.scan(m);
    if (cnt.R > 0) {
      t->reserve_more(cnt.R, cnt.I, cnt.V);
      TapeView v = tape_view(*t);
      hipLaunchKernelGGL(k_row_side_pack, dim3(grid_for(m)), dim3(kBlock), 0, 0, v, m, dl.data(), du.data(),
                         plan.new_AL, plan.new_AU, cnt.r.data(), cnt.i.data(), cnt.v.data());
      GP_LAUNCH_CHECK("row side pack");
      t->commit(cnt.R, cnt.I, cnt.V);
    }
  }
  hipLaunchKernelGGL(k_redundant_apply, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, full.data(), dl.data(), du.data(),
                     plan.keep_row_mask, plan.new_AL, plan.new_AU);
  GP_LAUNCH_CHECK("redundant apply");
  plan.has_row_action = 1;
  plan.has_redundant_row_deletion = plan.has_redundant_row_deletion || f[1] != 0;
  plan.has_row_deletion = plan.has_row_deletion || f[1] != 0;
  plan.has_change = 1;
}

// ---------------------------------------------------------------- bound strengthening
namespace {
__device__ inline double residual_min(double fin, int32_t ninf, double term) {
  if (isfinite(term)) return ninf == 0 ? fin - term : -INFINITY;
  if (term < 0.0) return ninf > 1 ? -INFINITY : fin;
  return fin;
}
__device__ inline double residual_max(double fin, int32_t pinf, double term) {
  if (isfinite(term)) return pinf == 0 ? fin - term : INFINITY;
  if (term > 0.0) return pinf > 1 ? INFINITY : fin;
  return fin;
}
__device__ inline void implied_bounds(double a, double lj, double uj, double al, double au, double min_fin,
                                      int32_t min_ninf, double max_fin, int32_t max_pinf, double& il, double& iu) {
  const double tmin = a >= 0.0 ? a * lj : a * uj;
  const double tmax = a >= 0.0 ? a * uj : a * lj;
  const double rmin = residual_min(min_fin, min_ninf, tmin);
  const double rmax = residual_max(max_fin, max_pinf, tmax);
  il = -INFINITY;
  iu = INFINITY;
  if (a > 0.0) {
    if (isfinite(al) && isfinite(rmax)) il = (al - rmax) / a;
    if (isfinite(au) && isfinite(rmin)) iu = (au - rmin) / a;
  } else {
    if (isfinite(au) && isfinite(rmin)) il = (au - rmin) / a;
    if (isfinite(al) && isfinite(rmax)) iu = (al - rmax) / a;
  }
}

// Candidate bounds: G lanes per row, each entry independent; the per-column
// result is a total-order max / min, so the lane schedule does not matter.
// entry_lu holds the (l, u) pair of each used entry at the entry's position,
// stored by the activity sweep (which uses the same entries), so the bounds
// are read coalesced instead of gathered per column.
template <int G>
__global__ void __launch_bounds__(kBlock) k_bs_candidates(gp_csr A, const uint8_t* keep_row, Bits keep_col,
                                                           const int32_t* row_nnz, const double* AL, const double* AU,
                                                           const double2* entry_lu, const double* min_fin,
                                                           const double* max_fin, const int32_t* min_ninf,
                                                           const int32_t* max_pinf, double zero_tol, double* cl,
                                                           double* cu) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; i < A.rows; i += groups) {
    if (!keep_row[i] || row_nnz[i] <= 1) continue;
    const double al = AL[i], au = AU[i], mnf = min_fin[i], mxf = max_fin[i];
    const int32_t mni = min_ninf[i], mxp = max_pinf[i];
    const gp_offset p1 = A.row_ptr[i + 1];
    for (gp_offset p = A.row_ptr[i] + lane; p < p1; p += G) {
      const int32_t j = ld_stream(A.col_idx + p);
      const double a = ld_stream(A.vals + p);
      if (!keep_col[j]) continue;
      if (fabs(a) <= zero_tol) continue;
      const double2 b = entry_lu[p];
      const double lj = b.x, uj = b.y;
      double il, iu;
      implied_bounds(a, lj, uj, al, au, mnf, mni, mxf, mxp, il, iu);
      double nl = lj, nu = uj;
      if (isfinite(il)) nl = cmax(nl, il);
      if (isfinite(iu)) nu = cmin(nu, iu);
      if (nl > lj) atomic_total_max(cl + j, nl);
      if (nu < uj) atomic_total_min(cu + j, nu);
    }
  }
}

// max |a| over the live entries of each column, G lanes per column.
template <int G>
__global__ void __launch_bounds__(kBlock) k_col_max_abs(gp_csr AT, Bits keep_row, double* vmax) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t j = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; j < AT.rows; j += groups) {
    double v = 0.0;
    const gp_offset p1 = AT.row_ptr[j + 1];
    for (gp_offset p = AT.row_ptr[j] + lane; p < p1; p += G) {
      const int32_t i = ld_stream(AT.col_idx + p);
      const double a = ld_stream(AT.vals + p);
      if (!keep_row[i]) continue;
      v = cmax(v, fabs(a));
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v = cmax(v, __shfl_xor(v, o, G));
    if (lane == 0) vmax[j] = v;
  }
}

// Column max |a| kept with the plan: after the first sweep, only columns
// with an entry in a row whose liveness changed since the last call are
// recomputed (a row deletion can only lower a column's maximum).
struct ColMaxCache {
  int32_t m = -1, n = -1;
  bool valid = false, masked = false;
  DVec<double> vmax;
  DVec<uint8_t> keep_row;
};

// marks the columns of rows whose mask changed and brings the cached mask up to date
__global__ void k_cm_dirty(gp_csr A, const uint8_t* keep_row, uint8_t* keep_row_old, uint8_t* dirty) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  const uint8_t now = keep_row[i];
  if (now == keep_row_old[i]) return;
  keep_row_old[i] = now;
  for (gp_offset p = A.row_ptr[i]; p < A.row_ptr[i + 1]; ++p) dirty[A.col_idx[p]] = 1;
}

__global__ void k_cm_list(int32_t n, const uint8_t* dirty, int32_t* list, int32_t* count) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n && dirty[j]) list[atomicAdd(count, 1)] = static_cast<int32_t>(j);
}

__global__ void k_cm_refresh(gp_csr AT, Bits keep_row, const int32_t* list, int32_t count, double* vmax) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int32_t j = list[k];
  double v = 0.0;
  for (gp_offset p = AT.row_ptr[j]; p < AT.row_ptr[j + 1]; ++p) {
    if (!keep_row[AT.col_idx[p]]) continue;
    v = cmax(v, fabs(AT.vals[p]));
  }
  vmax[j] = v;
}

template <int G>
struct BsCandidatesK {
  static constexpr auto kernel = k_bs_candidates<G>;
};
template <int G>
struct ColMaxAbsK {
  static constexpr auto kernel = k_col_max_abs<G>;
};

__global__ void k_bs_finalize(int32_t n, const uint8_t* keep_col, const double* l, const double* u,
                              const double* cl_in, const double* cu_in, const double* ref_l, const double* ref_u,
                              const double* c, const double* vmax, int preserve, double zero_tol, double ft,
                              double relax, double* fl, double* fu, uint8_t* lch, uint8_t* uch, uint8_t* fixed,
                              double* fval, int32_t* flags, BatchMarks bm) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (!keep_col[j]) {
    fl[j] = l[j];
    fu[j] = u[j];
    lch[j] = uch[j] = fixed[j] = 0;
    fval[j] = 0.0;
    return;
  }
  const double lj = l[j], uj = u[j];
  double cand_l = cl_in[j], cand_u = cu_in[j];
  double new_l = lj, new_u = uj;
  uint8_t lc = 0, uc = 0, fx = 0;
  double fat = 0.0;
  double el = lj, eu = uj;
  if (isfinite(cand_l)) el = cmax(el, cand_l);
  if (isfinite(cand_u)) eu = cmin(eu, cand_u);
  if (el > eu + ft) {
    wave_flag(flags + 0, 1);
    bm.col(j, 1);
    return;
  }
  const bool was_free = !isfinite(ref_l[j]) && !isfinite(ref_u[j]);
  const bool box = isfinite(cand_l) && isfinite(cand_u);
  const bool zero_cost = fabs(c[j]) <= zero_tol;
  const bool nearly_fixed = box && (cand_u <= cand_l || (cand_u - cand_l) * vmax[j] <= ft);
  if (preserve && was_free && box && zero_cost && !nearly_fixed) {
    cand_l = lj;
    cand_u = uj;
  }
  if (isfinite(cand_l) && cand_l > lj && isfinite(uj)) {
    if (cand_l >= uj + ft) {
      wave_flag(flags + 0, 1);
      bm.col(j, 1);
      return;
    }
    if (cand_l >= uj || (uj - cand_l) * vmax[j] <= ft) {
      fx = 1;
      fat = uj;
      new_l = uj;
      new_u = uj;
    }
  }
  if (!fx && isfinite(cand_l) && cand_l > lj) {
    const bool tighten = !isfinite(lj) || ((cand_l - lj > ft * 1.0e4) && (cand_l - lj > 1.0e-2 * fabs(lj)));
    if (tighten) {
      if (cand_l != nearbyint(cand_l)) cand_l -= 0.5 * ft * fabs(cand_l);
      new_l = cand_l;
      lc = 1;
    }
  }
  if (!fx && isfinite(cand_u) && cand_u < uj && isfinite(new_l)) {
    if (cand_u <= new_l - ft) {
      wave_flag(flags + 0, 1);
      bm.col(j, 1);
      return;
    }
    if (cand_u <= new_l || (cand_u - new_l) * vmax[j] <= ft) {
      fx = 1;
      fat = new_l;
      new_u = new_l;
    }
  }
  if (!fx && isfinite(cand_u) && cand_u < uj) {
    const bool tighten = !isfinite(uj) || ((uj - cand_u > ft * 1.0e4) && (uj - cand_u > 1.0e-2 * fabs(uj)));
    if (tighten) {
      if (cand_u != nearbyint(cand_u)) cand_u += relax * ft * fabs(cand_u);
      new_u = cand_u;
      uc = 1;
    }
  }
  if (fx) {
    lc = uc = 0;
    wave_flag(flags + 2, 1);
  }
  if (lc || uc) wave_flag(flags + 1, 1);
  fl[j] = new_l;
  fu[j] = new_u;
  lch[j] = lc;
  uch[j] = uc;
  fixed[j] = fx;
  fval[j] = fat;
}

// Support row of each tightened bound: the smallest live row whose implied
// bound matches.  G lanes per column take partial minima (order-free).
template <int G>
__global__ void __launch_bounds__(kBlock) k_bs_support(
    gp_csr AT, const BsRow* rows, const double* l, const double* u, const uint8_t* lch, const uint8_t* uch,
    const double* fl, const double* fu, double zero_tol, double tol, double support_tol, int32_t* sl_out,
    int32_t* su_out, int64_t* r, int64_t* ci, int64_t* cv) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t j = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; j < AT.rows; j += groups) {
    int32_t sl = INT32_MAX, su = INT32_MAX;
    const bool need_l = lch[j], need_u = uch[j];
    if (need_l || need_u) {
      const double lj = l[j], uj = u[j], flj = fl[j], fuj = fu[j];
      const gp_offset p1 = AT.row_ptr[j + 1];
      for (gp_offset p = AT.row_ptr[j] + lane; p < p1; p += G) {
        const int32_t i = AT.col_idx[p];
        const double a = AT.vals[p];
        const BsRow R = rows[i];
        if (!R.live || R.row_nnz <= 1) continue;
        if (fabs(a) <= zero_tol) continue;
        double il, iu;
        implied_bounds(a, lj, uj, R.al, R.au, R.min_fin, R.min_ninf, R.max_fin, R.max_pinf, il, iu);
        if (need_l && flj > lj + tol && isfinite(il) && fabs(il - flj) <= support_tol) sl = i < sl ? i : sl;
        if (need_u && fuj < uj - tol && isfinite(iu) && fabs(iu - fuj) <= support_tol) su = i < su ? i : su;
      }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      const int32_t a = __shfl_xor(sl, o, G), b = __shfl_xor(su, o, G);
      sl = a < sl ? a : sl;
      su = b < su ? b : su;
    }
    if (lane == 0) {
      sl_out[j] = sl;
      su_out[j] = su;
      r[j] = (need_l ? 1 : 0) + (need_u ? 1 : 0);
      ci[j] = (need_l ? (sl == INT32_MAX ? 1 : 2) : 0) + (need_u ? (su == INT32_MAX ? 1 : 2) : 0);
      cv[j] = 4 * r[j];
    }
  }
}
template <int G>
struct BsSupportK {
  static constexpr auto kernel = k_bs_support<G>;
};

__global__ void k_bs_tape_pack(TapeView out, int32_t n, const uint8_t* lch, const uint8_t* uch, const int32_t* sl,
                               const int32_t* su, const double* l, const double* u, const double* fl,
                               const double* fu, const int64_t* r, const int64_t* ci, const int64_t* cv) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n || !(lch[j] || uch[j])) return;
  int64_t rec = out.rbase + r[j], ip = out.ibase + ci[j], vp = out.vbase + cv[j];
  double cur_l = l[j];
  const double cur_u = u[j];
  if (lch[j]) {
    out.types[rec] = sl[j] == INT32_MAX ? GP_RED_VARIABLE_BOUND_CHANGE : GP_RED_ROW_IMPLIED_BOUND_CHANGE;
    out.modes[rec] = GP_DUAL_MINIMAL;
    out.idx[ip++] = static_cast<int32_t>(j);
    if (sl[j] != INT32_MAX) out.idx[ip++] = sl[j];
    out.vals[vp++] = cur_l;
    out.vals[vp++] = cur_u;
    out.vals[vp++] = fl[j];
    out.vals[vp++] = cur_u;
    out.istarts[rec + 1] = static_cast<gp_offset>(ip);
    out.vstarts[rec + 1] = static_cast<gp_offset>(vp);
    cur_l = fl[j];
    ++rec;
  }
  if (uch[j]) {
    out.types[rec] = su[j] == INT32_MAX ? GP_RED_VARIABLE_BOUND_CHANGE : GP_RED_ROW_IMPLIED_BOUND_CHANGE;
    out.modes[rec] = GP_DUAL_MINIMAL;
    out.idx[ip++] = static_cast<int32_t>(j);
    if (su[j] != INT32_MAX) out.idx[ip++] = su[j];
    out.vals[vp++] = cur_l;
    out.vals[vp++] = cur_u;
    out.vals[vp++] = cur_l;
    out.vals[vp++] = fu[j];
    out.istarts[rec + 1] = static_cast<gp_offset>(ip);
    out.vstarts[rec + 1] = static_cast<gp_offset>(vp);
  }
}

__global__ void k_bs_apply(int32_t n, const double* fl, const double* fu, const uint8_t* fixed, double* l, double* u,
                           uint8_t* keep_col) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  l[j] = fl[j];
  u[j] = fu[j];
  if (fixed[j]) keep_col[j] = 0;
}

// fixed-column row shift and removal of rows left without live entries
__global__ void k_bs_fixed_rows(gp_csr A, const uint8_t* fixed, const double* fval, const uint8_t* keep_col,
                                double ft, uint8_t* keep_row, double* AL, double* AU, int32_t* flags, BatchMarks bm) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.rows || !keep_row[i]) return;
  RowAcc acc;
  int32_t live = 0;
  const gp_offset p0 = A.row_ptr[i];
  for (gp_offset p = p0; p < A.row_ptr[i + 1]; ++p) {
    acc.pos(p - p0);
    const int32_t j = A.col_idx[p];
    if (fixed[j]) acc.add(A.vals[p] * fval[j]);
    live += keep_col[j] != 0;
  }
  const double s = acc.value();
  const double al = AL[i] - s, au = AU[i] - s;
  AL[i] = al;
  AU[i] = au;
  if (live == 0) {
    if (al <= ft && au >= -ft) {
      keep_row[i] = 0;
    } else {
      wave_flag(flags + 3, 1);
      bm.row(i, 1);
    }
  }
}
// The same shift and removal as k_bs_fixed_rows through the LDS row fold:
// each column's bound pair is (fval, fval) and the fold's column mask is the
// fixed bits, so the staged term of a fixed entry is a * fval in CSR order
// with the same block positions.  A row's live count after the rule is its
// live count at rule start minus its fixed entries (fixed columns were live
// and are the only columns the rule removes).  Rows with no fixed entry keep
// AL / AU as they are (x - (+0) == x).
struct FixedRowsFold {
  const uint8_t* keep_row;
  const int32_t* row_nnz;
  double ft;
  uint8_t* keep_row_out;
  double* AL;
  double* AU;
  int32_t* flags;
  BatchMarks bm;
  struct State {
    RowAcc acc;
    int32_t fixed;
  };
  __device__ bool active(int64_t i) const { return keep_row[i] != 0; }
  __device__ bool use(double) const { return true; }
  __device__ void init(State& s) const {
    s.acc = RowAcc{};
    s.fixed = 0;
  }
  __device__ void add(State& s, int64_t k, double lo, double, bool skip) const {
    s.acc.pos(k);
    if (skip) return;
    s.acc.add(lo);
    ++s.fixed;
  }
  __device__ void finish(int64_t i, const State& s, bool active) const {
    if (!active) return;
    double al = AL[i], au = AU[i];
    if (s.fixed > 0) {
      const double sh = s.acc.value();
      al -= sh;
      au -= sh;
      AL[i] = al;
      AU[i] = au;
    }
    if (row_nnz[i] - s.fixed == 0) {
      if (al <= ft && au >= -ft) {
        keep_row_out[i] = 0;
      } else {
        wave_flag(flags + 3, 1);
        bm.row(i, 1);
      }
    }
  }
};
}  // namespace

void rule_bound_strengthening(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  const uint8_t* keep_col = s.live_masks ? plan.keep_col_mask : nullptr;
  DVec<double> min_fin(m), max_fin(m);
  DVec<int32_t> min_ninf(m), max_pinf(m);
  DVec<double> cl(n), cu(n), fl(n), fu(n), fval(n);
  DVec<uint8_t> lch(n), uch(n), fixed(n);
  // candidate bounds start as the current ones
  PackedBounds lu(plan.new_l, plan.new_u, n, cl.data(), cu.data());
  // the support-row search (tape only) reads each row's summary as one record
  DVec<BsRow> row_rec(tape_on(plan, p) && s.row_nnz ? static_cast<size_t>(m) : 0);
  // the sweep refreshes the entry-bounds store, and the candidates read each entry's pair from it
  EntryBounds* store = plan_store(plan);
  // without live masks every column counts as live: start the store over, so it covers them all
  if (!keep_col) store->drop();
  DVec<uint32_t> current;
  const StoreView view = store_view(store, lp.A, lu.data(), keep_col, p.zero_tol, true, current);
  row_activity_summary(lp.A, lu.data(), p.zero_tol, plan.keep_row_mask, keep_col, min_fin.data(), max_fin.data(),
                       min_ninf.data(), max_pinf.data(), row_rec.size() ? row_rec.data() : nullptr, plan.new_AL,
                       plan.new_AU, s.row_nnz, view);
  store_commit(store, lp.A, lu, p.zero_tol);  // lu's buffer becomes the snap
  const double2* entry_lu = store->pairs.data();
  {
    const PackedMask live_cols(keep_col, n);
    // A, the entries' (l, u) pairs, keep_row, row_nnz, AL, AU, the row summary (24 B), a mask bit
    // per column, candidate bounds cl / cu
    TimedScope ts(GP_KT_BOUND_CANDIDATES, 4.0 * (m + 1) + 28.0 * lp.A.nnz + 45.0 * m + 16.0 * n + n / 8.0);
    launch_grouped<BsCandidatesK>(pick_lanes(tuning().lanes_candidates, m ? static_cast<double>(lp.A.nnz) / m : 0.0), m, lp.A,
                                  plan.keep_row_mask, live_cols.view(), s.row_nnz, plan.new_AL, plan.new_AU,
                                  entry_lu, static_cast<const double*>(min_fin.data()),
                                  static_cast<const double*>(max_fin.data()),
                                  static_cast<const int32_t*>(min_ninf.data()),
                                  static_cast<const int32_t*>(max_pinf.data()), p.zero_tol, cl.data(), cu.data());
    GP_LAUNCH_CHECK("bound candidates");
  }
  Flags flags;
  {
    // outside the sweeps: candidates, bounds, cost and column max per column
    TimedScope outer(GP_KT_BOUNDS, 48.0 * n);
    const bool masked = s.live_masks != 0;
    const PackedMask live_rows(masked ? static_cast<const uint8_t*>(plan.keep_row_mask) : nullptr, m);
    RuleCaches& rc = rule_caches(plan);
    if (!rc.col_max) rc.col_max = new ColMaxCache();
    ColMaxCache& cm = *static_cast<ColMaxCache*>(rc.col_max);
    if (cm.m != m || cm.n != n || cm.masked != masked) {
      cm.valid = false;
      cm.m = m;
      cm.n = n;
      cm.masked = masked;
      cm.vmax.resize(n);
      cm.keep_row.resize(m);
    }
    bool full = !cm.valid, dirty_ran = false;
    if (!full && masked) {
      dirty_ran = true;
      DVec<uint8_t> dirty(n);
      DVec<int32_t> list(n);
      fill_u8(dirty.data(), 0, n);
      Flags cnt;
      hipLaunchKernelGGL(k_cm_dirty, dim3(grid_for(m)), dim3(kBlock), 0, 0, lp.A, plan.keep_row_mask,
                         cm.keep_row.data(), dirty.data());
      hipLaunchKernelGGL(k_cm_list, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, dirty.data(), list.data(), cnt.d.data());
      GP_LAUNCH_CHECK("col max dirty");
      const int32_t count = cnt.read()[0];
      if (static_cast<int64_t>(count) * 16 > n) {
        full = true;
      } else if (count > 0) {
        hipLaunchKernelGGL(k_cm_refresh, dim3(grid_for(count)), dim3(kBlock), 0, 0, lp.AT, live_rows.view(),
                           list.data(), count, cm.vmax.data());
        GP_LAUNCH_CHECK("col max refresh");
      }
    }
    if (full) {
      // AT row_ptr, col_idx, vals, a mask bit per row, 8 bytes out per column
      TimedScope ts(GP_KT_COL_MAX, 4.0 * (n + 1) + 12.0 * lp.AT.nnz + m / 8.0 + 8.0 * n);
      if (tuning().lanes_col_max == 0) {
        col_max_abs_tiles(lp.AT, live_rows.view(), cm.vmax.data());
      } else {
        launch_grouped<ColMaxAbsK>(tuning().lanes_col_max, n, lp.AT, live_rows.view(), cm.vmax.data());
        GP_LAUNCH_CHECK("col max abs");
      }
    }
    // k_cm_dirty brought the cached mask up to date when it ran
    if (masked && !dirty_ran) d2d(cm.keep_row.data(), plan.keep_row_mask, static_cast<size_t>(m));
    cm.valid = true;
    const double* vmax_d = cm.vmax.data();
    const double* ref_l = plan.propagation_reference_l ? plan.propagation_reference_l : (lp.l ? lp.l : plan.new_l);
    const double* ref_u = plan.propagation_reference_u ? plan.propagation_reference_u : (lp.u ? lp.u : plan.new_u);
    hipLaunchKernelGGL(k_bs_finalize, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, plan.keep_col_mask, plan.new_l,
                       plan.new_u, cl.data(), cu.data(), ref_l, ref_u, plan.new_c, vmax_d,
                       p.bound_strengthening_preserve_free_zero_cost ? 1 : 0, p.zero_tol, p.feasibility_tol,
                       p.bound_strengthening_upper_relaxation_factor, fl.data(), fu.data(), lch.data(), uch.data(),
                       fixed.data(), fval.data(), flags.d.data(), batch_marks());
    GP_LAUNCH_CHECK("bound finalize");
  }
  const int32_t* f = flags.read();
  if (f[0]) {
    plan.has_infeasible = 1;
    return;
  }
  if (!(f[1] || f[2])) {
    return;
  }
  const bool f2 = f[2] != 0;
  // support rows, records and fixings: the new bounds per column
  TimedScope ts(GP_KT_BOUNDS, 16.0 * n);
  if (tape_on(plan, p) && f[1]) {
    const double support_tol = cmax(10.0 * p.bound_tol, 1.0e-10);
    DVec<int32_t> sl(n), su(n);
    AppendCounts cnt(n);
    if (!row_rec.size()) throw Error(GP_ERR_INVALID_ARGUMENT, "bound strengthening tape needs live row counts");
    launch_grouped<BsSupportK>(lanes_for_mean(n ? static_cast<double>(lp.AT.nnz) / n : 0.0), n, lp.AT,
                               static_cast<const BsRow*>(row_rec.data()), static_cast<const double*>(plan.new_l),
                               static_cast<const double*>(plan.new_u), static_cast<const uint8_t*>(lch.data()),
                               static_cast<const uint8_t*>(uch.data()), static_cast<const double*>(fl.data()),
                               static_cast<const double*>(fu.data()), p.zero_tol, p.bound_tol, support_tol, sl.data(),
                               su.data(), cnt.r.data(), cnt.i.data(), cnt.v.data());
    GP_LAUNCH_CHECK("bound support rows");
    cnt.scan(n);
    if (cnt.R > 0) {
      DeviceTape* t = plan_tape(plan);
      t->reserve_more(cnt.R, cnt.I, cnt.V);
      TapeView v = tape_view(*t);
      hipLaunchKernelGGL(k_bs_tape_pack, dim3(grid_for(n)), dim3(kBlock), 0, 0, v, n, lch.data(), uch.data(),
                         sl.data(), su.data(), plan.new_l, plan.new_u, fl.data(), fu.data(), cnt.r.data(),
                         cnt.i.data(), cnt.v.data());
      GP_LAUNCH_CHECK("bound tape pack");
      t->commit(cnt.R, cnt.I, cnt.V);
    }
  }
  hipLaunchKernelGGL(k_bs_apply, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, fl.data(), fu.data(), fixed.data(),
                     plan.new_l, plan.new_u, plan.keep_col_mask);
  GP_LAUNCH_CHECK("bound apply");
  if (f2) {
    Flags f3;
    if (s.live_masks && s.row_nnz) {
      const PackedBounds fv(fval.data(), fval.data(), n);
      const PackedMask fixed_bits(fixed.data(), n);
      launch_row_fold(lp.A, fv.data(), fixed_bits.view(),
                      FixedRowsFold{plan.keep_row_mask, s.row_nnz, p.feasibility_tol, plan.keep_row_mask,
                                    plan.new_AL, plan.new_AU, f3.d.data(), batch_marks()});
    } else {
      hipLaunchKernelGGL(k_bs_fixed_rows, dim3(grid_for(m)), dim3(kBlock), 0, 0, lp.A, fixed.data(), fval.data(),
                         plan.keep_col_mask, p.feasibility_tol, plan.keep_row_mask, plan.new_AL, plan.new_AU,
                         f3.d.data(), batch_marks());
    }
    GP_LAUNCH_CHECK("bound fixed rows");
    if (f3.read()[3]) {
      plan.has_infeasible = 1;
      return;
    }
  }
  double obj = 0.0;
  if (f2)
    obj = fixed_variable_records(tape_on(plan, p) ? plan_tape(plan) : nullptr, fixed.data(), fval.data(),
                                 plan.keep_row_mask, plan.new_c, lp.AT, n, true, FixedObjective::Extended);
  plan.obj_constant_delta += obj;
  plan.has_row_action = plan.has_row_action || f2;
  plan.has_col_action = plan.has_col_action || f2;
  plan.has_change = 1;
}

RuleCaches& rule_caches(gp_plan& plan) {
  if (!plan.rule_cache) plan.rule_cache = new RuleCaches();
  return *static_cast<RuleCaches*>(plan.rule_cache);
}

EntryBounds* plan_store(gp_plan& plan) {
  RuleCaches& rc = rule_caches(plan);
  if (rc.store) return rc.store;
  if (!rc.own_store) rc.own_store = new EntryBounds();
  return rc.own_store;
}

void col_max_cache_free(void* cache) { delete static_cast<ColMaxCache*>(cache); }

void rule_cache_free(void* cache) {
  if (!cache) return;
  RuleCaches* rc = static_cast<RuleCaches*>(cache);
  sc_cache_free(rc->sc);
  col_max_cache_free(rc->col_max);
  delete rc->own_store;
  delete rc;
}

void apply_rule(int32_t rule, gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  switch (rule) {
    case GP_RULE_FIXED_VARIABLES: rule_fixed_variables(plan, lp, p); break;
    case GP_RULE_EMPTY_COLUMNS: rule_empty_columns(plan, lp, s, p); break;
    case GP_RULE_EMPTY_ROWS: rule_empty_rows(plan, lp, s, p); break;
    case GP_RULE_SINGLETON_ROWS: rule_singleton_rows(plan, lp, s, p); break;
    case GP_RULE_SINGLETON_COLUMNS: rule_singleton_columns(plan, lp, s, p); break;
    case GP_RULE_BOUND_STRENGTHENING: rule_bound_strengthening(plan, lp, s, p); break;
    case GP_RULE_REDUNDANT_ROWS: rule_redundant_rows(plan, lp, s, p); break;
    case GP_RULE_DUAL_FIX: rule_dual_fix(plan, lp, s, p); break;
    case GP_RULE_DOUBLETON_EQUATIONS: rule_doubleton_equations(plan, lp, s, p); break;
    case GP_RULE_PARALLEL_ROWS: rule_parallel_rows(plan, lp, s, p); break;
    case GP_RULE_PARALLEL_COLUMNS: rule_parallel_columns(plan, lp, s, p); break;
    default: throw std::invalid_argument("unknown rule id");
  }
}

}  // namespace gp
