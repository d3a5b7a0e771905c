// gp_startup.hip -- the structural Startup rules that share one
// deletion-only plan (gpu_presolve.cu:2361-2410):
//   column domination             rule_column_domination.cu:138-377
//   two-row variable elimination  rule_two_row_variable_elimination.cu:256-641
//   zero-cost product variables   rule_zero_cost_product_variables.cu:150-567
// and the FmeCol records of the projected columns (device_recovery.cu:409-479).
//
// Every rule is one thread per column: the certificates read a column's two or
// three incident rows of at most eight entries, so the work is a handful of
// gathers per candidate and an early exit for the rest.  Counts are
// wave-aggregated atomics; the tape is count -> scan -> pack.  Sums are formed
// in the order of the CPU restatement (row order of A, incidence order of
// A^T, no contraction), so decisions and records are bit-identical to it.
#include "gp_startup.hpp"

#include "gp_tape.hpp"

namespace gp {
namespace {

constexpr int kMaxCertRow = 8;
constexpr int kMaxAffine = kMaxCertRow - 1;
constexpr int kMaxDiff = 2 * kMaxAffine;

__device__ inline int64_t gtid() { return static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; }
__device__ inline bool pos_inf(double v) { return isinf(v) && v > 0.0; }
__device__ inline bool neg_inf(double v) { return isinf(v) && v < 0.0; }

// Column data the analyses read: the LP's matrix, the plan's vectors when the
// analysis runs against a pending plan (projection_view), the plan's masks.
struct ProjView {
  gp_csr A, AT;
  const double *c, *l, *u, *AL, *AU;
  const uint8_t* keep_row;
  const uint8_t* keep_col;
};

// ============================================================== column domination
enum : int32_t { kBadRowBounds = 1, kBadColDomain = 2, kBadObjective = 4, kBadValue = 8, kNoWitness = 16 };

__device__ inline bool positive_integer(double v) { return isfinite(v) && v >= 1.0 && floor(v) == v; }

__device__ inline int32_t cover_row_failure(const ProjView& v, int32_t i) {
  if (v.AL[i] != 1.0 || !pos_inf(v.AU[i])) return kBadRowBounds;
  for (gp_offset p = v.A.row_ptr[i]; p < v.A.row_ptr[i + 1]; ++p)
    if (v.c[v.A.col_idx[p]] == 1.0) return 0;
  return kNoWitness;
}
__device__ inline int32_t cover_col_failure(const ProjView& v, int32_t j) {
  int32_t f = 0;
  if (v.l[j] != 0.0 || v.u[j] != 1.0) f |= kBadColDomain;
  if (!positive_integer(v.c[j])) f |= kBadObjective;
  return f;
}

__device__ inline int64_t sampled(int64_t s, int64_t samples, int64_t count) { return s * count / samples; }

// quick_probe_rows / _cols / _values (rule_column_domination.cu:50-110) in one
// launch: sampled rows, columns, then entries of A and of A^T.
__global__ void k_cover_probe(ProjView v, int32_t rp, int32_t cp, int32_t np, int32_t* failure) {
  const int64_t t = gtid();
  int32_t f = 0;
  if (t < rp) {
    f = cover_row_failure(v, static_cast<int32_t>(sampled(t, rp, v.A.rows)));
  } else if (t < rp + cp) {
    f = cover_col_failure(v, static_cast<int32_t>(sampled(t - rp, cp, v.A.cols)));
  } else if (t < rp + cp + 2 * static_cast<int64_t>(np)) {
    const int64_t s = t - rp - cp;
    const int64_t p = sampled(s < np ? s : s - np, np, v.A.nnz);
    f = (s < np ? v.A.vals[p] : v.AT.vals[p]) != 1.0 ? kBadValue : 0;
  }
  if (f) atomicOr(failure, f);
}

// The full certificate (validate_rows_and_witnesses, validate_columns,
// validate_matrix_values) and mark_dominated_columns in one grid-stride pass:
// rows, then columns (c_j > degree marks a deletion and is counted), then
// entries of A and A^T.
__global__ void k_cover_check(ProjView v, uint8_t* del, int32_t* failure, int32_t* count) {
  const int64_t m = v.A.rows, n = v.A.cols, nnz = v.A.nnz;
  const int64_t total = m + n + nnz;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t t = gtid(); t < total; t += stride) {
    int32_t f = 0;
    if (t < m) {
      f = cover_row_failure(v, static_cast<int32_t>(t));
    } else if (t < m + n) {
      const int32_t j = static_cast<int32_t>(t - m);
      f = cover_col_failure(v, j);
      const int32_t degree = static_cast<int32_t>(v.AT.row_ptr[j + 1] - v.AT.row_ptr[j]);
      const bool d = v.c[j] > static_cast<double>(degree);
      del[j] = d ? 1 : 0;
      if (d) wave_count(count);
    } else {
      const int64_t p = t - m - n;
      if (v.A.vals[p] != 1.0 || v.AT.vals[p] != 1.0) f = kBadValue;
    }
    if (f) atomicOr(failure, f);
  }
}

// apply_dominated_columns (rule_column_domination.cu:204-216)
__global__ void k_cover_apply(int32_t n, const uint8_t* del, uint8_t* keep_col, double* l, double* u) {
  const int64_t j = gtid();
  if (j >= n || !del[j]) return;
  keep_col[j] = 0;
  l[j] = 0.0;
  u[j] = 0.0;
}

// ============================================================== projections
// projection_rows (projection_support.cuh:24-35): the K live rows of a column
// in A^T order; without a row mask the column must have exactly K entries.
template <int K>
__device__ inline bool projection_rows(const ProjView& v, int32_t col, int32_t* rows) {
  const gp_offset a = v.AT.row_ptr[col], b = v.AT.row_ptr[col + 1];
  if (!v.keep_row && b - a != K) return false;
  int size = 0;
  for (gp_offset q = a; q < b; ++q) {
    const int32_t r = v.AT.col_idx[q];
    if (v.keep_row && !v.keep_row[r]) continue;
    if (size == K) return false;
    rows[size++] = r;
  }
  return size == K;
}

struct Affine {
  double constant;
  int32_t length;
  int32_t cols[kMaxAffine];
  double coeffs[kMaxAffine];
};

__device__ inline bool add_term(int32_t col, double coeff, int32_t* cols, double* coeffs, int32_t& length, int cap) {
  if (!isfinite(coeff)) return false;
  for (int32_t k = 0; k < length; ++k)
    if (cols[k] == col) {
      coeffs[k] = coeffs[k] + coeff;
      return isfinite(coeffs[k]);
    }
  if (length >= cap) return false;
  cols[length] = col;
  coeffs[length] = coeff;
  ++length;
  return true;
}

// build_row_affine + classify_one_sided_row
// (rule_two_row_variable_elimination.cu:73-180)
__device__ bool row_affine(const ProjView& v, int32_t row, int32_t col, int32_t max_nnz, double zero_tol, Affine& out,
                           bool& gives_lower) {
  const gp_offset first = v.A.row_ptr[row], last = v.A.row_ptr[row + 1];
  int32_t nnz = static_cast<int32_t>(last - first);
  if (v.keep_col) {
    nnz = 0;
    for (gp_offset p = first; p < last; ++p)
      if (v.keep_col[v.A.col_idx[p]] && ++nnz > max_nnz) return false;
  }
  if (nnz < 2 || nnz > max_nnz || nnz > kMaxCertRow) return false;
  int pivots = 0;
  double pivot = 0.0;
  for (gp_offset p = first; p < last; ++p) {
    const int32_t j = v.A.col_idx[p];
    if (v.keep_col && !v.keep_col[j]) continue;
    if (j == col) {
      ++pivots;
      pivot = v.A.vals[p];
    }
  }
  if (pivots != 1) return false;
  if (!isfinite(pivot) || fabs(pivot) <= zero_tol) return false;
  double bound;
  if (isfinite(v.AL[row]) && pos_inf(v.AU[row])) {
    gives_lower = pivot > 0.0;
    bound = v.AL[row];
  } else if (neg_inf(v.AL[row]) && isfinite(v.AU[row])) {
    gives_lower = pivot < 0.0;
    bound = v.AU[row];
  } else {
    return false;
  }
  out.constant = bound / pivot;
  out.length = 0;
  if (!isfinite(out.constant)) return false;
  for (gp_offset p = first; p < last; ++p) {
    const int32_t j = v.A.col_idx[p];
    if (v.keep_col && !v.keep_col[j]) continue;
    if (j == col) continue;
    if (!add_term(j, -v.A.vals[p] / pivot, out.cols, out.coeffs, out.length, kMaxAffine)) return false;
  }
  return true;
}

// affine_interval (:182-208)
__device__ bool affine_interval(const Affine& f, const double* l, const double* u, double& lo_out, double& hi_out) {
  double lo = f.constant, hi = f.constant;
  for (int32_t k = 0; k < f.length; ++k) {
    const double lj = l[f.cols[k]], uj = u[f.cols[k]], q = f.coeffs[k];
    if (!isfinite(lj) || !isfinite(uj) || lj > uj) return false;
    lo = lo + (q >= 0.0 ? q * lj : q * uj);
    hi = hi + (q >= 0.0 ? q * uj : q * lj);
  }
  if (!isfinite(lo) || !isfinite(hi)) return false;
  lo_out = lo;
  hi_out = hi;
  return true;
}

// strictly_le_with_guard (:210-218)
__device__ inline bool le_guard(double lhs, double rhs, double tol) {
  if (!isfinite(lhs) || !isfinite(rhs)) return false;
  const double scale = 1.0 + fmax(fabs(lhs), fabs(rhs));
  return rhs - lhs >= tol * scale;
}

// difference_max (:220-254)
__device__ bool difference_max(const Affine& lo, const Affine& up, const double* l, const double* u, double& out) {
  int32_t cols[kMaxDiff];
  double coeffs[kMaxDiff];
  int32_t len = 0;
  for (int32_t k = 0; k < lo.length; ++k)
    if (!add_term(lo.cols[k], lo.coeffs[k], cols, coeffs, len, kMaxDiff)) return false;
  for (int32_t k = 0; k < up.length; ++k)
    if (!add_term(up.cols[k], -up.coeffs[k], cols, coeffs, len, kMaxDiff)) return false;
  double val = lo.constant - up.constant;
  for (int32_t k = 0; k < len; ++k) {
    const double lj = l[cols[k]], uj = u[cols[k]];
    if (!isfinite(lj) || !isfinite(uj) || lj > uj) return false;
    val = val + (coeffs[k] >= 0.0 ? coeffs[k] * uj : coeffs[k] * lj);
  }
  if (!isfinite(val)) return false;
  out = val;
  return true;
}

// detect_bounded_two_row_candidates (:256-343)
__device__ bool two_row_candidate(const ProjView& v, int32_t col, int32_t max_nnz, double zero_tol, double tol,
                                  int32_t* rows) {
  const double lj = v.l[col], uj = v.u[col];
  if (v.c[col] != 0.0 || !isfinite(lj) || !isfinite(uj) || !le_guard(lj, uj, tol) ||
      !projection_rows<2>(v, col, rows))
    return false;
  if (rows[0] == rows[1]) return false;
  Affine f0, f1;
  bool lower0 = false, lower1 = false;
  if (!row_affine(v, rows[0], col, max_nnz, zero_tol, f0, lower0) ||
      !row_affine(v, rows[1], col, max_nnz, zero_tol, f1, lower1) || lower0 == lower1)
    return false;
  const Affine& lo = lower0 ? f0 : f1;
  const Affine& up = lower0 ? f1 : f0;
  double lo_min, lo_max, up_min, up_max, diff;
  if (!affine_interval(lo, v.l, v.u, lo_min, lo_max) || !affine_interval(up, v.l, v.u, up_min, up_max)) return false;
  if (!le_guard(lo_max, uj, tol) || !le_guard(lj, up_min, tol)) return false;
  if (!difference_max(lo, up, v.l, v.u, diff) || !le_guard(diff, 0.0, tol)) return false;
  return true;
}

struct McRow {
  bool valid, gives_lower;
  int32_t e0, e1;
};

// match_mccormick_row (rule_zero_cost_product_variables.cu:73-148)
__device__ McRow mccormick_row(const ProjView& v, int32_t row, int32_t z) {
  McRow r{false, false, -1, -1};
  const gp_offset first = v.A.row_ptr[row], last = v.A.row_ptr[row + 1];
  int32_t nnz = static_cast<int32_t>(last - first);
  if (v.keep_col) {
    nnz = 0;
    for (gp_offset p = first; p < last; ++p)
      if (v.keep_col[v.A.col_idx[p]] && ++nnz > 3) return r;
  }
  if (nnz != 2 && nnz != 3) return r;
  double pivot = 0.0;
  int pivots = 0;
  for (gp_offset p = first; p < last; ++p) {
    const int32_t j = v.A.col_idx[p];
    if (v.keep_col && !v.keep_col[j]) continue;
    if (j == z) {
      pivot = v.A.vals[p];
      ++pivots;
    }
  }
  if (pivots != 1 || !isfinite(pivot) || pivot == 0.0) return r;
  double bound;
  if (isfinite(v.AL[row]) && pos_inf(v.AU[row])) {
    r.gives_lower = pivot > 0.0;
    bound = v.AL[row];
  } else if (neg_inf(v.AL[row]) && isfinite(v.AU[row])) {
    r.gives_lower = pivot < 0.0;
    bound = v.AU[row];
  } else {
    return r;
  }
  const int32_t want_nnz = r.gives_lower ? 3 : 2;
  const double want_bound = r.gives_lower ? -pivot : 0.0;
  if (nnz != want_nnz || bound != want_bound) return r;
  int32_t ends = 0;
  for (gp_offset p = first; p < last; ++p) {
    const int32_t j = v.A.col_idx[p];
    if (v.keep_col && !v.keep_col[j]) continue;
    if (j == z) continue;
    if (v.A.vals[p] != -pivot || ends >= 2) return r;
    if (ends == 0)
      r.e0 = j;
    else
      r.e1 = j;
    ++ends;
  }
  if (ends != (r.gives_lower ? 2 : 1) || (r.gives_lower && r.e0 == r.e1)) return r;
  r.valid = true;
  return r;
}

// detect_orphan_mccormick_candidates (:150-240)
__device__ bool mccormick_candidate(const ProjView& v, int32_t z, int32_t* rows, int32_t& ex, int32_t& ey) {
  if (v.c[z] != 0.0 || v.l[z] != 0.0 || v.u[z] != 1.0 || !projection_rows<3>(v, z, rows)) return false;
  if (rows[0] == rows[1] || rows[0] == rows[2] || rows[1] == rows[2]) return false;
  McRow pat[3];
  for (int k = 0; k < 3; ++k) pat[k] = mccormick_row(v, rows[k], z);
  int lower = -1, ups[2] = {-1, -1}, nup = 0;
  for (int k = 0; k < 3; ++k) {
    if (!pat[k].valid) return false;
    if (pat[k].gives_lower) {
      if (lower >= 0) return false;
      lower = k;
    } else {
      if (nup >= 2) return false;
      ups[nup++] = k;
    }
  }
  if (lower < 0 || nup != 2) return false;
  const int32_t x = pat[lower].e0, y = pat[lower].e1;
  const int32_t ux = pat[ups[0]].e0, uy = pat[ups[1]].e0;
  if (x < 0 || y < 0 || x == y || !((ux == x && uy == y) || (ux == y && uy == x)) || v.l[x] != 0.0 ||
      v.u[x] != 1.0 || v.l[y] != 0.0 || v.u[y] != 1.0)
    return false;
  ex = x;
  ey = y;
  return true;
}

template <int K>
__global__ void k_proj_detect(ProjView v, int32_t max_nnz, double zero_tol, double tol, uint8_t* cand, int32_t* rows,
                              int32_t* ex, int32_t* ey, int32_t* live) {
  const int64_t j64 = gtid();
  const int32_t n = v.A.cols;
  if (j64 >= n) return;
  const int32_t j = static_cast<int32_t>(j64);
  cand[j] = 0;
  if (K == 3) ex[j] = ey[j] = -1;
  if (v.keep_col && !v.keep_col[j]) return;
  if (v.keep_col) wave_count(live);
  int32_t* rj = rows + static_cast<int64_t>(j) * K;
  bool ok;
  if (K == 2) {
    ok = two_row_candidate(v, j, max_nnz, zero_tol, tol, rj);
  } else {
    int32_t x = -1, y = -1;
    ok = mccormick_candidate(v, j, rj, x, y);
    ex[j] = x;
    ey[j] = y;
  }
  cand[j] = ok ? 1 : 0;
}

// count_candidate_rows (:345-357)
template <int K>
__global__ void k_proj_rows(int32_t n, const uint8_t* cand, const int32_t* rows, int32_t* row_count) {
  const int64_t j = gtid();
  if (j >= n || !cand[j]) return;
  for (int k = 0; k < K; ++k) atomicAdd(&row_count[rows[j * K + k]], 1);
}

// select_nonconflicting_candidates / select_orphan_mccormick_candidates: every
// row claimed by this candidate only; a product's endpoints are no candidates.
template <int K>
__global__ void k_proj_select(int32_t n, const uint8_t* cand, const int32_t* rows, const int32_t* ex,
                              const int32_t* ey, const int32_t* row_count, uint8_t* sel, int32_t* selected) {
  const int64_t j = gtid();
  if (j >= n) return;
  bool ok = cand[j] != 0;
  if (ok) {
    for (int k = 0; k < K; ++k) ok = ok && row_count[rows[j * K + k]] == 1;
    if (K == 3) ok = ok && !cand[ex[j]] && !cand[ey[j]];
  }
  sel[j] = ok ? 1 : 0;
  if (ok) wave_count(selected);
}

// lower_row (device_recovery.cu:409-431): the first incident row bounding the
// column from below, with its pivot and side.
template <int K>
__device__ int32_t lower_row(const ProjView& v, int32_t col, const int32_t* rows, double& pivot, double& bound) {
  for (int k = 0; k < K; ++k) {
    const int32_t r = rows[k];
    double a = 0.0;
    int seen = 0;
    for (gp_offset p = v.A.row_ptr[r]; p < v.A.row_ptr[r + 1]; ++p)
      if (v.A.col_idx[p] == col) {
        a = v.A.vals[p];
        ++seen;
      }
    if (seen != 1 || !isfinite(a) || a == 0.0) return -1;
    const bool lower = isfinite(v.AL[r]) && pos_inf(v.AU[r]) && a > 0.0;
    const bool upper = neg_inf(v.AL[r]) && isfinite(v.AU[r]) && a < 0.0;
    if (lower || upper) {
      pivot = a;
      bound = lower ? v.AL[r] : v.AU[r];
      return r;
    }
  }
  return -1;
}

// projection_counts (device_recovery.cu:432-452): one record per selected
// column with K + 8 + N indices and 2 + N values, N = live support.
template <int K>
__global__ void k_proj_counts(ProjView v, const uint8_t* sel, const int32_t* rows, int64_t* r, int64_t* ni,
                              int64_t* nv, int32_t* error) {
  const int64_t j = gtid();
  if (j >= v.A.cols) return;
  r[j] = ni[j] = nv[j] = 0;
  if (!sel[j]) return;
  double pivot = 0.0, bound = 0.0;
  const int32_t row = lower_row<K>(v, static_cast<int32_t>(j), rows + j * K, pivot, bound);
  if (row < 0) {
    atomicOr(error, 1);
    return;
  }
  int64_t N = 0;
  for (gp_offset p = v.A.row_ptr[row]; p < v.A.row_ptr[row + 1]; ++p) {
    const int32_t c = v.A.col_idx[p];
    N += c != j && (!v.keep_col || v.keep_col[c]);
  }
  if (K == 3 && N != 2) atomicOr(error, 1);
  r[j] = 1;
  ni[j] = K + 8 + N;
  nv[j] = 2 + N;
}

// projection_pack (device_recovery.cu:453-479): indices [K, col, K rows, N,
// 2, 1, 0, 1, N, support], values [lower (two rows) or 0, bound / pivot,
// -a / pivot ...].
template <int K>
__global__ void k_proj_pack(TapeView out, ProjView v, const uint8_t* sel, const int32_t* rows, const int64_t* r,
                            const int64_t* ni, const int64_t* nv) {
  const int64_t j = gtid();
  if (j >= v.A.cols || !sel[j]) return;
  double pivot = 0.0, bound = 0.0;
  const int32_t row = lower_row<K>(v, static_cast<int32_t>(j), rows + j * K, pivot, bound);
  if (row < 0) return;  // reported by the counts
  const int64_t rec = out.rbase + r[j];
  int32_t* ix = out.idx + out.ibase + ni[j];
  double* va = out.vals + out.vbase + nv[j];
  int32_t N = 0;
  for (gp_offset p = v.A.row_ptr[row]; p < v.A.row_ptr[row + 1]; ++p) {
    const int32_t c = v.A.col_idx[p];
    if (c != j && (!v.keep_col || v.keep_col[c])) {
      ix[K + 8 + N] = c;
      va[2 + N] = -v.A.vals[p] / pivot;
      ++N;
    }
  }
  ix[0] = K;
  ix[1] = static_cast<int32_t>(j);
  for (int k = 0; k < K; ++k) ix[2 + k] = rows[j * K + k];
  ix[2 + K] = N;
  ix[3 + K] = 2;
  ix[4 + K] = 1;
  ix[5 + K] = 0;
  ix[6 + K] = 1;
  ix[7 + K] = N;
  va[0] = K == 2 ? v.l[j] : 0.0;
  va[1] = bound / pivot;
  out.types[rec] = GP_RED_FME_COL;
  out.modes[rec] = GP_DUAL_MINIMAL;
  out.istarts[rec + 1] = static_cast<gp_offset>(out.ibase + ni[j] + K + 8 + N);
  out.vstarts[rec + 1] = static_cast<gp_offset>(out.vbase + nv[j] + 2 + N);
}

// apply_two_row_variable_elimination / apply_zero_cost_product_variables
template <int K>
__global__ void k_proj_apply(int32_t n, const uint8_t* sel, const int32_t* rows, uint8_t* keep_row,
                             uint8_t* keep_col) {
  const int64_t j = gtid();
  if (j >= n || !sel[j]) return;
  keep_col[j] = 0;
  for (int k = 0; k < K; ++k) keep_row[rows[j * K + k]] = 0;
}

bool lp_usable(const gp_device_lp& lp) {
  return lp.A.rows > 0 && lp.A.cols > 0 && lp.A.nnz > 0 && lp.AT.rows == lp.A.cols && lp.AT.cols == lp.A.rows &&
         lp.AT.nnz == lp.A.nnz && lp.A.row_ptr && lp.A.col_idx && lp.A.vals && lp.AT.row_ptr && lp.AT.col_idx &&
         lp.AT.vals && lp.c && lp.AL && lp.AU && lp.l && lp.u;
}

// projection_view (projection_support.cuh:11-21)
ProjView make_view(const gp_device_lp& lp, const gp_plan* plan, bool masks) {
  ProjView v{lp.A, lp.AT, lp.c, lp.l, lp.u, lp.AL, lp.AU, nullptr, nullptr};
  if (plan) {
    if (plan->new_c) v.c = plan->new_c;
    if (plan->new_l) v.l = plan->new_l;
    if (plan->new_u) v.u = plan->new_u;
    if (plan->new_AL) v.AL = plan->new_AL;
    if (plan->new_AU) v.AU = plan->new_AU;
    if (masks) {
      v.keep_row = plan->keep_row_mask;
      v.keep_col = plan->keep_col_mask;
    }
  }
  return v;
}

StartupOutcome column_domination(const gp_device_lp& lp, const gp_params& p, const PlanSource& get_plan) {
  StartupOutcome out;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  const int32_t min_fixed = p.column_domination_min_fixed_cols > 1 ? p.column_domination_min_fixed_cols : 1;
  if (!p.enable_column_domination || !lp_usable(lp) || n < min_fixed) return out;
  const ProjView v = make_view(lp, nullptr, false);
  Flags f;  // [0] probe failure, [1] full failure, [2] dominated count
  {
    // quick_probe_column_domination (rule_column_domination.cu:218-275)
    auto clampi = [](int64_t a, int64_t b) { return a < b ? a : b; };
    const int32_t rp = static_cast<int32_t>(clampi(m, p.column_domination_probe_rows) > 1 ? clampi(m, p.column_domination_probe_rows) : 1);
    const int32_t cp = static_cast<int32_t>(clampi(n, p.column_domination_probe_cols) > 1 ? clampi(n, p.column_domination_probe_cols) : 1);
    const int32_t np = static_cast<int32_t>(clampi(lp.A.nnz, p.column_domination_probe_nnz) > 1 ? clampi(lp.A.nnz, p.column_domination_probe_nnz) : 1);
    const int64_t threads = static_cast<int64_t>(rp) + cp + 2 * static_cast<int64_t>(np);
    hipLaunchKernelGGL(k_cover_probe, dim3(grid_for(threads)), dim3(kBlock), 0, 0, v, rp, cp, np, f.d.data());
    GP_LAUNCH_CHECK("covering probe");
    if (f.read()[0] != 0) return out;
  }
  DVec<uint8_t> del(static_cast<size_t>(n));
  const int64_t total = static_cast<int64_t>(m) + n + lp.A.nnz;
  hipLaunchKernelGGL(k_cover_check, dim3(grid_capped(total)), dim3(kBlock), 0, 0, v, del.data(), f.d.data() + 1,
                     f.d.data() + 2);
  GP_LAUNCH_CHECK("covering certificate");
  const int32_t* h = f.read();
  if (h[1] != 0) return out;
  const int32_t k = h[2];
  const double ratio = static_cast<double>(k) / static_cast<double>(n);
  if (k < min_fixed || ratio + 1.0e-15 < p.column_domination_min_candidate_ratio) return out;
  out.applicable = true;
  out.count = k;
  gp_plan* plan = get_plan();
  if (!plan || !plan->keep_col_mask || !plan->new_l || !plan->new_u) return out;
  hipLaunchKernelGGL(k_cover_apply, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, del.data(), plan->keep_col_mask,
                     plan->new_l, plan->new_u);
  GP_LAUNCH_CHECK("covering apply");
  plan->has_change = plan->has_col_action = 1;
  out.primal_only = GP_PRIMAL_ONLY_COLUMN_DOMINATION;
  return out;
}

template <int K>
StartupOutcome projection(const gp_device_lp& lp, const gp_params& p, gp_plan* pending, const PlanSource& get_plan) {
  StartupOutcome out;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  const int enabled = K == 2 ? p.enable_two_row_variable_elimination : p.enable_zero_cost_product_variables;
  if (!enabled || !lp_usable(lp)) return out;
  const ProjView v = make_view(lp, pending, pending != nullptr);
  int32_t max_nnz = p.two_row_variable_elimination_max_row_nnz;
  if (max_nnz > kMaxCertRow) max_nnz = kMaxCertRow;
  if (max_nnz < 2) max_nnz = 2;
  double tol = 10.0 * p.feasibility_tol;
  if (100.0 * p.zero_tol > tol) tol = 100.0 * p.zero_tol;
  if (p.two_row_variable_elimination_safety_tol > tol) tol = p.two_row_variable_elimination_safety_tol;

  DVec<uint8_t> cand(static_cast<size_t>(n)), sel(static_cast<size_t>(n));
  DVec<int32_t> rows(static_cast<size_t>(n) * K), ex(K == 3 ? static_cast<size_t>(n) : 1),
      ey(K == 3 ? static_cast<size_t>(n) : 1), row_count(static_cast<size_t>(m));
  Flags f;  // [0] selected, [1] live columns, [2] record certificate error
  GP_CHECK(hipMemsetAsync(row_count.data(), 0, sizeof(int32_t) * static_cast<size_t>(m), 0));
  hipLaunchKernelGGL(k_proj_detect<K>, dim3(grid_for(n)), dim3(kBlock), 0, 0, v, max_nnz, p.zero_tol, tol,
                     cand.data(), rows.data(), ex.data(), ey.data(), f.d.data() + 1);
  GP_LAUNCH_CHECK("projection candidates");
  hipLaunchKernelGGL(k_proj_rows<K>, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, cand.data(), rows.data(),
                     row_count.data());
  GP_LAUNCH_CHECK("projection rows");
  hipLaunchKernelGGL(k_proj_select<K>, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, cand.data(), rows.data(),
                     ex.data(), ey.data(), row_count.data(), sel.data(), f.d.data());
  GP_LAUNCH_CHECK("projection select");
  const int32_t* h = f.read();
  const int32_t selected = h[0], live = h[1];
  const int32_t min_c = K == 2 ? p.two_row_variable_elimination_min_candidates
                               : p.zero_cost_product_variables_min_candidates;
  const double min_r = K == 2 ? p.two_row_variable_elimination_min_candidate_ratio
                              : p.zero_cost_product_variables_min_candidate_ratio;
  const double denom = pending ? static_cast<double>(live > 1 ? live : 1) : static_cast<double>(n);
  if (!(selected >= (min_c > 1 ? min_c : 1) && static_cast<double>(selected) / denom + 1.0e-15 >= min_r)) return out;
  out.applicable = true;
  out.count = selected;
  gp_plan* plan = get_plan();
  if (!plan || !plan->keep_row_mask || !plan->keep_col_mask) return out;
  DeviceTape* tape = plan_tape(*plan);
  if (tape && p.record_postsolve_tape) {
    // records read the plan's vectors and its current column mask
    ProjView bv = make_view(lp, plan, false);
    bv.keep_col = plan->keep_col_mask;
    AppendCounts ac(n);
    hipLaunchKernelGGL(k_proj_counts<K>, dim3(grid_for(n)), dim3(kBlock), 0, 0, bv, sel.data(), rows.data(),
                       ac.r.data(), ac.i.data(), ac.v.data(), f.d.data() + 2);
    GP_LAUNCH_CHECK("projection record counts");
    ac.scan(n);
    if (read1(f.d.data() + 2) != 0) throw std::runtime_error("projection tape certificate changed before commit");
    if (ac.R > 0) {
      tape->reserve_more(ac.R, ac.I, ac.V);
      hipLaunchKernelGGL(k_proj_pack<K>, dim3(grid_for(n)), dim3(kBlock), 0, 0, tape_view(*tape), bv, sel.data(),
                         rows.data(), ac.r.data(), ac.i.data(), ac.v.data());
      GP_LAUNCH_CHECK("projection record pack");
      tape->commit(ac.R, ac.I, ac.V);
    }
  }
  hipLaunchKernelGGL(k_proj_apply<K>, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, sel.data(), rows.data(),
                     plan->keep_row_mask, plan->keep_col_mask);
  GP_LAUNCH_CHECK("projection apply");
  plan->has_change = plan->has_row_action = plan->has_col_action = 1;
  out.primal_only = GP_PRIMAL_ONLY_PROJECTED_AUXILIARY;
  return out;
}

}  // namespace

StartupOutcome startup_rule(int32_t rule, const gp_device_lp& lp, const gp_params& p, gp_plan* pending,
                            const PlanSource& get_plan) {
  TimedScope ts(GP_KT_RULES_OTHER, 0.0);
  switch (rule) {
    case GP_STARTUP_COLUMN_DOMINATION:
      return column_domination(lp, p, get_plan);
    case GP_STARTUP_TWO_ROW_ELIMINATION:
      return projection<2>(lp, p, pending, get_plan);
    case GP_STARTUP_ZERO_COST_PRODUCTS:
      return projection<3>(lp, p, pending, get_plan);
    default:
      throw std::invalid_argument("unknown Startup rule");
  }
}

}  // namespace gp

using namespace gp;

extern "C" int gp_startup_rule(int32_t rule, const gp_device_lp* lp, const gp_params* params, gp_plan* plan,
                               int32_t pending, int32_t* candidate_count, int32_t* applicable) {
  return guarded([&] {
    if (!lp || !params) throw std::invalid_argument("null argument");
    if (pending && !plan) throw std::invalid_argument("pending analysis needs a plan");
    if (rule < GP_STARTUP_COLUMN_DOMINATION || rule > GP_STARTUP_ZERO_COST_PRODUCTS)
      throw std::invalid_argument("unknown Startup rule");
    const StartupOutcome o = startup_rule(rule, *lp, *params, pending ? plan : nullptr, [plan] { return plan; });
    GP_CHECK(hipDeviceSynchronize());
    if (candidate_count) *candidate_count = o.count;
    if (applicable) *applicable = o.applicable ? 1 : 0;
  });
}
