/*
 * oracle_startup.c -- CPU restatement of the structural Startup rules that
 * share one deletion-only plan (gpu_presolve.cu:2361-2410):
 *   column domination            rule_column_domination.cu:138-377
 *   two-row variable elimination rule_two_row_variable_elimination.cu:64-641
 *   zero-cost product variables  rule_zero_cost_product_variables.cu:60-567
 * and the FmeCol records they write (device_recovery.cu:409-479,817-844).
 * TEST INFRASTRUCTURE ONLY.
 *
 * Every decision is a comparison of input values or of sums formed in the
 * order written here (row order of A, incidence order of A^T), so the HIP
 * engine reproduces masks, counts and records bit for bit.
 */
#include "oracle_internal.h"

#include <stdlib.h>
#include <string.h>

#define OR_MAX_CERT_ROW 8
#define OR_MAX_AFFINE (OR_MAX_CERT_ROW - 1)
#define OR_MAX_DIFF (2 * OR_MAX_AFFINE)

static int pos_inf(double v) { return isinf(v) && v > 0.0; }
static int neg_inf(double v) { return isinf(v) && v < 0.0; }

static int lp_usable(const gp_device_lp* lp) {
  return lp->A.rows > 0 && lp->A.cols > 0 && lp->A.nnz > 0 && lp->AT.rows == lp->A.cols &&
         lp->AT.cols == lp->A.rows && lp->AT.nnz == lp->A.nnz && lp->A.row_ptr && lp->A.col_idx &&
         lp->A.vals && lp->AT.row_ptr && lp->AT.col_idx && lp->c && lp->AL && lp->AU && lp->l && lp->u;
}

/* ---- column domination (rule_column_domination.cu:138-377) --------------
 * Unit set covering: every row 1 <= a x, every entry 1, every column a
 * [0, 1] box with a positive integer cost, every row holding a cost-one
 * column.  A column whose cost exceeds its degree is then fixed to 0.  The
 * reference first checks a sample of rows, columns and entries; every sample
 * check is one of the full checks, so the decision is the full check's. */
static int covering_certificate(const gp_device_lp* lp) {
  const gp_csr* A = &lp->A;
  for (int32_t i = 0; i < A->rows; ++i) {
    if (lp->AL[i] != 1.0 || !pos_inf(lp->AU[i])) return 0;
    int witness = 0;
    for (gp_offset p = A->row_ptr[i]; p < A->row_ptr[i + 1]; ++p)
      if (lp->c[A->col_idx[p]] == 1.0) {
        witness = 1;
        break;
      }
    if (!witness) return 0;
  }
  for (int32_t j = 0; j < A->cols; ++j) {
    const double c = lp->c[j];
    if (lp->l[j] != 0.0 || lp->u[j] != 1.0) return 0;
    if (!(isfinite(c) && c >= 1.0 && floor(c) == c)) return 0;
  }
  for (gp_offset p = 0; p < A->nnz; ++p)
    if (A->vals[p] != 1.0 || lp->AT.vals[p] != 1.0) return 0;
  return 1;
}

static int column_domination(const gp_device_lp* lp, const gp_params* p, gp_plan* plan, int32_t* count) {
  const int32_t n = lp->A.cols;
  const int32_t min_fixed = p->column_domination_min_fixed_cols > 1 ? p->column_domination_min_fixed_cols : 1;
  *count = 0;
  if (!p->enable_column_domination || !lp_usable(lp) || n < min_fixed) return 0;
  if (!covering_certificate(lp)) return 0;
  uint8_t* del = (uint8_t*)or_xcalloc((size_t)n, 1);
  int32_t k = 0;
  for (int32_t j = 0; j < n; ++j) {
    const int32_t degree = (int32_t)(lp->AT.row_ptr[j + 1] - lp->AT.row_ptr[j]);
    del[j] = lp->c[j] > (double)degree;
    k += del[j];
  }
  const double ratio = (double)k / (double)n;
  const int ok = k >= min_fixed && !(ratio + 1.0e-15 < p->column_domination_min_candidate_ratio);
  if (ok && plan && plan->keep_col_mask && plan->new_l && plan->new_u) {
    for (int32_t j = 0; j < n; ++j)
      if (del[j]) {
        plan->keep_col_mask[j] = 0;
        plan->new_l[j] = 0.0;
        plan->new_u[j] = 0.0;
      }
    plan->has_change = plan->has_col_action = 1;
  }
  free(del);
  *count = ok ? k : 0;
  return ok;
}

/* ---- projection support (projection_support.cuh:11-35) ------------------- */

/* The K live rows of column `col`, in A^T order; without a row mask the
 * column must have exactly K entries. */
static int projection_rows(const gp_csr* AT, const uint8_t* keep_row, int32_t col, int K, int32_t* rows) {
  const gp_offset a = AT->row_ptr[col], b = AT->row_ptr[col + 1];
  if (!keep_row && b - a != K) return 0;
  int size = 0;
  for (gp_offset q = a; q < b; ++q) {
    if (keep_row && !keep_row[AT->col_idx[q]]) continue;
    if (size == K) return 0;
    rows[size++] = AT->col_idx[q];
  }
  return size == K;
}

typedef struct or_affine {
  double constant;
  int32_t length;
  int32_t cols[OR_MAX_AFFINE];
  double coeffs[OR_MAX_AFFINE];
} or_affine;

static int add_term(int32_t col, double coeff, int32_t* cols, double* coeffs, int32_t* length, int32_t cap) {
  if (!isfinite(coeff)) return 0;
  for (int32_t k = 0; k < *length; ++k)
    if (cols[k] == col) {
      coeffs[k] += coeff;
      return isfinite(coeffs[k]);
    }
  if (*length >= cap) return 0;
  cols[*length] = col;
  coeffs[*length] = coeff;
  ++*length;
  return 1;
}

/* build_row_affine (rule_two_row_variable_elimination.cu:123-180): the one
 * side of a one-sided row solved for the eliminated column,
 * x_col >= or <= bound / pivot - sum_k (a_k / pivot) x_k. */
static int row_affine(const gp_device_lp* lp, const uint8_t* keep_col, int32_t row, int32_t col, int32_t max_nnz,
                      double zero_tol, or_affine* out, int* gives_lower) {
  const gp_csr* A = &lp->A;
  const gp_offset first = A->row_ptr[row], last = A->row_ptr[row + 1];
  int32_t nnz = (int32_t)(last - first);
  if (keep_col) {
    nnz = 0;
    for (gp_offset p = first; p < last; ++p)
      if (keep_col[A->col_idx[p]] && ++nnz > max_nnz) return 0;
  }
  if (nnz < 2 || nnz > max_nnz || nnz > OR_MAX_CERT_ROW) return 0;
  int pivots = 0;
  double pivot = 0.0;
  for (gp_offset p = first; p < last; ++p) {
    if (keep_col && !keep_col[A->col_idx[p]]) continue;
    if (A->col_idx[p] == col) {
      ++pivots;
      pivot = A->vals[p];
    }
  }
  if (pivots != 1) return 0;
  /* classify_one_sided_row (:73-94) */
  if (!isfinite(pivot) || fabs(pivot) <= zero_tol) return 0;
  double bound;
  if (isfinite(lp->AL[row]) && pos_inf(lp->AU[row])) {
    *gives_lower = pivot > 0.0;
    bound = lp->AL[row];
  } else if (neg_inf(lp->AL[row]) && isfinite(lp->AU[row])) {
    *gives_lower = pivot < 0.0;
    bound = lp->AU[row];
  } else {
    return 0;
  }
  out->constant = bound / pivot;
  out->length = 0;
  if (!isfinite(out->constant)) return 0;
  for (gp_offset p = first; p < last; ++p) {
    if (keep_col && !keep_col[A->col_idx[p]]) continue;
    if (A->col_idx[p] == col) continue;
    if (!add_term(A->col_idx[p], -A->vals[p] / pivot, out->cols, out->coeffs, &out->length, OR_MAX_AFFINE))
      return 0;
  }
  return 1;
}

/* affine_interval (:182-208) */
static int affine_interval(const or_affine* f, const double* l, const double* u, double* lo_out, double* hi_out) {
  double lo = f->constant, hi = f->constant;
  for (int32_t k = 0; k < f->length; ++k) {
    const double lj = l[f->cols[k]], uj = u[f->cols[k]], q = f->coeffs[k];
    if (!isfinite(lj) || !isfinite(uj) || lj > uj) return 0;
    lo += q >= 0.0 ? q * lj : q * uj;
    hi += q >= 0.0 ? q * uj : q * lj;
  }
  if (!isfinite(lo) || !isfinite(hi)) return 0;
  *lo_out = lo;
  *hi_out = hi;
  return 1;
}

/* strictly_le_with_guard (:210-218) */
static int le_guard(double lhs, double rhs, double tol) {
  if (!isfinite(lhs) || !isfinite(rhs)) return 0;
  const double scale = 1.0 + fmax(fabs(lhs), fabs(rhs));
  return rhs - lhs >= tol * scale;
}

/* difference_max (:220-254): max over the box of lower(x) - upper(x), terms
 * merged in the order lower then upper. */
static int difference_max(const or_affine* lo, const or_affine* up, const double* l, const double* u, double* out) {
  int32_t cols[OR_MAX_DIFF];
  double coeffs[OR_MAX_DIFF];
  int32_t len = 0;
  for (int32_t k = 0; k < lo->length; ++k)
    if (!add_term(lo->cols[k], lo->coeffs[k], cols, coeffs, &len, OR_MAX_DIFF)) return 0;
  for (int32_t k = 0; k < up->length; ++k)
    if (!add_term(up->cols[k], -up->coeffs[k], cols, coeffs, &len, OR_MAX_DIFF)) return 0;
  double v = lo->constant - up->constant;
  for (int32_t k = 0; k < len; ++k) {
    const double lj = l[cols[k]], uj = u[cols[k]];
    if (!isfinite(lj) || !isfinite(uj) || lj > uj) return 0;
    v += coeffs[k] >= 0.0 ? coeffs[k] * uj : coeffs[k] * lj;
  }
  if (!isfinite(v)) return 0;
  *out = v;
  return 1;
}

/* detect_bounded_two_row_candidates (:256-343) */
static int two_row_candidate(const gp_device_lp* lp, const uint8_t* keep_row, const uint8_t* keep_col, int32_t col,
                             int32_t max_nnz, double zero_tol, double tol, int32_t* rows) {
  const double lj = lp->l[col], uj = lp->u[col];
  if (lp->c[col] != 0.0 || !isfinite(lj) || !isfinite(uj) || !le_guard(lj, uj, tol) ||
      !projection_rows(&lp->AT, keep_row, col, 2, rows))
    return 0;
  if (rows[0] == rows[1]) return 0;
  or_affine f0, f1;
  int lower0 = 0, lower1 = 0;
  if (!row_affine(lp, keep_col, rows[0], col, max_nnz, zero_tol, &f0, &lower0) ||
      !row_affine(lp, keep_col, rows[1], col, max_nnz, zero_tol, &f1, &lower1) || lower0 == lower1)
    return 0;
  const or_affine* lo = lower0 ? &f0 : &f1;
  const or_affine* up = lower0 ? &f1 : &f0;
  double lo_min, lo_max, up_min, up_max, diff;
  if (!affine_interval(lo, lp->l, lp->u, &lo_min, &lo_max) || !affine_interval(up, lp->l, lp->u, &up_min, &up_max))
    return 0;
  if (!le_guard(lo_max, uj, tol) || !le_guard(lj, up_min, tol)) return 0;
  if (!difference_max(lo, up, lp->l, lp->u, &diff) || !le_guard(diff, 0.0, tol)) return 0;
  return 1;
}

/* ---- zero-cost products (rule_zero_cost_product_variables.cu:66-240) ------ */

typedef struct or_mc_row {
  int valid, gives_lower;
  int32_t e0, e1;
} or_mc_row;

/* match_mccormick_row (:73-148): z >= x + y - 1 (3 entries, side -pivot) or
 * z <= x (2 entries, side 0), scaled by any nonzero pivot, endpoints -pivot. */
static or_mc_row mccormick_row(const gp_device_lp* lp, const uint8_t* keep_col, int32_t row, int32_t z) {
  or_mc_row r = {0, 0, -1, -1};
  const gp_csr* A = &lp->A;
  const gp_offset first = A->row_ptr[row], last = A->row_ptr[row + 1];
  int32_t nnz = (int32_t)(last - first);
  if (keep_col) {
    nnz = 0;
    for (gp_offset p = first; p < last; ++p)
      if (keep_col[A->col_idx[p]] && ++nnz > 3) return r;
  }
  if (nnz != 2 && nnz != 3) return r;
  double pivot = 0.0;
  int pivots = 0;
  for (gp_offset p = first; p < last; ++p) {
    if (keep_col && !keep_col[A->col_idx[p]]) continue;
    if (A->col_idx[p] == z) {
      pivot = A->vals[p];
      ++pivots;
    }
  }
  if (pivots != 1 || !isfinite(pivot) || pivot == 0.0) return r;
  double bound;
  if (isfinite(lp->AL[row]) && pos_inf(lp->AU[row])) {
    r.gives_lower = pivot > 0.0;
    bound = lp->AL[row];
  } else if (neg_inf(lp->AL[row]) && isfinite(lp->AU[row])) {
    r.gives_lower = pivot < 0.0;
    bound = lp->AU[row];
  } else {
    return r;
  }
  const int32_t want_nnz = r.gives_lower ? 3 : 2;
  const double want_bound = r.gives_lower ? -pivot : 0.0;
  if (nnz != want_nnz || bound != want_bound) return r;
  int32_t ends = 0;
  for (gp_offset p = first; p < last; ++p) {
    if (keep_col && !keep_col[A->col_idx[p]]) continue;
    const int32_t j = A->col_idx[p];
    if (j == z) continue;
    if (A->vals[p] != -pivot || ends >= 2) return r;
    if (ends == 0)
      r.e0 = j;
    else
      r.e1 = j;
    ++ends;
  }
  if (ends != (r.gives_lower ? 2 : 1) || (r.gives_lower && r.e0 == r.e1)) return r;
  r.valid = 1;
  return r;
}

/* detect_orphan_mccormick_candidates (:150-240) */
static int mccormick_candidate(const gp_device_lp* lp, const uint8_t* keep_row, const uint8_t* keep_col, int32_t z,
                               int32_t* rows, int32_t* ex, int32_t* ey) {
  *ex = *ey = -1;
  if (lp->c[z] != 0.0 || lp->l[z] != 0.0 || lp->u[z] != 1.0 || !projection_rows(&lp->AT, keep_row, z, 3, rows))
    return 0;
  if (rows[0] == rows[1] || rows[0] == rows[2] || rows[1] == rows[2]) return 0;
  or_mc_row pat[3];
  for (int k = 0; k < 3; ++k) pat[k] = mccormick_row(lp, keep_col, rows[k], z);
  int lower = -1, ups[2] = {-1, -1}, nup = 0;
  for (int k = 0; k < 3; ++k) {
    if (!pat[k].valid) return 0;
    if (pat[k].gives_lower) {
      if (lower >= 0) return 0;
      lower = k;
    } else {
      if (nup >= 2) return 0;
      ups[nup++] = k;
    }
  }
  if (lower < 0 || nup != 2) return 0;
  const int32_t x = pat[lower].e0, y = pat[lower].e1;
  const int32_t ux = pat[ups[0]].e0, uy = pat[ups[1]].e0;
  if (x < 0 || y < 0 || x == y || !((ux == x && uy == y) || (ux == y && uy == x)) || lp->l[x] != 0.0 ||
      lp->u[x] != 1.0 || lp->l[y] != 0.0 || lp->u[y] != 1.0)
    return 0;
  *ex = x;
  *ey = y;
  return 1;
}

/* ---- FmeCol record of a projected column (device_recovery.cu:409-479) ----
 * indices [K, col, K rows, N, 2, 1, 0, 1, N, N support columns], values
 * [lower (two rows) or 0 (product), bound / pivot, -a / pivot ...] from the
 * first incident row that bounds the column from below; support = that row's
 * other live columns in row order. */
static void projection_record(or_tape* t, const gp_device_lp* lp, const uint8_t* keep_col, int32_t col, int K,
                              const int32_t* rows) {
  const gp_csr* A = &lp->A;
  int32_t row = -1;
  double pivot = 0.0, bound = 0.0;
  for (int k = 0; k < K && row < 0; ++k) {
    const int32_t r = rows[k];
    double v = 0.0;
    int seen = 0;
    for (gp_offset p = A->row_ptr[r]; p < A->row_ptr[r + 1]; ++p)
      if (A->col_idx[p] == col) {
        v = A->vals[p];
        ++seen;
      }
    if (seen != 1 || !isfinite(v) || v == 0.0) break;
    const int lower = isfinite(lp->AL[r]) && pos_inf(lp->AU[r]) && v > 0.0;
    const int upper = neg_inf(lp->AL[r]) && isfinite(lp->AU[r]) && v < 0.0;
    if (lower || upper) {
      row = r;
      pivot = v;
      bound = lower ? lp->AL[r] : lp->AU[r];
    }
  }
  if (row < 0) return; /* the certificate guarantees one */
  int32_t N = 0;
  for (gp_offset p = A->row_ptr[row]; p < A->row_ptr[row + 1]; ++p)
    N += A->col_idx[p] != col && (!keep_col || keep_col[A->col_idx[p]]);
  int32_t* ix = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(K + 8 + N));
  double* v = (double*)or_xmalloc(sizeof(double) * (size_t)(2 + N));
  ix[0] = K;
  ix[1] = col;
  for (int k = 0; k < K; ++k) ix[2 + k] = rows[k];
  ix[2 + K] = N;
  ix[3 + K] = 2;
  ix[4 + K] = 1;
  ix[5 + K] = 0;
  ix[6 + K] = 1;
  ix[7 + K] = N;
  v[0] = K == 2 ? lp->l[col] : 0.0;
  v[1] = bound / pivot;
  int32_t j = 0;
  for (gp_offset p = A->row_ptr[row]; p < A->row_ptr[row + 1]; ++p)
    if (A->col_idx[p] != col && (!keep_col || keep_col[A->col_idx[p]])) {
      ix[K + 8 + j] = A->col_idx[p];
      v[2 + j] = -A->vals[p] / pivot;
      ++j;
    }
  or_tape_push(t, GP_RED_FME_COL, GP_DUAL_MINIMAL, ix, K + 8 + N, v, 2 + N);
  free(ix);
  free(v);
}

/* analyze_* + build_*_plan of the two projection rules: candidates, rows
 * claimed by one candidate only (products also need both endpoints to be
 * non-candidates), the count threshold against the live columns, records,
 * then the deletions. */
static int projection_rule(int K, const gp_device_lp* view, const gp_device_lp* build_view, const gp_params* p,
                           gp_plan* plan, int pending, int32_t* count) {
  const int32_t m = view->A.rows, n = view->A.cols;
  *count = 0;
  const int enabled = K == 2 ? p->enable_two_row_variable_elimination : p->enable_zero_cost_product_variables;
  if (!enabled || !lp_usable(view)) return 0;
  const uint8_t* keep_row = pending ? plan->keep_row_mask : NULL;
  const uint8_t* keep_col = pending ? plan->keep_col_mask : NULL;
  int32_t max_nnz = p->two_row_variable_elimination_max_row_nnz;
  if (max_nnz > OR_MAX_CERT_ROW) max_nnz = OR_MAX_CERT_ROW;
  if (max_nnz < 2) max_nnz = 2;
  double tol = 10.0 * p->feasibility_tol;
  if (100.0 * p->zero_tol > tol) tol = 100.0 * p->zero_tol;
  if (p->two_row_variable_elimination_safety_tol > tol) tol = p->two_row_variable_elimination_safety_tol;

  int32_t* rows = (int32_t*)or_xcalloc((size_t)n * (size_t)K, sizeof(int32_t));
  uint8_t* cand = (uint8_t*)or_xcalloc((size_t)n, 1);
  int32_t* ex = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)n);
  int32_t* ey = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)n);
  int32_t* row_count = (int32_t*)or_xcalloc((size_t)m, sizeof(int32_t));
  int32_t live = 0;
  for (int32_t j = 0; j < n; ++j) {
    ex[j] = ey[j] = -1;
    if (keep_col && !keep_col[j]) continue;
    ++live;
    int32_t* rj = rows + (size_t)j * K;
    cand[j] = K == 2 ? two_row_candidate(view, keep_row, keep_col, j, max_nnz, p->zero_tol, tol, rj)
                     : mccormick_candidate(view, keep_row, keep_col, j, rj, &ex[j], &ey[j]);
  }
  for (int32_t j = 0; j < n; ++j)
    if (cand[j])
      for (int k = 0; k < K; ++k) row_count[rows[(size_t)j * K + k]]++;
  uint8_t* sel = (uint8_t*)or_xcalloc((size_t)n, 1);
  int32_t selected = 0;
  for (int32_t j = 0; j < n; ++j) {
    if (!cand[j]) continue;
    int ok = 1;
    for (int k = 0; k < K; ++k) ok &= row_count[rows[(size_t)j * K + k]] == 1;
    if (K == 3) ok &= !cand[ex[j]] && !cand[ey[j]];
    sel[j] = (uint8_t)ok;
    selected += ok;
  }
  const int32_t min_c = K == 2 ? p->two_row_variable_elimination_min_candidates
                               : p->zero_cost_product_variables_min_candidates;
  const double min_r = K == 2 ? p->two_row_variable_elimination_min_candidate_ratio
                              : p->zero_cost_product_variables_min_candidate_ratio;
  const double denom = pending ? (double)(live > 1 ? live : 1) : (double)n;
  const int ok = selected >= (min_c > 1 ? min_c : 1) && (double)selected / denom + 1.0e-15 >= min_r;
  if (ok && plan) {
    or_tape* t = (or_tape*)plan->tape;
    if (t && p->record_postsolve_tape)
      for (int32_t j = 0; j < n; ++j)
        if (sel[j]) projection_record(t, build_view, plan->keep_col_mask, j, K, rows + (size_t)j * K);
    for (int32_t j = 0; j < n; ++j)
      if (sel[j]) {
        plan->keep_col_mask[j] = 0;
        for (int k = 0; k < K; ++k) plan->keep_row_mask[rows[(size_t)j * K + k]] = 0;
      }
    plan->has_change = plan->has_row_action = plan->has_col_action = 1;
  }
  *count = ok ? selected : 0;
  free(rows);
  free(cand);
  free(ex);
  free(ey);
  free(row_count);
  free(sel);
  return ok;
}

int or_startup_rule(int32_t rule, const gp_device_lp* lp, const gp_params* p, gp_plan* plan, int pending,
                    int32_t* count) {
  int32_t dummy = 0;
  if (!count) count = &dummy;
  if (pending && !plan) return -1;
  /* projection_view (projection_support.cuh:11-21): the plan's vectors.  The
   * analysis reads them when it runs against the plan's masks; the records
   * of an accepted rule always read them. */
  gp_device_lp view = *lp;
  if (plan) {
    if (plan->new_c) view.c = plan->new_c;
    if (plan->new_l) view.l = plan->new_l;
    if (plan->new_u) view.u = plan->new_u;
    if (plan->new_AL) view.AL = plan->new_AL;
    if (plan->new_AU) view.AU = plan->new_AU;
  }
  switch (rule) {
    case GP_STARTUP_COLUMN_DOMINATION:
      return column_domination(lp, p, plan, count);
    case GP_STARTUP_TWO_ROW_ELIMINATION:
      return projection_rule(2, pending ? &view : lp, &view, p, plan, pending, count);
    case GP_STARTUP_ZERO_COST_PRODUCTS:
      return projection_rule(3, pending ? &view : lp, &view, p, plan, pending, count);
    default:
      return -1;
  }
}
