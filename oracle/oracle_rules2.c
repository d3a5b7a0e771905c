/*
 * oracle_rules2.c -- CPU restatement of the column-side and structural
 * general rules: singleton columns, dual fix, doubleton equations (batch
 * form), parallel rows and parallel columns.
 * TEST INFRASTRUCTURE ONLY (parity checker for the HIP rules).
 *
 * Same execution-order policy as oracle_rules.c: sums the reference forms
 * with atomicAdd / CUB (row shifts, objective deltas) use or_rowacc (CSR
 * order) and or_det_sum; interval max/min use or_cmax / or_cmin.  Work the
 * reference does per hash group or per selected pair is sequential inside
 * the group, exactly as the reference kernels run it.
 */
#include "oracle_internal.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>

static int tape_on(const gp_plan* plan, const gp_params* p) {
  return plan->tape != NULL && p->record_postsolve_tape;
}

double or_fixed_objective_ld(const uint8_t* fixed, const uint8_t* excluded, const double* c, const double* v,
                             int32_t n) {
  long double acc = 0.0L;
  for (int32_t j = 0; j < n; ++j)
    if (fixed[j] && (!excluded || !excluded[j])) acc += (long double)c[j] * (long double)v[j];
  return (double)acc;
}

/* append_compact_fixed_variable_tape_from_device (fixed_variable_tape.cu:141-327):
 * FixedVariable records of the selected columns in column order, idx [col,
 * kept rows...], vals [v, c, coeffs...]; returns the long-double objective
 * when objective != 0, else 0. */
double or_fixed_variable_tape(or_tape* t, const uint8_t* fixed, const uint8_t* excluded, const double* val,
                              const uint8_t* keep_row, const double* c, const gp_csr* AT, int32_t n, int objective) {
  if (n <= 0 || (!t && !objective)) return 0.0;
  if (t) {
    for (int32_t j = 0; j < n; ++j) {
      if (!fixed[j] || (excluded && excluded[j])) continue;
      const gp_offset a0 = AT->row_ptr[j], a1 = AT->row_ptr[j + 1];
      int32_t* ix = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(a1 - a0 + 1));
      double* vx = (double*)or_xmalloc(sizeof(double) * (size_t)(a1 - a0 + 2));
      int64_t ni = 0, nv = 0;
      ix[ni++] = j;
      vx[nv++] = val[j];
      vx[nv++] = c[j];
      for (gp_offset q = a0; q < a1; ++q) {
        const int32_t r = AT->col_idx[q];
        if (!keep_row[r]) continue;
        ix[ni++] = r;
        vx[nv++] = AT->vals[q];
      }
      or_tape_push(t, GP_RED_FIXED_VARIABLE, GP_DUAL_MINIMAL, ix, ni, vx, nv);
      free(ix);
      free(vx);
    }
  }
  return objective ? or_fixed_objective_ld(fixed, excluded, c, val, n) : 0.0;
}

/* FixedVariable records with their kept-row payload, in column order
 * (fixed_variable_tape.cu:36-120): idx [col, rows...], vals [v, c, coeffs...]. */
static void push_fixed_with_rows(or_tape* t, const uint8_t* sel, const double* val,
                                 const uint8_t* keep_row, const double* c, const gp_csr* AT) {
  for (int32_t j = 0; j < AT->rows; ++j) {
    if (!sel[j]) continue;
    const gp_offset a0 = AT->row_ptr[j], a1 = AT->row_ptr[j + 1];
    int32_t* ix = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(a1 - a0 + 1));
    double* vx = (double*)or_xmalloc(sizeof(double) * (size_t)(a1 - a0 + 2));
    int64_t ni = 0, nv = 0;
    ix[ni++] = j;
    vx[nv++] = val[j];
    vx[nv++] = c[j];
    for (gp_offset q = a0; q < a1; ++q) {
      const int32_t r = AT->col_idx[q];
      if (!keep_row[r]) continue;
      ix[ni++] = r;
      vx[nv++] = AT->vals[q];
    }
    or_tape_push(t, GP_RED_FIXED_VARIABLE, GP_DUAL_MINIMAL, ix, ni, vx, nv);
    free(ix);
    free(vx);
  }
}

/* row_shift[i] = sum over stored entries of kept row i whose column has
 * sel[col] of a * val[col], blocked CSR order; 0 for rows not kept. */
static void row_shift_by_columns(const gp_csr* A, const uint8_t* keep_row, const uint8_t* sel,
                                 const double* val, double* shift) {
  for (int32_t i = 0; i < A->rows; ++i) {
    shift[i] = 0.0;
    if (!keep_row[i]) continue;
    or_rowacc acc;
    or_rowacc_init(&acc);
    const gp_offset p0 = A->row_ptr[i];
    for (gp_offset q = p0; q < A->row_ptr[i + 1]; ++q) {
      or_rowacc_pos(&acc, q - p0);
      const int32_t j = A->col_idx[q];
      if (sel[j]) or_rowacc_add(&acc, A->vals[q] * val[j]);
    }
    shift[i] = or_rowacc_value(&acc);
  }
}

/* ======================================================================
 * singleton columns  (rule_singleton_columns.cu:518-1190, 1562-1827)
 * ====================================================================== */
enum { SC_NONE = 0, SC_SET_ROW_SIDE = 1, SC_TIGHTEN_LHS = 2, SC_TIGHTEN_RHS = 3, SC_ELIM_FREE = 4,
       SC_ELIM_ABOVE = 5, SC_ELIM_BELOW = 6 };

typedef struct sc_activity {
  double fmin, fmax;
  int32_t min_neg, min_pos, max_neg, max_pos;
  uint8_t invalid;
} sc_activity;

/* _kernel_singleton_row_activity_cache (:555-635): live entries only, no
 * zero filter; NaN terms mark the row invalid. */
static void sc_row_activity(const gp_csr* A, int32_t i, const uint8_t* keep_col, const double* l,
                            const double* u, sc_activity* out) {
  or_rowacc mn, mx;
  or_rowacc_init(&mn);
  or_rowacc_init(&mx);
  memset(out, 0, sizeof(*out));
  const gp_offset p0 = A->row_ptr[i];
  for (gp_offset q = p0; q < A->row_ptr[i + 1]; ++q) {
    or_rowacc_pos(&mn, q - p0);
    or_rowacc_pos(&mx, q - p0);
    const int32_t j = A->col_idx[q];
    if (!keep_col[j]) continue;
    const double a = A->vals[q];
    const double tmin = a >= 0.0 ? a * l[j] : a * u[j];
    const double tmax = a >= 0.0 ? a * u[j] : a * l[j];
    if (isnan(tmin) || isnan(tmax)) {
      out->invalid = 1;
      continue;
    }
    if (isinf(tmin)) {
      if (tmin < 0.0) out->min_neg++;
      else out->min_pos++;
    } else {
      or_rowacc_add(&mn, tmin);
    }
    if (isinf(tmax)) {
      if (tmax < 0.0) out->max_neg++;
      else out->max_pos++;
    } else {
      or_rowacc_add(&mx, tmax);
    }
  }
  out->fmin = or_rowacc_value(&mn);
  out->fmax = or_rowacc_value(&mx);
}

/* _singleton_column_activity_bounds_from_cache (:756-816) */
static int sc_rest_bounds(const sc_activity* act, double a, double lj, double uj, double* rest_min,
                          double* rest_max) {
  if (act->invalid) return 0;
  const double tmin = a >= 0.0 ? a * lj : a * uj;
  const double tmax = a >= 0.0 ? a * uj : a * lj;
  if (isnan(tmin) || isnan(tmax)) return 0;
  int32_t min_neg = act->min_neg, min_pos = act->min_pos, max_neg = act->max_neg, max_pos = act->max_pos;
  double vmin = act->fmin, vmax = act->fmax;
  if (isinf(tmin)) {
    if (tmin < 0.0) --min_neg;
    else --min_pos;
  } else {
    vmin -= tmin;
  }
  if (isinf(tmax)) {
    if (tmax < 0.0) --max_neg;
    else --max_pos;
  } else {
    vmax -= tmax;
  }
  if (min_neg < 0 || min_pos < 0 || max_neg < 0 || max_pos < 0 || (min_neg > 0 && min_pos > 0) ||
      (max_neg > 0 && max_pos > 0))
    return 0;
  *rest_min = min_neg > 0 ? -INFINITY : (min_pos > 0 ? INFINITY : vmin);
  *rest_max = max_neg > 0 ? -INFINITY : (max_pos > 0 ? INFINITY : vmax);
  return !isnan(*rest_min) && !isnan(*rest_max);
}

static int sc_free_above_ineq(double a, double lhs, double rhs, double ub, double rmin, double rmax,
                              double tol) {
  if (!isfinite(ub)) return 1;
  double iu = INFINITY;
  if (a > 0.0 && isfinite(rhs)) iu = (rhs - rmin) / a;
  else if (a < 0.0 && isfinite(lhs)) iu = (lhs - rmax) / a;
  return iu <= ub + tol;
}

static int sc_free_below_ineq(double a, double lhs, double rhs, double lb, double rmin, double rmax,
                              double tol) {
  if (!isfinite(lb)) return 1;
  double il = -INFINITY;
  if (a > 0.0 && isfinite(lhs)) il = (lhs - rmax) / a;
  else if (a < 0.0 && isfinite(rhs)) il = (rhs - rmin) / a;
  return il >= lb - tol;
}

static int sc_ineq_action(double cj, double a, double lhs, double rhs, int above, int below, double zt,
                          double* side) {
  if (above && below) {
    if ((cj > zt && a > 0.0) || (cj < -zt && a < 0.0)) *side = lhs;
    else if ((cj > zt && a < 0.0) || (cj < -zt && a > 0.0)) *side = rhs;
    else *side = isfinite(lhs) ? lhs : rhs;
    return isfinite(*side) ? SC_SET_ROW_SIDE : SC_NONE;
  }
  if (((cj < -zt && a > 0.0 && above) || (cj > zt && a < 0.0 && below)) && isfinite(rhs)) {
    *side = rhs;
    return SC_TIGHTEN_LHS;
  }
  if (((cj > zt && a > 0.0 && below) || (cj < -zt && a < 0.0 && above)) && isfinite(lhs)) {
    *side = lhs;
    return SC_TIGHTEN_RHS;
  }
  *side = 0.0;
  return SC_NONE;
}

int or_rule_singleton_columns(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                              const gp_params* p) {
  if (plan->has_infeasible || plan->has_unbounded || !p->enable_singleton_columns) return 0;
  const int32_t m = lp->A.rows, n = lp->A.cols;
  if (m == 0 || n == 0) return 0;
  const uint8_t* keep_row = plan->keep_row_mask;
  const uint8_t* keep_col = plan->keep_col_mask;
  const double zt = p->zero_tol, tol = p->bound_tol;
  /* precheck (:1576-1601) */
  int32_t count = 0;
  for (int32_t j = 0; j < n; ++j) {
    if (!keep_col[j] || !s->singleton_column_mask[j]) continue;
    const int32_t r = s->singleton_column_row[j];
    if (r >= 0 && r < m && keep_row[r]) ++count;
  }
  if (count == 0) return 0;

  uint8_t* needed = (uint8_t*)or_xcalloc((size_t)m, 1);
  sc_activity* act = (sc_activity*)or_xcalloc((size_t)m, sizeof(sc_activity));
  int32_t* owner = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)m);
  uint8_t* action = (uint8_t*)or_xcalloc((size_t)n, 1);
  double* side = (double*)or_xcalloc((size_t)n, sizeof(double));
  uint8_t* row_delete = (uint8_t*)or_xcalloc((size_t)m, 1);
  uint8_t* lower_change = (uint8_t*)or_xcalloc((size_t)m, 1);
  uint8_t* upper_change = (uint8_t*)or_xcalloc((size_t)m, 1);
  uint8_t* col_delete = (uint8_t*)or_xcalloc((size_t)n, 1);
  double* AL_new = (double*)or_xcalloc((size_t)m, sizeof(double));
  double* AU_new = (double*)or_xcalloc((size_t)m, sizeof(double));
  double* contrib = (double*)or_xcalloc((size_t)n, sizeof(double));
  double* c_new = (double*)or_xcalloc((size_t)n, sizeof(double));
  for (int32_t i = 0; i < m; ++i) owner[i] = INT_MAX;
  /* activity cache for the support rows (:637-655) */
  for (int32_t j = 0; j < n; ++j) {
    if (!keep_col[j] || !s->singleton_column_mask[j]) continue;
    const int32_t r = s->singleton_column_row[j];
    if (r >= 0 && r < m && keep_row[r]) needed[r] = 1;
  }
  for (int32_t i = 0; i < m; ++i)
    if (needed[i] && keep_row[i]) sc_row_activity(&lp->A, i, keep_col, plan->new_l, plan->new_u, &act[i]);

  /* propose (:905-1017) */
  int status0 = 0, status1 = 0, status2 = 0;
  for (int32_t j = 0; j < n; ++j) {
    if (!keep_col[j] || !s->singleton_column_mask[j]) continue;
    const int32_t row = s->singleton_column_row[j];
    if (row < 0 || row >= m || !keep_row[row]) continue;
    const double a = s->singleton_column_val[j];
    if (fabs(a) <= zt) continue;
    int32_t live = 0;
    for (gp_offset q = lp->A.row_ptr[row]; q < lp->A.row_ptr[row + 1]; ++q)
      if (keep_col[lp->A.col_idx[q]] && ++live > 1) break;
    if (live <= 1) continue;
    const double lhs = plan->new_AL[row], rhs = plan->new_AU[row], cj = plan->new_c[j];
    const double lj = plan->new_l[j], uj = plan->new_u[j];
    if ((cj > zt && a > zt && !isfinite(lhs) && !isfinite(lj)) ||
        (cj > zt && a < -zt && !isfinite(rhs) && !isfinite(lj)) ||
        (cj < -zt && a < -zt && !isfinite(lhs) && !isfinite(uj)) ||
        (cj < -zt && a > zt && !isfinite(rhs) && !isfinite(uj))) {
      status0 = 1;
      continue;
    }
    double rmin = 0.0, rmax = 0.0;
    if (!sc_rest_bounds(&act[row], a, lj, uj, &rmin, &rmax)) continue;
    const int is_eq = isfinite(lhs) && isfinite(rhs) && fabs(lhs - rhs) <= tol;
    if (is_eq) {
      const double x1 = (rhs - rmin) / a, x2 = (rhs - rmax) / a;
      const double ilb = fmin(x1, x2), iub = fmax(x1, x2);
      const int above = !isfinite(uj) || iub <= uj + tol;
      const int below = !isfinite(lj) || ilb >= lj - tol;
      if (!(above || below)) continue;
      action[j] = above && below ? SC_ELIM_FREE : (above ? SC_ELIM_ABOVE : SC_ELIM_BELOW);
      side[j] = rhs;
    } else {
      const int above = sc_free_above_ineq(a, lhs, rhs, uj, rmin, rmax, tol);
      const int below = sc_free_below_ineq(a, lhs, rhs, lj, rmin, rmax, tol);
      double sd = 0.0;
      const int act_ = sc_ineq_action(cj, a, lhs, rhs, above, below, zt, &sd);
      if (act_ == SC_NONE) continue;
      action[j] = (uint8_t)act_;
      side[j] = sd;
    }
    const int32_t key = action[j] == SC_ELIM_FREE ? j - INT_MAX : j;
    if (key < owner[row]) owner[row] = key;
  }

  /* stage (:1021-1073) */
  if (!status0) {
    for (int32_t row = 0; row < m; ++row) {
      const int32_t j = owner[row] < 0 ? owner[row] + INT_MAX : owner[row];
      if (j < 0 || j >= n) continue;
      owner[row] = j;
      const double lhs = plan->new_AL[row], rhs = plan->new_AU[row];
      AL_new[row] = lhs;
      AU_new[row] = rhs;
      const int ac = action[j];
      if (ac == SC_ELIM_FREE || ac == SC_ELIM_ABOVE || ac == SC_ELIM_BELOW) {
        const double a = s->singleton_column_val[j];
        contrib[j] = plan->new_c[j] * rhs / a;
        if (ac == SC_ELIM_FREE) {
          row_delete[row] = 1;
        } else {
          const int above = ac == SC_ELIM_ABOVE;
          const double bound = above ? plan->new_l[j] : plan->new_u[j];
          const int keep_lower = above ? a < 0.0 : a > 0.0;
          const double shifted = AU_new[row] - a * bound;
          if (keep_lower) {
            AL_new[row] = shifted;
            AU_new[row] = INFINITY;
          } else {
            AL_new[row] = -INFINITY;
            AU_new[row] = shifted;
          }
        }
        col_delete[j] = 1;
        status1 = 1;
        continue;
      }
      if (ac == SC_SET_ROW_SIDE) {
        const double sd = side[j];
        if (!isfinite(lhs) || fabs(lhs - sd) > tol) lower_change[row] = 1;
        if (!isfinite(rhs) || fabs(rhs - sd) > tol) upper_change[row] = 1;
        AL_new[row] = sd;
        AU_new[row] = sd;
      } else if (ac == SC_TIGHTEN_LHS) {
        lower_change[row] = 1;
        AL_new[row] = rhs;
      } else if (ac == SC_TIGHTEN_RHS) {
        upper_change[row] = 1;
        AU_new[row] = lhs;
      }
      status2 = 1;
    }
  }

  if (status0) {
    plan->has_unbounded = 1;
  } else if (status1 || status2) {
    if (status1) {
      /* _kernel_apply_singleton_objective_updates_deterministic (:1075-1104) */
      for (int32_t j = 0; j < n; ++j) {
        if (!keep_col[j] || col_delete[j]) continue;
        double updated = plan->new_c[j];
        for (gp_offset q = lp->AT.row_ptr[j]; q < lp->AT.row_ptr[j + 1]; ++q) {
          const int32_t o = owner[lp->AT.col_idx[q]];
          if (o < 0 || o >= n || !col_delete[o]) continue;
          updated += -(plan->new_c[o] * lp->AT.vals[q] / s->singleton_column_val[o]);
        }
        c_new[j] = updated;
      }
    }
    if (status1 && tape_on(plan, p)) {
      /* SingletonColumnSubstitution records in row order (:1191-1374) */
      or_tape* t = (or_tape*)plan->tape;
      for (int32_t row = 0; row < m; ++row) {
        const int32_t j = owner[row];
        if (j < 0 || j >= n || !col_delete[j]) continue;
        const gp_offset r0 = lp->A.row_ptr[row], r1 = lp->A.row_ptr[row + 1];
        int32_t* ix = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(r1 - r0 + 3));
        double* vx = (double*)or_xmalloc(sizeof(double) * (size_t)(r1 - r0 + 6));
        double pivot = 0.0;
        int64_t k = 0;
        for (gp_offset q = r0; q < r1; ++q) {
          const int32_t col = lp->A.col_idx[q];
          if (col == j) {
            pivot = lp->A.vals[q];
          } else if (keep_col[col]) {
            ix[3 + k] = col;
            vx[6 + k] = lp->A.vals[q];
            ++k;
          }
        }
        if (fabs(pivot) > zt) {
          ix[0] = j;
          ix[1] = row;
          ix[2] = (int32_t)k;
          vx[0] = pivot;
          vx[1] = side[j];
          vx[2] = plan->new_l[j];
          vx[3] = plan->new_u[j];
          vx[4] = plan->new_c[j];
          vx[5] = row_delete[row] ? 1.0 : 0.0;
          or_tape_push(t, GP_RED_SINGLETON_COLUMN_SUBSTITUTION, GP_DUAL_MINIMAL, ix, 3 + k, vx, 6 + k);
        }
        free(ix);
        free(vx);
      }
    }
    if (status2 && tape_on(plan, p)) {
      /* RowDualShift records (:459-509) */
      or_tape* t = (or_tape*)plan->tape;
      const double zero = 0.0;
      for (int32_t row = 0; row < m; ++row)
        if (lower_change[row] || upper_change[row])
          or_tape_push(t, GP_RED_ROW_DUAL_SHIFT, GP_DUAL_MINIMAL, &row, 1, &zero, 1);
    }
    /* commit (:1157-1189) */
    for (int32_t row = 0; row < m; ++row) {
      if (row_delete[row]) plan->keep_row_mask[row] = 0;
      if (owner[row] >= 0 && owner[row] < n) {
        plan->new_AL[row] = AL_new[row];
        plan->new_AU[row] = AU_new[row];
      }
    }
    for (int32_t j = 0; j < n; ++j) {
      if (col_delete[j]) plan->keep_col_mask[j] = 0;
      if (status1 && plan->keep_col_mask[j]) plan->new_c[j] = c_new[j];
    }
    plan->obj_constant_delta += or_det_sum(contrib, n);
    plan->has_row_action = 1;
    plan->has_col_action = plan->has_col_action || status1;
    plan->has_change = 1;
  }
  free(needed);
  free(act);
  free(owner);
  free(action);
  free(side);
  free(row_delete);
  free(lower_change);
  free(upper_change);
  free(col_delete);
  free(AL_new);
  free(AU_new);
  free(contrib);
  free(c_new);
  return 0;
}

/* ======================================================================
 * dual fix  (rule_dual_fix.cu:43-170, 186-296, 325-645)
 * ====================================================================== */
int or_rule_dual_fix(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s, const gp_params* p) {
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  const int32_t n = lp->A.cols, m = lp->A.rows;
  if (n == 0) return 0;
  const double zt = p->zero_tol;
  uint8_t* fixed = (uint8_t*)or_xcalloc((size_t)n, 1);
  uint8_t* inf = (uint8_t*)or_xcalloc((size_t)n, 1);
  uint8_t* finite_fixed = (uint8_t*)or_xcalloc((size_t)n, 1);
  double* val = (double*)or_xcalloc((size_t)n, sizeof(double));
  double* contrib = (double*)or_xcalloc((size_t)n, sizeof(double));
  double* shift = (double*)or_xcalloc((size_t)(m ? m : 1), sizeof(double));
  uint8_t* row_remove = (uint8_t*)or_xcalloc((size_t)(m ? m : 1), 1);
  uint8_t* keep_after = (uint8_t*)or_xcalloc((size_t)(m ? m : 1), 1);
  int status0 = 0, status1 = 0, status2 = 0;
  /* candidates (:186-296) */
  for (int32_t j = 0; j < n; ++j) {
    if (!plan->keep_col_mask[j]) continue;
    const double lj = plan->new_l[j], uj = plan->new_u[j];
    if (lj > uj + p->bound_tol) {
      if (status0 < 1) status0 = 1;
      continue;
    }
    const gp_offset a0 = lp->AT.row_ptr[j], a1 = lp->AT.row_ptr[j + 1];
    const int64_t degree = s->live_masks ? (int64_t)s->col_nnz[j] : (int64_t)(a1 - a0);
    if (fabs(plan->new_c[j]) > zt && degree > p->dual_fix_max_nonzero_cost_col_degree) continue;
    int down = 0, up = 0;
    for (gp_offset q = a0; q < a1; ++q) {
      const int32_t r = lp->AT.col_idx[q];
      if (!plan->keep_row_mask[r]) continue;
      const double a = lp->AT.vals[q];
      if (a > zt) {
        down = down || isfinite(plan->new_AL[r]);
        up = up || isfinite(plan->new_AU[r]);
      } else if (a < -zt) {
        down = down || isfinite(plan->new_AU[r]);
        up = up || isfinite(plan->new_AL[r]);
      }
      if (down && up) break;
    }
    const double cj = plan->new_c[j];
    int fix = 0, infinite = 0;
    double v = 0.0;
    if (cj > zt) {
      if (!down) {
        if (!isfinite(lj)) {
          status0 = 2;
          continue;
        }
        v = lj;
        fix = 1;
      }
    } else if (cj < -zt) {
      if (!up) {
        if (!isfinite(uj)) {
          status0 = 2;
          continue;
        }
        v = uj;
        fix = 1;
      }
    } else if (fabs(cj) <= zt) {
      if (!down) {
        fix = 1;
        if (isfinite(lj)) v = lj;
        else {
          v = -INFINITY;
          infinite = 1;
        }
      } else if (!up) {
        fix = 1;
        if (isfinite(uj)) v = uj;
        else {
          v = INFINITY;
          infinite = 1;
        }
      }
    }
    if (fix) {
      fixed[j] = 1;
      inf[j] = (uint8_t)infinite;
      finite_fixed[j] = (uint8_t)!infinite;
      val[j] = v;
      contrib[j] = infinite ? 0.0 : cj * v;
      status1 = 1;
      if (infinite) status2 = 1;
    }
  }
  if (status0 == 1) {
    plan->has_infeasible = 1;
  } else if (status0 == 2) {
    plan->has_unbounded = 1;
  } else if (status1) {
    row_shift_by_columns(&lp->A, plan->keep_row_mask, finite_fixed, val, shift);
    int row_changed = 0;
    for (int32_t i = 0; i < m; ++i) {
      if (!plan->keep_row_mask[i]) continue;
      for (gp_offset q = lp->A.row_ptr[i]; q < lp->A.row_ptr[i + 1]; ++q)
        if (inf[lp->A.col_idx[q]]) {
          row_remove[i] = 1;
          row_changed = 1;
          break;
        }
    }
    if (tape_on(plan, p)) {
      or_tape* t = (or_tape*)plan->tape;
      if (status2) {
        /* InfiniteBoundDualFix records (device_recovery.cu:350-410) */
        const uint8_t* entry = plan->rule_entry_keep_col;
        for (int32_t j = 0; j < n; ++j) {
          if (!fixed[j] || !inf[j]) continue;
          int64_t ni = 3, nv = 1, icap = 3, vcap = 1;
          for (gp_offset q = lp->AT.row_ptr[j]; q < lp->AT.row_ptr[j + 1]; ++q) {
            const int32_t r = lp->AT.col_idx[q];
            if (!plan->keep_row_mask[r]) continue;
            icap += 1 + (lp->A.row_ptr[r + 1] - lp->A.row_ptr[r]);
            vcap += 1 + (lp->A.row_ptr[r + 1] - lp->A.row_ptr[r]);
          }
          int32_t* ix = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)icap);
          double* vx = (double*)or_xmalloc(sizeof(double) * (size_t)vcap);
          const int positive = val[j] > 0.0;
          ix[0] = positive ? 1 : -1;
          ix[1] = j;
          vx[0] = positive ? plan->new_l[j] : plan->new_u[j];
          int32_t active = 0;
          for (gp_offset q = lp->AT.row_ptr[j]; q < lp->AT.row_ptr[j + 1]; ++q) {
            const int32_t r = lp->AT.col_idx[q];
            if (!plan->keep_row_mask[r]) continue;
            ++active;
            const int64_t header = ni++;
            ix[header] = 0;
            vx[nv++] = isfinite(plan->new_AL[r]) ? plan->new_AL[r] : plan->new_AU[r];
            for (gp_offset e = lp->A.row_ptr[r]; e < lp->A.row_ptr[r + 1]; ++e) {
              if (entry && !entry[lp->A.col_idx[e]]) continue;
              ix[header]++;
              ix[ni++] = lp->A.col_idx[e];
              vx[nv++] = lp->A.vals[e];
            }
          }
          ix[2] = active;
          or_tape_push(t, GP_RED_INFINITE_BOUND_DUAL_FIX, GP_DUAL_EXACT, ix, ni, vx, nv);
          free(ix);
          free(vx);
        }
      }
      for (int32_t i = 0; i < m; ++i) keep_after[i] = (uint8_t)(plan->keep_row_mask[i] && !row_remove[i]);
      if (p->record_postsolve_tape_cpu)
        push_fixed_with_rows(t, finite_fixed, val, keep_after, plan->new_c, &lp->AT);
    }
    /* apply (:413-441) */
    for (int32_t j = 0; j < n; ++j) {
      if (!fixed[j]) continue;
      plan->new_l[j] = val[j];
      plan->new_u[j] = val[j];
      plan->keep_col_mask[j] = 0;
    }
    for (int32_t i = 0; i < m; ++i) {
      plan->new_AL[i] -= shift[i];
      plan->new_AU[i] -= shift[i];
      if (row_remove[i]) plan->keep_row_mask[i] = 0;
    }
    plan->obj_constant_delta += or_det_sum(contrib, n);
    plan->has_row_action = plan->has_row_action || row_changed;
    plan->has_col_action = 1;
    plan->has_change = 1;
  }
  free(fixed);
  free(inf);
  free(finite_fixed);
  free(val);
  free(contrib);
  free(shift);
  free(row_remove);
  free(keep_after);
  return 0;
}

/* ======================================================================
 * doubleton equations, batch form  (rule_doubleton_equations.cu:150-272,
 * 824-1340, 1512-1812)
 * ====================================================================== */
#define DBL_MAX_RATIO_PIVOT 1.0e3
#define DBL_MATCHING_ROUNDS 8

static int dbl_integral(double num, double den, double tol) {
  if (fabs(den) <= tol) return 0;
  const double r = fabs(num / den);
  if (!isfinite(r)) return 0;
  return fabs(r - nearbyint(r)) <= tol;
}

/* _doubleton_equations_fill_in_proxy_device_live (:843-890) */
static int32_t dbl_fill_in(const gp_csr* AT, const uint8_t* keep_row, double zt, int32_t keep, int32_t elim) {
  int32_t fill = -1;
  gp_offset jj = AT->row_ptr[keep], kk = AT->row_ptr[elim];
  const gp_offset jstop = AT->row_ptr[keep + 1], kstop = AT->row_ptr[elim + 1];
  if (jstop <= jj || kstop <= kk) return fill;
#define NEXT_LIVE(pos, stop) \
  while ((pos) < (stop) && !(keep_row[AT->col_idx[pos]] && fabs(AT->vals[pos]) > zt)) ++(pos)
  while (jj < jstop && kk < kstop) {
    gp_offset jl = jj, kl = kk;
    NEXT_LIVE(jl, jstop);
    NEXT_LIVE(kl, kstop);
    if (jl >= jstop || kl >= kstop) break;
    const int32_t rj = AT->col_idx[jl], rk = AT->col_idx[kl];
    if (rj == rk) {
      jj = jl + 1;
      kk = kl + 1;
    } else if (rk < rj) {
      kk = kl + 1;
      ++fill;
    } else {
      jj = jl + 1;
    }
  }
  while (kk < kstop) {
    gp_offset kl = kk;
    NEXT_LIVE(kl, kstop);
    if (kl >= kstop) break;
    ++fill;
    kk = kl + 1;
  }
#undef NEXT_LIVE
  return fill;
}

typedef struct dbl_delta {
  int64_t key;
  double val;
} dbl_delta;

static int dbl_delta_cmp(const void* a, const void* b) {
  const int64_t x = ((const dbl_delta*)a)->key, y = ((const dbl_delta*)b)->key;
  return x < y ? -1 : x > y;
}

/* ======================================================================
 * doubleton equations, single-row form (doubleton_equations_scan off;
 * rule_doubleton_equations.cu:1948-2128 with :150-243, 245-571, 1814-1945)
 * ====================================================================== */
/* _doubleton_equations_fill_in_proxy_device (:216-243): every entry of the
 * two columns of A^T counts, live or not */
static int32_t dbl_fill_in_all(const gp_csr* AT, int32_t keep, int32_t elim) {
  int32_t fill = -1;
  gp_offset jj = AT->row_ptr[keep], kk = AT->row_ptr[elim];
  const gp_offset jstop = AT->row_ptr[keep + 1], kstop = AT->row_ptr[elim + 1];
  if (jstop <= jj || kstop <= kk) return fill;
  while (jj < jstop && kk < kstop) {
    const int32_t rj = AT->col_idx[jj], rk = AT->col_idx[kk];
    if (rj == rk) {
      ++jj;
      ++kk;
    } else if (rk < rj) {
      ++kk;
      ++fill;
    } else {
      ++jj;
    }
  }
  return fill + (int32_t)(kstop - kk);
}

/* The single-row form eliminates one doubleton equation per call: the lowest
 * live equality row with two live entries whose pivot and fill-in pass
 * (:273-363).  The reference picks that row by atomicMin and lets every
 * thread that lowered the minimum write the row's data; here the data are
 * those of the minimum row, the selection the reference's kernel intends. */
static int dbl_single_row(gp_plan* plan, const gp_device_lp* lp, const gp_params* p) {
  const int32_t m = lp->A.rows, n = lp->A.cols;
  const gp_csr* A = plan->has_new_A ? &plan->new_A : &lp->A;
  const gp_csr* AT = &lp->AT;
  const double zt = p->zero_tol, bt = p->bound_tol;
  const double ratio_tol = fmax(1.0e-9, bt);
  uint8_t* keep_row = plan->keep_row_mask;
  uint8_t* keep_col = plan->keep_col_mask;
  /* live column counts (:245-271) */
  int32_t* col_nnz = (int32_t*)or_xcalloc((size_t)n, sizeof(int32_t));
  for (int32_t j = 0; j < n; ++j) {
    if (!keep_col[j]) continue;
    int32_t c = 0;
    for (gp_offset q = AT->row_ptr[j]; q < AT->row_ptr[j + 1]; ++q)
      if (keep_row[AT->col_idx[q]] && fabs(AT->vals[q]) > zt) ++c;
    col_nnz[j] = c;
  }
  int32_t target = -1, ec = -1, kc = -1;
  double ev = 0.0, kv = 0.0;
  for (int32_t i = 0; i < m && target < 0; ++i) {
    if (!keep_row[i]) continue;
    const double lhs = plan->new_AL[i], rhs = plan->new_AU[i];
    if (!isfinite(lhs) || !isfinite(rhs) || fabs(lhs - rhs) > bt) continue;
    int32_t live = 0, c1 = -1, c2 = -1;
    double v1 = 0.0, v2 = 0.0;
    for (gp_offset q = A->row_ptr[i]; q < A->row_ptr[i + 1]; ++q) {
      const int32_t col = A->col_idx[q];
      const double v = A->vals[q];
      if (keep_col[col] && fabs(v) > zt) {
        ++live;
        if (live == 1) {
          c1 = col;
          v1 = v;
        } else if (live == 2) {
          c2 = col;
          v2 = v;
        } else {
          break;
        }
      }
    }
    if (live != 2) continue;
    /* _choose_doubleton_equations_columns (:170-214) */
    const int i12 = dbl_integral(v1, v2, ratio_tol), i21 = dbl_integral(v2, v1, ratio_tol);
    int32_t e, k;
    double evv, kvv;
    if (col_nnz[c1] == 1 && col_nnz[c2] != 1) {
      e = c1, evv = v1, k = c2, kvv = v2;
    } else if (col_nnz[c1] != 1 && col_nnz[c2] == 1) {
      e = c2, evv = v2, k = c1, kvv = v1;
    } else if (i12 && !i21) {
      e = c2, evv = v2, k = c1, kvv = v1;
    } else if (i21 && !i12) {
      e = c1, evv = v1, k = c2, kvv = v2;
    } else if (col_nnz[c1] < col_nnz[c2]) {
      e = c1, evv = v1, k = c2, kvv = v2;
    } else {
      e = c2, evv = v2, k = c1, kvv = v1;
    }
    /* _is_acceptable_doubleton_equations_pivot (:161-168) */
    if (fabs(evv) <= 0.0) continue;
    const double pr = fabs(kvv / evv);
    if (!(pr <= DBL_MAX_RATIO_PIVOT && pr >= 1.0 / DBL_MAX_RATIO_PIVOT)) continue;
    if (col_nnz[e] > col_nnz[k] + p->doubleton_equations_max_fill_in_proxy + 1) continue;
    if (dbl_fill_in_all(AT, k, e) > p->doubleton_equations_max_fill_in_proxy) continue;
    target = i, ec = e, kc = k, ev = evv, kv = kvv;
  }
  free(col_nnz);
  if (target < 0) return 0;
  /* _kernel_doubleton_equations_apply (:365-461) */
  const double rhs = plan->new_AU[target];
  const double old_el = plan->new_l[ec], old_eu = plan->new_u[ec], elim_obj = plan->new_c[ec];
  const double old_kl = plan->new_l[kc], old_ku = plan->new_u[kc];
  const double alpha = -kv / ev, beta = rhs / ev;
  double ml = -INFINITY, mu = INFINITY;
  if (isfinite(old_el)) {
    const double b = (old_el - beta) / alpha;
    if (alpha > 0.0) ml = b;
    else mu = b;
  }
  if (isfinite(old_eu)) {
    const double b = (old_eu - beta) / alpha;
    if (alpha > 0.0) mu = b;
    else ml = b;
  }
  if (ml > mu) {
    const double tmp = ml;
    ml = mu;
    mu = tmp;
  }
  const double kl = isfinite(ml) ? fmax(old_kl, ml) : old_kl;
  const double ku = isfinite(mu) ? fmin(old_ku, mu) : old_ku;
  if (kl > ku + bt) {
    plan->has_infeasible = 1;
    return 0;
  }
  /* every row: the first entry of the eliminated column in that row, times beta */
  for (int32_t t = 0; t < m; ++t) {
    double shift = 0.0;
    for (gp_offset q = AT->row_ptr[ec]; q < AT->row_ptr[ec + 1]; ++q) {
      if (AT->col_idx[q] == t) {
        shift = AT->vals[q] * beta;
        break;
      }
    }
    plan->new_AL[t] -= shift;
    plan->new_AU[t] -= shift;
  }
  plan->new_l[kc] = kl;
  plan->new_u[kc] = ku;
  plan->new_c[kc] += elim_obj * alpha;
  keep_row[target] = 0;
  keep_col[ec] = 0;
  /* DoubletonEquation record (:2059-2101; capture_scalar_recovery :571-575) */
  if (tape_on(plan, p)) {
    const int32_t ix[4] = {ec, kc, target, 0};
    const double vx[12] = {ev, kv, rhs, old_el, old_eu, old_kl, old_ku, kl, ku, elim_obj, 0.0, 0.0};
    or_tape_push((or_tape*)plan->tape, GP_RED_DOUBLETON_EQUATION, GP_DUAL_MINIMAL, ix, 4, vx, 12);
  }
  /* build_doubleton_scalar_new_A (:1894-1945) with
   * _kernel_doubleton_equations_row_counts / _kernel_copy_doubleton_equations_rows
   * (:463-521, 1814-1892): the target row is copied as it is; in every other
   * row the keep entry becomes old_keep + alpha * a_elim (dropped when
   * |.| <= zero_tol, inserted in column order when the row had an eliminated
   * entry but no keep entry); the eliminated column's entries stay */
  gp_csr out;
  out.rows = m;
  out.cols = n;
  out.row_ptr = (gp_offset*)or_xcalloc((size_t)m + 1, sizeof(gp_offset));
  out.col_idx = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(A->nnz + m + 1));
  out.vals = (double*)or_xmalloc(sizeof(double) * (size_t)(A->nnz + m + 1));
  gp_offset w = 0;
  for (int32_t r = 0; r < m; ++r) {
    const gp_offset first = A->row_ptr[r], last = A->row_ptr[r + 1];
    if (r == target) {
      for (gp_offset q = first; q < last; ++q) {
        out.col_idx[w] = A->col_idx[q];
        out.vals[w++] = A->vals[q];
      }
      out.row_ptr[r + 1] = w;
      continue;
    }
    int elim_present = 0, keep_present = 0;
    double are = 0.0, old_keep = 0.0;
    for (gp_offset q = first; q < last; ++q) {
      if (A->col_idx[q] == ec) {
        elim_present = 1;
        are = A->vals[q];
      } else if (A->col_idx[q] == kc) {
        keep_present = 1;
        old_keep = A->vals[q];
      }
    }
    const double new_keep = old_keep + alpha * are;
    const int insert_keep = elim_present && !keep_present && fabs(new_keep) > zt;
    int inserted = !insert_keep;
    for (gp_offset q = first; q < last; ++q) {
      const int32_t col = A->col_idx[q];
      if (insert_keep && !inserted && col > kc) {
        out.col_idx[w] = kc;
        out.vals[w++] = new_keep;
        inserted = 1;
      }
      if (col == kc) {
        if (fabs(new_keep) > zt) {
          out.col_idx[w] = kc;
          out.vals[w++] = new_keep;
        }
      } else {
        out.col_idx[w] = col;
        out.vals[w++] = A->vals[q];
      }
    }
    if (insert_keep && !inserted) {
      out.col_idx[w] = kc;
      out.vals[w++] = new_keep;
    }
    out.row_ptr[r + 1] = w;
  }
  out.nnz = w;
  if (plan->has_new_A) or_csr_free(&plan->new_A);
  plan->new_A = out;
  plan->has_new_A = 1;
  plan->obj_constant_delta += elim_obj * beta;
  plan->has_row_action = 1;
  plan->has_col_action = 1;
  plan->has_change = 1;
  return 0;
}

int or_rule_doubleton_equations(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                                const gp_params* p) {
  (void)s;
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  if (!p->doubleton_equations_scan) {
    if (lp->A.rows == 0 || lp->A.cols == 0) return 0;
    return dbl_single_row(plan, lp, p);
  }
  const int32_t m = lp->A.rows, n = lp->A.cols;
  if (m == 0 || n == 0) return 0;
  const gp_csr* A = plan->has_new_A ? &plan->new_A : &lp->A;
  const gp_csr* AT = &lp->AT;
  const double zt = p->zero_tol, bt = p->bound_tol;
  const double ratio_tol = fmax(1.0e-9, bt);
  uint8_t* keep_row = plan->keep_row_mask;
  uint8_t* keep_col = plan->keep_col_mask;

  int32_t* col_nnz = (int32_t*)or_xcalloc((size_t)n, sizeof(int32_t));
  uint8_t* candidate = (uint8_t*)or_xcalloc((size_t)m, 1);
  uint8_t* active = (uint8_t*)or_xcalloc((size_t)m, 1);
  uint8_t* selected = (uint8_t*)or_xcalloc((size_t)m, 1);
  uint8_t* blocked = (uint8_t*)or_xcalloc((size_t)n, 1);
  int32_t* owner = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)n);
  int32_t* cand_elim = (int32_t*)or_xcalloc((size_t)m, sizeof(int32_t));
  int32_t* cand_keep = (int32_t*)or_xcalloc((size_t)m, sizeof(int32_t));
  double* cand_elim_val = (double*)or_xcalloc((size_t)m, sizeof(double));
  double* cand_keep_val = (double*)or_xcalloc((size_t)m, sizeof(double));

  /* live column counts (:245-271) */
  for (int32_t j = 0; j < n; ++j) {
    if (!keep_col[j]) continue;
    int32_t c = 0;
    for (gp_offset q = AT->row_ptr[j]; q < AT->row_ptr[j + 1]; ++q)
      if (keep_row[AT->col_idx[q]] && fabs(AT->vals[q]) > zt) ++c;
    col_nnz[j] = c;
  }
  /* candidates (:892-958) and the fill-in filter (:960-987) */
  int32_t acceptable = 0;
  for (int32_t i = 0; i < m; ++i) {
    if (!keep_row[i]) continue;
    const double lhs = plan->new_AL[i], rhs = plan->new_AU[i];
    if (!isfinite(lhs) || !isfinite(rhs) || fabs(lhs - rhs) > bt) continue;
    int32_t live = 0, c1 = -1, c2 = -1;
    double v1 = 0.0, v2 = 0.0;
    for (gp_offset q = A->row_ptr[i]; q < A->row_ptr[i + 1]; ++q) {
      const int32_t col = A->col_idx[q];
      const double v = A->vals[q];
      if (keep_col[col] && fabs(v) > zt) {
        ++live;
        if (live == 1) {
          c1 = col;
          v1 = v;
        } else if (live == 2) {
          c2 = col;
          v2 = v;
        } else {
          break;
        }
      }
    }
    if (live != 2) continue;
    const int i12 = dbl_integral(v1, v2, ratio_tol), i21 = dbl_integral(v2, v1, ratio_tol);
    int32_t ec, kc;
    double ev, kv;
    if (col_nnz[c1] == 1 && col_nnz[c2] != 1) {
      ec = c1, ev = v1, kc = c2, kv = v2;
    } else if (col_nnz[c1] != 1 && col_nnz[c2] == 1) {
      ec = c2, ev = v2, kc = c1, kv = v1;
    } else if (i12 && !i21) {
      ec = c2, ev = v2, kc = c1, kv = v1;
    } else if (i21 && !i12) {
      ec = c1, ev = v1, kc = c2, kv = v2;
    } else if (col_nnz[c1] < col_nnz[c2]) {
      ec = c1, ev = v1, kc = c2, kv = v2;
    } else {
      ec = c2, ev = v2, kc = c1, kv = v1;
    }
    if (fabs(ev) <= 0.0) continue;
    const double pr = fabs(kv / ev);
    if (!(pr <= DBL_MAX_RATIO_PIVOT && pr >= 1.0 / DBL_MAX_RATIO_PIVOT)) continue;
    candidate[i] = 1;
    cand_elim[i] = ec;
    cand_keep[i] = kc;
    cand_elim_val[i] = ev;
    cand_keep_val[i] = kv;
    if (col_nnz[ec] > col_nnz[kc] + p->doubleton_equations_max_fill_in_proxy + 1) continue;
    if (dbl_fill_in(AT, keep_row, zt, kc, ec) <= p->doubleton_equations_max_fill_in_proxy) {
      active[i] = 1;
      ++acceptable;
    }
  }
  /* matching rounds (:989-1027) */
  for (int round = 0; round < DBL_MATCHING_ROUNDS; ++round) {
    for (int32_t j = 0; j < n; ++j) owner[j] = INT_MAX;
    for (int32_t i = 0; i < m; ++i) {
      if (!active[i]) continue;
      const int32_t kc = cand_keep[i], ec = cand_elim[i];
      if (!blocked[kc] && !blocked[ec]) {
        if (i < owner[kc]) owner[kc] = i;
        if (i < owner[ec]) owner[ec] = i;
      }
    }
    for (int32_t i = 0; i < m; ++i) {
      if (!active[i]) continue;
      const int32_t kc = cand_keep[i], ec = cand_elim[i];
      if (!blocked[kc] && !blocked[ec] && owner[kc] == i && owner[ec] == i) {
        selected[i] = 1;
        active[i] = 0;
        blocked[kc] = 1;
        blocked[ec] = 1;
      }
    }
  }
  int32_t k = 0;
  for (int32_t i = 0; i < m; ++i) k += selected[i];
  const double ratio = acceptable == 0 ? 0.0 : (double)k / (double)acceptable;
  if (k == 0 || (p->doubleton_equations_scan && (k < p->doubleton_equations_min_selected_per_batch ||
                                                  ratio < p->doubleton_equations_min_selected_ratio)))
    goto done;
  {
    int32_t* rows = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)k);
    double* tape_rec = (double*)or_xcalloc((size_t)k * 9, sizeof(double));
    uint8_t* subst = (uint8_t*)or_xcalloc((size_t)k, 1);
    uint8_t* fixedp = (uint8_t*)or_xcalloc((size_t)k, 1);
    double* obj_s = (double*)or_xcalloc((size_t)k, sizeof(double));
    double* obj_f = (double*)or_xcalloc((size_t)k, sizeof(double));
    int status0 = 0, status1 = 0;
    int32_t t = 0;
    for (int32_t i = 0; i < m; ++i)
      if (selected[i]) rows[t++] = i;
    /* capture (:770-803): old AU(row), elim l/u, keep l/u, elim c */
    for (t = 0; t < k; ++t) {
      const int32_t row = rows[t], kc = cand_keep[row], ec = cand_elim[row];
      double* r = tape_rec + (size_t)t * 9;
      r[0] = plan->new_AU[row];
      r[1] = plan->new_l[ec];
      r[2] = plan->new_u[ec];
      r[3] = plan->new_l[kc];
      r[4] = plan->new_u[kc];
      r[5] = plan->new_c[ec];
    }
    /* apply (:1029-1117); the selected pairs touch disjoint columns */
    for (t = 0; t < k; ++t) {
      const int32_t row = rows[t], kc = cand_keep[row], ec = cand_elim[row];
      const double kv = cand_keep_val[row], ev = cand_elim_val[row];
      const double alpha = -kv / ev, beta = plan->new_AU[row] / ev;
      double ml = -INFINITY, mu = INFINITY;
      if (isfinite(plan->new_l[ec])) {
        const double b = (plan->new_l[ec] - beta) / alpha;
        if (alpha > 0.0) ml = b;
        else mu = b;
      }
      if (isfinite(plan->new_u[ec])) {
        const double b = (plan->new_u[ec] - beta) / alpha;
        if (alpha > 0.0) mu = b;
        else ml = b;
      }
      if (ml > mu) {
        const double tmp = ml;
        ml = mu;
        mu = tmp;
      }
      const double kl = isfinite(ml) ? or_cmax(plan->new_l[kc], ml) : plan->new_l[kc];
      const double ku = isfinite(mu) ? or_cmin(plan->new_u[kc], mu) : plan->new_u[kc];
      if (kl > ku + bt) {
        status0 = 1;
        continue;
      }
      if (isfinite(kl) && isfinite(ku) && ku <= kl + bt) {
        const double at = 0.5 * (kl + ku);
        plan->new_l[kc] = at;
        plan->new_u[kc] = at;
        keep_col[kc] = 0;
        fixedp[t] = 1;
        obj_f[t] = plan->new_c[kc] * at;
      } else {
        plan->new_l[kc] = kl;
        plan->new_u[kc] = ku;
        plan->new_c[kc] += plan->new_c[ec] * alpha;
        keep_row[row] = 0;
        keep_col[ec] = 0;
        subst[t] = 1;
        obj_s[t] = plan->new_c[ec] * beta;
      }
      status1 = 1;
    }
    if (status0) {
      plan->has_infeasible = 1;
    } else if (status1) {
      int32_t subst_count = 0;
      for (t = 0; t < k; ++t) subst_count += subst[t];
      if (tape_on(plan, p)) {
        /* DoubletonEquation records in selected order (:641-669) */
        or_tape* tp = (or_tape*)plan->tape;
        for (t = 0; t < k; ++t) {
          if (!subst[t]) continue;
          const int32_t row = rows[t], kc = cand_keep[row], ec = cand_elim[row];
          const double* r = tape_rec + (size_t)t * 9;
          const int32_t ix[4] = {ec, kc, row, 0};
          const double vx[12] = {cand_elim_val[row], cand_keep_val[row], r[0], r[1], r[2], r[3], r[4],
                                 plan->new_l[kc], plan->new_u[kc], r[5], 0.0, 0.0};
          or_tape_push(tp, GP_RED_DOUBLETON_EQUATION, GP_DUAL_MINIMAL, ix, 4, vx, 12);
        }
      }
      /* row shifts: a * beta over the eliminated columns' other live rows and
       * a * fixed_at over fixed keep columns (:1164-1229), in CSR row order */
      int32_t* pair_of = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)n);
      for (int32_t j = 0; j < n; ++j) pair_of[j] = -1;
      for (t = 0; t < k; ++t) {
        const int32_t row = rows[t];
        if (subst[t]) pair_of[cand_elim[row]] = t;
        else if (fixedp[t]) pair_of[cand_keep[row]] = t;
      }
      double* shift = (double*)or_xcalloc((size_t)m, sizeof(double));
      int64_t total = 0;
      for (int32_t i = 0; i < m; ++i) {
        if (!keep_row[i]) continue;
        or_rowacc acc;
        or_rowacc_init(&acc);
        const gp_offset p0 = A->row_ptr[i];
        for (gp_offset q = p0; q < A->row_ptr[i + 1]; ++q) {
          or_rowacc_pos(&acc, q - p0);
          const int32_t j = A->col_idx[q];
          const int32_t tt = pair_of[j];
          const double a = A->vals[q];
          if (tt < 0 || !(fabs(a) > zt)) continue;
          const int32_t prow = rows[tt];
          if (subst[tt]) {
            if (prow == i) continue;
            or_rowacc_add(&acc, a * (plan->new_AU[prow] / cand_elim_val[prow]));
            ++total;
          } else {
            or_rowacc_add(&acc, a * plan->new_l[cand_keep[prow]]);
          }
        }
        shift[i] = or_rowacc_value(&acc);
      }
      if (subst_count > 0 && total > 0) {
        /* new A = source A merged with alpha * a deltas at (support row, keep) */
        dbl_delta* d = (dbl_delta*)or_xmalloc(sizeof(dbl_delta) * (size_t)total);
        int64_t w = 0;
        for (t = 0; t < k; ++t) {
          if (!subst[t]) continue;
          const int32_t row = rows[t], kc = cand_keep[row], ec = cand_elim[row];
          const double alpha = -cand_keep_val[row] / cand_elim_val[row];
          for (gp_offset q = AT->row_ptr[ec]; q < AT->row_ptr[ec + 1]; ++q) {
            const int32_t sr = AT->col_idx[q];
            const double a = AT->vals[q];
            if (sr != row && keep_row[sr] && fabs(a) > zt) {
              d[w].key = (int64_t)sr * (int64_t)n + kc;
              d[w].val = alpha * a;
              ++w;
            }
          }
        }
        qsort(d, (size_t)w, sizeof(dbl_delta), dbl_delta_cmp);
        gp_offset* drp = (gp_offset*)or_xcalloc((size_t)m + 1, sizeof(gp_offset));
        for (int64_t e = 0; e < w; ++e) drp[d[e].key / n + 1]++;
        for (int32_t i = 0; i < m; ++i) drp[i + 1] += drp[i];
        gp_csr out;
        out.rows = m;
        out.cols = n;
        out.row_ptr = (gp_offset*)or_xcalloc((size_t)m + 1, sizeof(gp_offset));
        out.col_idx = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(A->nnz + w + 1));
        out.vals = (double*)or_xmalloc(sizeof(double) * (size_t)(A->nnz + w + 1));
        gp_offset wr = 0;
        for (int32_t i = 0; i < m; ++i) {
          gp_offset pp = A->row_ptr[i], qq = drp[i];
          const gp_offset pe = A->row_ptr[i + 1], qe = drp[i + 1];
          while (pp < pe || qq < qe) {
            const int32_t sc = pp < pe ? A->col_idx[pp] : INT_MAX;
            const int32_t dc = qq < qe ? (int32_t)(d[qq].key % n) : INT_MAX;
            const int32_t next = sc < dc ? sc : dc;
            double merged = 0.0;
            while (pp < pe && A->col_idx[pp] == next) merged += A->vals[pp++];
            while (qq < qe && (int32_t)(d[qq].key % n) == next) merged += d[qq++].val;
            if (fabs(merged) > zt) {
              out.col_idx[wr] = next;
              out.vals[wr] = merged;
              ++wr;
            }
          }
          out.row_ptr[i + 1] = wr;
        }
        out.nnz = wr;
        if (plan->has_new_A) or_csr_free(&plan->new_A);
        plan->new_A = out;
        plan->has_new_A = 1;
        free(drp);
        free(d);
      }
      for (int32_t i = 0; i < m; ++i) {
        plan->new_AL[i] -= shift[i];
        plan->new_AU[i] -= shift[i];
      }
      plan->has_row_action = subst_count > 0;
      plan->has_col_action = 1;
      plan->has_change = 1;
      plan->obj_constant_delta += or_det_sum(obj_s, k);
      plan->obj_constant_delta += or_det_sum(obj_f, k);
      free(shift);
      free(pair_of);
    }
    free(rows);
    free(tape_rec);
    free(subst);
    free(fixedp);
    free(obj_s);
    free(obj_f);
  }
done:
  free(col_nnz);
  free(candidate);
  free(active);
  free(selected);
  free(blocked);
  free(owner);
  free(cand_elim);
  free(cand_keep);
  free(cand_elim_val);
  free(cand_keep_val);
  return 0;
}

/* ======================================================================
 * parallel rows  (rule_parallel_rows.cu:141-466, 648-888)
 * ====================================================================== */
#define FNV_OFFSET 0xcbf29ce484222325ULL
#define FNV_PRIME 0x100000001b3ULL

static uint64_t fnv_mix(uint64_t h, uint64_t x) { return (h ^ x) * FNV_PRIME; }
static uint64_t dbits(double v) {
  uint64_t b;
  memcpy(&b, &v, sizeof(b));
  return b;
}

typedef struct hash_item {
  uint64_t h;
  int32_t id;
} hash_item;

static int hash_item_cmp(const void* a, const void* b) {
  const hash_item* x = (const hash_item*)a;
  const hash_item* y = (const hash_item*)b;
  if (x->h != y->h) return x->h < y->h ? -1 : 1;
  return x->id < y->id ? -1 : x->id > y->id;
}

/* _kernel_parallel_row_hashes (:141-196): FNV over (col, bits(a / pivot)) of
 * the live entries; XOR-aggregated when nnz > 32 m. */
static uint64_t row_hash(const gp_csr* A, int32_t i, const uint8_t* keep_row, const uint8_t* keep_col,
                         double tol, int use_xor) {
  if (!keep_row[i]) return (uint64_t)(int64_t)i;
  const gp_offset p1 = A->row_ptr[i + 1];
  gp_offset first = A->row_ptr[i];
  while (keep_col && first < p1 && !keep_col[A->col_idx[first]]) ++first;
  int64_t len = 0;
  for (gp_offset q = first; q < p1; ++q)
    if (!keep_col || keep_col[A->col_idx[q]]) ++len;
  if (len <= 0) return fnv_mix(FNV_OFFSET, (uint64_t)(int64_t)i);
  const double pivot = A->vals[first];
  if (fabs(pivot) <= tol) return fnv_mix(FNV_OFFSET, (uint64_t)(int64_t)i);
  uint64_t h = fnv_mix(FNV_OFFSET, (uint64_t)len), agg = 0;
  for (gp_offset q = first; q < p1; ++q) {
    const int32_t col = A->col_idx[q];
    if (keep_col && !keep_col[col]) continue;
    const uint64_t nb = dbits(A->vals[q] / pivot);
    if (use_xor) {
      agg ^= fnv_mix(fnv_mix(FNV_OFFSET, (uint64_t)(int64_t)col), nb);
    } else {
      h = fnv_mix(h, (uint64_t)(int64_t)col);
      h = fnv_mix(h, nb);
    }
  }
  return use_xor ? fnv_mix(h, agg) : h;
}

/* compute_parallel_row_hashes (rule_parallel_rows.cu:583-627): one hash per row. */
void or_parallel_row_hashes(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, double tol,
                            uint64_t* out) {
  const int use_xor = (int64_t)A->nnz > 32 * (int64_t)A->rows;
  for (int32_t i = 0; i < A->rows; ++i) out[i] = row_hash(A, i, keep_row, keep_col, tol, use_xor);
}

/* _parallel_row_ratio (:312-346): ratio = a_i / a_k on the first live pair,
 * every pair within the absolute tolerance. */
static int row_ratio(const gp_csr* A, int32_t i, int32_t k, const uint8_t* keep_col, double tol,
                     double* out) {
  gp_offset li = A->row_ptr[i], le = A->row_ptr[i + 1], ri = A->row_ptr[k], re = A->row_ptr[k + 1];
#define SKIP(pos, end) while (keep_col && (pos) < (end) && !keep_col[A->col_idx[pos]]) ++(pos)
  SKIP(li, le);
  SKIP(ri, re);
  if (li == le && ri == re) {
    *out = 1.0;
    return 1;
  }
  double ratio = 0.0;
  int set = 0;
  for (; li < le && ri < re;) {
    if (A->col_idx[li] != A->col_idx[ri]) return 0;
    const double ai = A->vals[li], ak = A->vals[ri];
    if (!set) {
      if (fabs(ak) <= tol) return 0;
      ratio = ai / ak;
      set = 1;
    }
    if (fabs(ai - ratio * ak) > tol) return 0;
    ++li;
    SKIP(li, le);
    ++ri;
    SKIP(ri, re);
  }
#undef SKIP
  *out = ratio;
  return set && li == le && ri == re;
}

static void scaled_interval(double lo, double hi, double ratio, double* sl, double* su) {
  const double rl = lo * ratio, ru = hi * ratio;
  if (ratio >= 0.0) {
    *sl = rl;
    *su = ru;
  } else {
    *sl = ru;
    *su = rl;
  }
}

typedef struct prow_rec {
  int32_t step, kept, deleted;
  double ratio, kal, kau, dal, dau;
} prow_rec;

static int prow_rec_cmp(const void* a, const void* b) {
  const prow_rec* x = (const prow_rec*)a;
  const prow_rec* y = (const prow_rec*)b;
  if (x->step != y->step) return x->step < y->step ? -1 : 1;
  return x->deleted < y->deleted ? -1 : x->deleted > y->deleted;
}

int or_rule_parallel_rows(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                          const gp_params* p) {
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  const int32_t m = lp->A.rows;
  if (m <= 1) return 0;
  const uint8_t* keep_col = s->live_masks ? plan->keep_col_mask : NULL;
  const uint8_t* keep_row = plan->keep_row_mask;
  const double ctol = p->zero_tol, itol = p->feasibility_tol;
  const int use_xor = (int64_t)lp->A.nnz > 32 * (int64_t)m;
  double* AL = plan->new_AL;
  double* AU = plan->new_AU;
  hash_item* items = (hash_item*)or_xmalloc(sizeof(hash_item) * (size_t)m);
  for (int32_t i = 0; i < m; ++i) {
    items[i].h = row_hash(&lp->A, i, keep_row, keep_col, ctol, use_xor);
    items[i].id = i;
  }
  qsort(items, (size_t)m, sizeof(hash_item), hash_item_cmp);
  uint8_t* row_delete = (uint8_t*)or_xcalloc((size_t)m, 1);
  prow_rec* recs = (prow_rec*)or_xcalloc((size_t)m, sizeof(prow_rec));
  int32_t nrec = 0;
  int status = 0;
  /* _kernel_parallel_row_groups (:361-466): one sequential pass per group */
  for (int32_t gs = 0; gs < m;) {
    int32_t ge = gs;
    while (ge + 1 < m && items[ge + 1].h == items[gs].h) ++ge;
    const int32_t next = ge + 1;
    if (ge > gs) {
      int32_t seed = -1;
      int seed_eq = 0;
      for (int32_t a = gs; a <= ge; ++a) {
        const int32_t row = items[a].id;
        if (!keep_row[row]) continue;
        const int eq = isfinite(AL[row]) && isfinite(AU[row]) && fabs(AL[row] - AU[row]) <= itol;
        if (seed < 0 || (eq && !seed_eq)) {
          seed = row;
          seed_eq = eq;
        }
      }
      if (seed >= 0) {
        int32_t rep = seed;
        double rsl = AL[seed], rsu = AU[seed];
        for (int32_t a = gs; a <= ge; ++a) {
          const int32_t row = items[a].id;
          if (row == seed || !keep_row[row]) continue;
          double r = 0.0;
          if (!row_ratio(&lp->A, seed, row, keep_col, ctol, &r)) continue;
          double sl, su;
          scaled_interval(AL[row], AU[row], r, &sl, &su);
          if (rsl <= sl + itol && su <= rsu + itol) {
            rep = row;
            rsl = sl;
            rsu = su;
          }
        }
        double rl = AL[rep], ru = AU[rep];
        int found = 0, abandoned = 0;
        int32_t step = 0;
        for (int32_t a = gs; a <= ge; ++a) {
          const int32_t row = items[a].id;
          if (row == rep || !keep_row[row]) continue;
          double r = 0.0;
          if (!row_ratio(&lp->A, rep, row, keep_col, ctol, &r)) continue;
          found = 1;
          double sl, su;
          scaled_interval(AL[row], AU[row], r, &sl, &su);
          if (or_cmax(rl, sl) > or_cmin(ru, su) + itol) {
            status = 1;
            abandoned = 1;
            break;
          }
          ++step;
          prow_rec* rc = &recs[nrec++];
          rc->step = step;
          rc->kept = rep;
          rc->deleted = row;
          rc->ratio = r;
          rc->kal = rl;
          rc->kau = ru;
          rc->dal = AL[row];
          rc->dau = AU[row];
          rl = or_cmax(rl, sl);
          ru = or_cmin(ru, su);
          row_delete[row] = 1;
        }
        if (found && !abandoned) {
          AL[rep] = rl;
          AU[rep] = ru;
        }
      }
    }
    gs = next;
  }
  if (status) {
    plan->has_infeasible = 1;
  } else {
    int32_t changed = 0;
    for (int32_t i = 0; i < m; ++i)
      if (row_delete[i]) {
        plan->keep_row_mask[i] = 0;
        ++changed;
      }
    if (changed) {
      if (tape_on(plan, p)) {
        qsort(recs, (size_t)nrec, sizeof(prow_rec), prow_rec_cmp);
        or_tape* t = (or_tape*)plan->tape;
        for (int32_t r = 0; r < nrec; ++r) {
          const int32_t ix[2] = {recs[r].kept, recs[r].deleted};
          const double vx[5] = {recs[r].ratio, recs[r].kal, recs[r].kau, recs[r].dal, recs[r].dau};
          or_tape_push(t, GP_RED_PARALLEL_ROW, GP_DUAL_EXACT, ix, 2, vx, 5);
        }
      }
      plan->has_row_action = 1;
      plan->has_change = 1;
    }
  }
  free(items);
  free(row_delete);
  free(recs);
  return 0;
}

/* ======================================================================
 * parallel columns  (rule_parallel_columns.cu:189-612, 636-1127)
 * ====================================================================== */
static int col_live(const gp_csr* AT, gp_offset q, const uint8_t* keep_row, double zt) {
  return keep_row[AT->col_idx[q]] && fabs(AT->vals[q]) > zt;
}

/* _kernel_parallel_column_hashes (:189-232) */
static uint64_t col_hash(const gp_csr* AT, int32_t j, const uint8_t* keep_col, const uint8_t* keep_row,
                         double tol) {
  if (!keep_col[j]) return (uint64_t)(int64_t)j;
  int32_t len = 0;
  double pivot = 0.0;
  int set = 0;
  for (gp_offset q = AT->row_ptr[j]; q < AT->row_ptr[j + 1]; ++q) {
    if (!col_live(AT, q, keep_row, tol)) continue;
    ++len;
    if (!set) {
      pivot = AT->vals[q];
      set = 1;
    }
  }
  if (!set || len == 0) return fnv_mix(FNV_OFFSET, (uint64_t)(int64_t)j);
  uint64_t h = fnv_mix(FNV_OFFSET, (uint64_t)(int64_t)len);
  for (gp_offset q = AT->row_ptr[j]; q < AT->row_ptr[j + 1]; ++q) {
    if (!col_live(AT, q, keep_row, tol)) continue;
    h = fnv_mix(h, (uint64_t)(int64_t)AT->col_idx[q]);
    h = fnv_mix(h, dbits(AT->vals[q] / pivot));
  }
  return h;
}

static gp_offset next_live(const gp_csr* AT, gp_offset q, gp_offset stop, const uint8_t* keep_row, double zt) {
  while (q < stop && !col_live(AT, q, keep_row, zt)) ++q;
  return q;
}

/* _parallel_column_ratio (:252-304): ratio = a_k / a_j, entries compared at
 * their own scale. */
static int col_ratio(const gp_csr* AT, int32_t j, int32_t k, const uint8_t* keep_row, double zt,
                     double ctol, double* out) {
  gp_offset pj = AT->row_ptr[j], pk = AT->row_ptr[k];
  const gp_offset sj = AT->row_ptr[j + 1], sk = AT->row_ptr[k + 1];
  double ratio = 0.0;
  int set = 0;
  for (;;) {
    pj = next_live(AT, pj, sj, keep_row, zt);
    pk = next_live(AT, pk, sk, keep_row, zt);
    if (pj >= sj || pk >= sk) break;
    if (AT->col_idx[pj] != AT->col_idx[pk]) return 0;
    const double aj = AT->vals[pj], ak = AT->vals[pk];
    if (!set) {
      if (fabs(aj) <= zt) return 0;
      ratio = ak / aj;
      if (fabs(ratio) <= zt || !isfinite(ratio)) return 0;
      set = 1;
    }
    const double expected = ratio * aj;
    const double scale = fmax(fabs(ak), fabs(expected));
    if (!isfinite(expected) || fabs(ak - expected) > ctol * scale) return 0;
    ++pj;
    ++pk;
  }
  pj = next_live(AT, pj, sj, keep_row, zt);
  pk = next_live(AT, pk, sk, keep_row, zt);
  if (pj < sj || pk < sk || !set) return 0;
  *out = ratio;
  return 1;
}

static double merged_lower(double tl, double sl, double su, double r) {
  if (r > 0.0) {
    if (isfinite(tl) && isfinite(sl)) return tl + r * sl;
  } else if (isfinite(tl) && isfinite(su)) {
    return tl + r * su;
  }
  return -INFINITY;
}
static double merged_upper(double tu, double sl, double su, double r) {
  if (r > 0.0) {
    if (isfinite(tu) && isfinite(su)) return tu + r * su;
  } else if (isfinite(tu) && isfinite(sl)) {
    return tu + r * sl;
  }
  return INFINITY;
}

static int ratio_out_of_guard(double ratio, double max_abs) {
  const double ar = fabs(ratio);
  return max_abs < 1.0 || ar > max_abs || ar * max_abs < 1.0;
}

int or_rule_parallel_columns(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                             const gp_params* p) {
  (void)s;
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  const int32_t m = lp->A.rows, n = lp->A.cols;
  if (n <= 1) return 0;
  const gp_csr* AT = &lp->AT;
  const uint8_t* keep_col = plan->keep_col_mask;
  const uint8_t* keep_row = plan->keep_row_mask;
  const double zt = p->zero_tol;
  const double ctol = fmax(p->zero_tol, p->bound_tol), otol = ctol;
  double* c = plan->new_c;
  double* l = plan->new_l;
  double* u = plan->new_u;
  hash_item* items = (hash_item*)or_xmalloc(sizeof(hash_item) * (size_t)n);
  for (int32_t j = 0; j < n; ++j) {
    items[j].h = col_hash(AT, j, keep_col, keep_row, zt);
    items[j].id = j;
  }
  qsort(items, (size_t)n, sizeof(hash_item), hash_item_cmp);
  /* scale guard (:347-399) */
  double active_max = p->parallel_columns_max_abs_ratio;
  const int guard = isfinite(p->parallel_columns_max_abs_ratio) && p->parallel_columns_scale_guard_min_groups > 1;
  if (guard) {
    int32_t groups = 0;
    for (int32_t gs = 0; gs < n && groups < p->parallel_columns_scale_guard_min_groups;) {
      int32_t ge = gs;
      while (ge + 1 < n && items[ge + 1].h == items[gs].h) ++ge;
      const int32_t size = ge - gs + 1;
      if (size >= 2 && size <= p->parallel_columns_scale_guard_max_group_size) {
        int hit = 0;
        for (int32_t a = gs; a < ge && !hit; ++a) {
          const int32_t j = items[a].id;
          if (!keep_col[j]) continue;
          for (int32_t b = a + 1; b <= ge; ++b) {
            const int32_t k = items[b].id;
            if (!keep_col[k]) continue;
            double r = 0.0;
            if (!col_ratio(AT, j, k, keep_row, zt, ctol, &r)) continue;
            if (fabs(c[k] - r * c[j]) > otol) continue;
            if (ratio_out_of_guard(r, p->parallel_columns_max_abs_ratio)) {
              hit = 1;
              break;
            }
          }
        }
        groups += hit;
      }
      gs = ge + 1;
    }
    if (groups < p->parallel_columns_scale_guard_min_groups) active_max = INFINITY;
  } else {
    active_max = p->parallel_columns_max_abs_ratio;
  }

  uint8_t* col_delete = (uint8_t*)or_xcalloc((size_t)n, 1);
  uint8_t* fixed = (uint8_t*)or_xcalloc((size_t)n, 1);
  double* fval = (double*)or_xcalloc((size_t)n, sizeof(double));
  int32_t* merge_to = (int32_t*)or_xcalloc((size_t)n, sizeof(int32_t));
  double* mrec = (double*)or_xcalloc((size_t)n * 5, sizeof(double));
  int status0 = 0, status1 = 0, status2 = 0;
  /* _kernel_parallel_column_groups (:400-571) */
  for (int32_t gs = 0; gs < n;) {
    int32_t ge = gs;
    while (ge + 1 < n && items[ge + 1].h == items[gs].h) ++ge;
    const int32_t next = ge + 1;
    int abandon = 0;
    for (int32_t a = gs; a < ge && !abandon; ++a) {
      const int32_t j = items[a].id;
      if (!keep_col[j] || col_delete[j] || fixed[j]) continue;
      for (int32_t b = a + 1; b <= ge; ++b) {
        const int32_t k = items[b].id;
        if (!keep_col[k] || col_delete[k] || fixed[k]) continue;
        if (isfinite(l[j]) && isfinite(u[j]) && isfinite(l[k]) && isfinite(u[k])) {
          const gp_offset fj = next_live(AT, AT->row_ptr[j], AT->row_ptr[j + 1], keep_row, zt);
          const gp_offset fk = next_live(AT, AT->row_ptr[k], AT->row_ptr[k + 1], keep_row, zt);
          if (fj < AT->row_ptr[j + 1] && fk < AT->row_ptr[k + 1]) {
            const double fr = AT->vals[fk] / AT->vals[fj];
            if (fabs(c[k] - fr * c[j]) > otol) continue;
          }
        }
        double r = 0.0;
        if (!col_ratio(AT, j, k, keep_row, zt, ctol, &r)) continue;
        if (isfinite(active_max) && ratio_out_of_guard(r, active_max)) continue;
        const double gap = c[k] - r * c[j];
        const double tl = l[j], tu = u[j], sl = l[k], su = u[k];
        if (fabs(gap) <= otol) {
          merge_to[k] = j;
          double* mr = mrec + (size_t)k * 5;
          mr[0] = r;
          mr[1] = sl;
          mr[2] = su;
          mr[3] = tl;
          mr[4] = tu;
          l[j] = merged_lower(tl, sl, su, r);
          u[j] = merged_upper(tu, sl, su, r);
          col_delete[k] = 1;
          status1 = 1;
          continue;
        }
        int xk_lo = 0, xk_up = 0, xj_lo = 0, xj_up = 0;
        if (gap > otol) {
          if (r > 0.0) {
            xk_lo = !isfinite(tu);
            xj_up = !isfinite(sl);
          } else {
            xk_lo = !isfinite(tl);
            xj_lo = !isfinite(sl);
          }
        } else {
          if (r > 0.0) {
            xk_up = !isfinite(tl);
            xj_lo = !isfinite(su);
          } else {
            xk_up = !isfinite(tu);
            xj_up = !isfinite(su);
          }
        }
        if (xk_lo || xk_up) {
          const double v = xk_lo ? sl : su;
          if (!isfinite(v)) {
            status0 = 1;
            abandon = 1;
            break;
          }
          fixed[k] = 1;
          fval[k] = v;
          status2 = 1;
          continue;
        }
        if (xj_lo || xj_up) {
          const double v = xj_lo ? tl : tu;
          if (!isfinite(v)) {
            status0 = 1;
            abandon = 1;
            break;
          }
          fixed[j] = 1;
          fval[j] = v;
          status2 = 1;
          break;
        }
      }
    }
    gs = next;
  }
  if (status0) {
    plan->has_unbounded = 1;
  } else if (status1 || status2) {
    if (status1 && tape_on(plan, p)) {
      or_tape* t = (or_tape*)plan->tape;
      for (int32_t k = 0; k < n; ++k) {
        if (!col_delete[k]) continue;
        const int32_t ix[2] = {k, merge_to[k]};
        or_tape_push(t, GP_RED_PARALLEL_COLUMN, GP_DUAL_MINIMAL, ix, 2, mrec + (size_t)k * 5, 5);
      }
    }
    double* shift = (double*)or_xcalloc((size_t)(m ? m : 1), sizeof(double));
    if (status2) {
      if (tape_on(plan, p)) push_fixed_with_rows((or_tape*)plan->tape, fixed, fval, keep_row, c, AT);
      row_shift_by_columns(&lp->A, keep_row, fixed, fval, shift);
      plan->has_row_action = 1;
    }
    double* contrib = (double*)or_xcalloc((size_t)n, sizeof(double));
    for (int32_t j = 0; j < n; ++j) {
      if (!col_delete[j] && !fixed[j]) continue;
      plan->keep_col_mask[j] = 0;
      if (fixed[j]) {
        l[j] = fval[j];
        u[j] = fval[j];
        contrib[j] = c[j] * fval[j];
      }
    }
    for (int32_t i = 0; i < m; ++i) {
      plan->new_AL[i] -= shift[i];
      plan->new_AU[i] -= shift[i];
    }
    plan->obj_constant_delta += or_det_sum(contrib, n);
    plan->has_col_action = 1;
    plan->has_change = 1;
    free(shift);
    free(contrib);
  }
  free(items);
  free(col_delete);
  free(fixed);
  free(fval);
  free(merge_to);
  free(mrec);
  return 0;
}
