/*
 * oracle_rules.c -- CPU restatement of the general presolve rules.
 * TEST INFRASTRUCTURE ONLY (parity checker for the HIP rules).
 *
 * Each rule follows the reference kernel sequence it names.  Two deliberate,
 * documented differences keep the result independent of execution order:
 *  - sums that the reference forms with atomicAdd / CUB reductions (row
 *    shifts, error budgets, objective deltas) use the fixed orders of
 *    or_rowacc (row sums, CSR order) and or_det_sum (global sums);
 *  - candidate bound max/min use the IEEE total order (-0 < +0).
 */
#include "oracle_internal.h"

#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int tape_on(const gp_plan* plan, const gp_params* p) {
  return plan->tape != NULL && p->record_postsolve_tape;
}

/* ---- infeasible and fixed variables -------------------------------------
 * rule_infeasible_and_fixed_variables.cu:65-348 */
int or_rule_fixed_variables(gp_plan* plan, const gp_device_lp* lp, const gp_params* p) {
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  const int32_t n = lp->A.cols, m = lp->A.rows;
  if (n == 0) return 0;
  uint8_t* fixed = (uint8_t*)or_xcalloc((size_t)n, 1);
  double* val = (double*)or_xcalloc((size_t)n, sizeof(double));
  double* radius = (double*)or_xcalloc((size_t)n, sizeof(double));
  double* contrib = (double*)or_xcalloc((size_t)n, sizeof(double));
  double* row_error = (double*)or_xcalloc((size_t)(m ? m : 1), sizeof(double));
  double* row_shift = (double*)or_xcalloc((size_t)(m ? m : 1), sizeof(double));
  /* candidates (:65-113) */
  for (int32_t j = 0; j < n; ++j) {
    if (!plan->keep_col_mask[j]) continue;
    const double lj = plan->new_l[j], uj = plan->new_u[j];
    if (isfinite(lj) && isfinite(uj) && lj <= uj && uj - lj <= p->bound_tol) {
      const double v = lj + 0.5 * (uj - lj);
      fixed[j] = 1;
      val[j] = v;
      radius[j] = or_cmax(v - lj, uj - v);
      contrib[j] = plan->new_c[j] * v;
    }
  }
  /* per-row uncertainty budget: sum |a| * radius over candidate columns */
  for (int32_t i = 0; i < m; ++i) {
    if (!plan->keep_row_mask[i]) continue;
    or_rowacc acc;
    or_rowacc_init(&acc);
    const gp_offset p0 = lp->A.row_ptr[i];
    for (gp_offset q = p0; q < lp->A.row_ptr[i + 1]; ++q) {
      or_rowacc_pos(&acc, q - p0);
      const int32_t j = lp->A.col_idx[q];
      if (fixed[j] && radius[j] > 0.0) or_rowacc_add(&acc, fabs(lp->A.vals[q]) * radius[j]);
    }
    row_error[i] = or_rowacc_value(&acc);
  }
  /* filter (:114-150): approximate fixings touching an over-budget row are dropped */
  const double budget = 0.1 * p->feasibility_tol;
  int32_t count = 0;
  for (int32_t j = 0; j < n; ++j) {
    if (!fixed[j]) continue;
    if (plan->new_l[j] != plan->new_u[j]) {
      for (gp_offset q = lp->AT.row_ptr[j]; q < lp->AT.row_ptr[j + 1]; ++q) {
        const int32_t r = lp->AT.col_idx[q];
        if (plan->keep_row_mask[r] && !(row_error[r] <= budget)) {
          fixed[j] = 0;
          contrib[j] = 0.0;
          break;
        }
      }
    }
    count += fixed[j];
  }
  if (count == 0) goto done;
  /* row shifts: sum a * v over accepted fixings, CSR order */
  for (int32_t i = 0; i < m; ++i) {
    if (!plan->keep_row_mask[i]) continue;
    or_rowacc acc;
    or_rowacc_init(&acc);
    const gp_offset p0 = lp->A.row_ptr[i];
    for (gp_offset q = p0; q < lp->A.row_ptr[i + 1]; ++q) {
      or_rowacc_pos(&acc, q - p0);
      const int32_t j = lp->A.col_idx[q];
      if (fixed[j]) or_rowacc_add(&acc, lp->A.vals[q] * val[j]);
    }
    row_shift[i] = or_rowacc_value(&acc);
  }
  if (tape_on(plan, p) && p->record_postsolve_tape_cpu) {
    or_tape* t = (or_tape*)plan->tape;
    for (int32_t j = 0; j < n; ++j) {
      if (!fixed[j]) continue;
      const gp_offset a0 = lp->AT.row_ptr[j], a1 = lp->AT.row_ptr[j + 1];
      int32_t* ix = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(a1 - a0 + 1));
      double* vx = (double*)or_xmalloc(sizeof(double) * (size_t)(a1 - a0 + 2));
      int64_t ni = 0, nv = 0;
      ix[ni++] = j;
      vx[nv++] = val[j];
      vx[nv++] = plan->new_c[j];
      for (gp_offset q = a0; q < a1; ++q) {
        const int32_t r = lp->AT.col_idx[q];
        if (!plan->keep_row_mask[r]) continue;
        ix[ni++] = r;
        vx[nv++] = lp->AT.vals[q];
      }
      or_tape_push(t, GP_RED_FIXED_VARIABLE, GP_DUAL_MINIMAL, ix, ni, vx, nv);
      free(ix);
      free(vx);
    }
  }
  /* apply (:151-176) */
  for (int32_t j = 0; j < n; ++j) {
    if (!fixed[j]) continue;
    plan->new_l[j] = val[j];
    plan->new_u[j] = val[j];
    plan->keep_col_mask[j] = 0;
  }
  for (int32_t i = 0; i < m; ++i) {
    plan->new_AL[i] -= row_shift[i];
    plan->new_AU[i] -= row_shift[i];
  }
  plan->obj_constant_delta += or_det_sum(contrib, n);
  plan->has_col_action = 1;
  plan->has_change = 1;
done:
  free(fixed);
  free(val);
  free(radius);
  free(contrib);
  free(row_error);
  free(row_shift);
  return 0;
}

/* ---- empty columns --------------------------------------------------------
 * empty_rows_and_columns/columns.cu:64-250 */
int or_rule_empty_columns(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                          const gp_params* p) {
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  const int32_t n = lp->A.cols;
  if (n == 0) return 0;
  uint8_t* fixed = (uint8_t*)or_xcalloc((size_t)n, 1);
  double* val = (double*)or_xcalloc((size_t)n, sizeof(double));
  double* contrib = (double*)or_xcalloc((size_t)n, sizeof(double));
  int infeasible = 0, unbounded = 0, any = 0;
  for (int32_t j = 0; j < n; ++j) {
    if (!(plan->keep_col_mask[j] && s->empty_col_mask[j])) continue;
    const double cj = plan->new_c[j], lj = plan->new_l[j], uj = plan->new_u[j];
    if (lj > uj + p->bound_tol) {
      infeasible = 1;
      continue;
    }
    if ((cj > p->zero_tol && !isfinite(lj)) || (cj < -p->zero_tol && !isfinite(uj))) {
      unbounded = 1;
      continue;
    }
    double v = 0.0;
    if (cj > p->zero_tol) v = lj;
    else if (cj < -p->zero_tol) v = uj;
    else if (isfinite(lj)) v = lj;
    else if (isfinite(uj)) v = uj;
    fixed[j] = 1;
    val[j] = v;
    contrib[j] = cj * v;
    any = 1;
  }
  if (infeasible) {
    plan->has_infeasible = 1;
  } else if (unbounded) {
    plan->has_unbounded = 1;
  } else if (any) {
    if (tape_on(plan, p) && p->record_postsolve_tape_cpu) {
      for (int32_t j = 0; j < n; ++j) {
        if (!fixed[j]) continue;
        const double vx[2] = {val[j], plan->new_c[j]};
        or_tape_push((or_tape*)plan->tape, GP_RED_FIXED_VARIABLE, GP_DUAL_MINIMAL, &j, 1, vx, 2);
      }
    }
    for (int32_t j = 0; j < n; ++j) {
      if (!fixed[j]) continue;
      plan->new_l[j] = val[j];
      plan->new_u[j] = val[j];
      plan->keep_col_mask[j] = 0;
    }
    plan->obj_constant_delta += or_det_sum(contrib, n);
    plan->has_col_action = 1;
    plan->has_change = 1;
  }
  free(fixed);
  free(val);
  free(contrib);
  return 0;
}

/* ---- empty rows -------------------------------------------------------------
 * empty_rows_and_columns/rows.cu:25-120 */
int or_rule_empty_rows(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s, const gp_params* p) {
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  const int32_t m = lp->A.rows;
  if (m == 0) return 0;
  const double tol = p->feasibility_tol;
  int infeasible = 0, removable = 0;
  for (int32_t i = 0; i < m; ++i) {
    if (!(plan->keep_row_mask[i] && s->row_nnz[i] == 0)) continue;
    if (plan->new_AL[i] <= tol && plan->new_AU[i] >= -tol) removable = 1;
    else infeasible = 1;
  }
  if (infeasible) {
    plan->has_infeasible = 1;
    return 0;
  }
  if (removable) {
    for (int32_t i = 0; i < m; ++i)
      if (plan->keep_row_mask[i] && s->row_nnz[i] == 0 && plan->new_AL[i] <= tol &&
          plan->new_AU[i] >= -tol)
        plan->keep_row_mask[i] = 0;
    plan->has_row_action = 1;
    plan->has_row_deletion = 1;
    plan->has_change = 1;
  }
  return 0;
}

/* ---- singleton rows ---------------------------------------------------------
 * rule_singleton_rows.cu:76-278 */
int or_rule_singleton_rows(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                           const gp_params* p) {
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  const int32_t m = lp->A.rows, n = lp->A.cols;
  if (m == 0) return 0;
  double* cl = (double*)or_xmalloc(sizeof(double) * (size_t)(n ? n : 1));
  double* cu = (double*)or_xmalloc(sizeof(double) * (size_t)(n ? n : 1));
  uint8_t* remove = (uint8_t*)or_xcalloc((size_t)m, 1);
  if (n) {
    memcpy(cl, plan->new_l, sizeof(double) * (size_t)n);
    memcpy(cu, plan->new_u, sizeof(double) * (size_t)n);
  }
  for (int32_t i = 0; i < m; ++i) {
    if (!(plan->keep_row_mask[i] && s->singleton_row_mask[i])) continue;
    const int32_t j = s->singleton_row_col[i];
    const double a = s->singleton_row_val[i];
    if (j < 0 || j >= n || !(fabs(a) > p->zero_tol)) continue;
    const double lo = plan->new_AL[i], hi = plan->new_AU[i];
    const double il = a > 0.0 ? lo / a : hi / a;
    const double iu = a > 0.0 ? hi / a : lo / a;
    or_cand_max(&cl[j], il);
    or_cand_min(&cu[j], iu);
    remove[i] = 1;
  }
  int infeasible = 0, changed_bounds = 0, removed = 0;
  for (int32_t j = 0; j < n; ++j) {
    if (cl[j] > cu[j] + p->bound_tol) {
      infeasible = 1;
      if (getenv("GP_ORACLE_TRACE_RULES")) {
        fprintf(stderr, "[oracle] singleton rows: col %d implied [%.17g, %.17g] bounds [%.17g, %.17g]\n", j, cl[j],
                cu[j], plan->new_l[j], plan->new_u[j]);
        for (int32_t i = 0; i < m; ++i)
          if (remove[i] && s->singleton_row_col[i] == j)
            fprintf(stderr, "[oracle]   row %d a %.17g AL %.17g AU %.17g\n", i, s->singleton_row_val[i],
                    plan->new_AL[i], plan->new_AU[i]);
      }
    }
    if (cl[j] != plan->new_l[j] || cu[j] != plan->new_u[j]) changed_bounds = 1;
  }
  for (int32_t i = 0; i < m; ++i) removed |= remove[i];
  if (infeasible) {
    plan->has_infeasible = 1;
  } else if (removed || changed_bounds) {
    if (tape_on(plan, p) && p->record_postsolve_tape_cpu && removed) {
      for (int32_t i = 0; i < m; ++i) {
        if (!remove[i]) continue;
        const int32_t ix[2] = {i, s->singleton_row_col[i]};
        const double vx[3] = {plan->new_AL[i], plan->new_AU[i], s->singleton_row_val[i]};
        or_tape_push((or_tape*)plan->tape, GP_RED_DELETED_ROW, GP_DUAL_MINIMAL, ix, 2, vx, 3);
      }
    }
    for (int32_t i = 0; i < m; ++i)
      if (remove[i]) plan->keep_row_mask[i] = 0;
    if (n) {
      memcpy(plan->new_l, cl, sizeof(double) * (size_t)n);
      memcpy(plan->new_u, cu, sizeof(double) * (size_t)n);
    }
    plan->has_row_action = 1;
    plan->has_row_deletion = plan->has_row_deletion || removed;
    plan->has_change = 1;
  }
  free(cl);
  free(cu);
  free(remove);
  return 0;
}

/* ---- infeasible and redundant rows -----------------------------------------
 * rule_infeasible_and_redundant_rows.cu:46-332.  Row min/max follow IEEE sum
 * semantics (any NaN term or +inf/-inf mix gives NaN) with the finite part
 * accumulated in the fixed row-sum order. */
typedef struct or_activity {
  or_rowacc fin;
  int32_t ninf, pinf, nan;
} or_activity;

static double activity_value(or_activity* a) {
  if (a->nan || (a->ninf && a->pinf)) return NAN;
  if (a->ninf) return -INFINITY;
  if (a->pinf) return INFINITY;
  return or_rowacc_value(&a->fin);
}

static void activity_add(or_activity* a, double t) {
  if (isfinite(t)) or_rowacc_add(&a->fin, t);
  else if (isnan(t)) a->nan++;
  else if (t < 0.0) a->ninf++;
  else a->pinf++;
}

int or_rule_redundant_rows(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                           const gp_params* p) {
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  const int32_t m = lp->A.rows;
  if (m == 0) return 0;
  const double tol = p->bound_tol;
  const uint8_t* keep_col = s->live_masks ? plan->keep_col_mask : NULL;
  uint8_t* full = (uint8_t*)or_xcalloc((size_t)m, 1);
  uint8_t* dl = (uint8_t*)or_xcalloc((size_t)m, 1);
  uint8_t* du = (uint8_t*)or_xcalloc((size_t)m, 1);
  int f0 = 0, f1 = 0, f2 = 0, f3 = 0;
  for (int32_t i = 0; i < m; ++i) {
    if (!plan->keep_row_mask[i] || s->row_nnz[i] <= 1) continue;
    const double al = plan->new_AL[i], au = plan->new_AU[i];
    const int lf = isfinite(al), uf = isfinite(au);
    if (lf && uf && fabs(au - al) <= tol) continue;
    or_activity lo, hi;
    memset(&lo, 0, sizeof(lo));
    memset(&hi, 0, sizeof(hi));
    or_rowacc_init(&lo.fin);
    or_rowacc_init(&hi.fin);
    const gp_offset p0 = lp->A.row_ptr[i];
    for (gp_offset q = p0; q < lp->A.row_ptr[i + 1]; ++q) {
      or_rowacc_pos(&lo.fin, q - p0);
      or_rowacc_pos(&hi.fin, q - p0);
      const int32_t j = lp->A.col_idx[q];
      if (keep_col && !keep_col[j]) continue;
      const double a = lp->A.vals[q];
      const double lj = plan->new_l[j], uj = plan->new_u[j];
      activity_add(&lo, a >= 0.0 ? a * lj : a * uj);
      activity_add(&hi, a >= 0.0 ? a * uj : a * lj);
    }
    const double rmin = activity_value(&lo), rmax = activity_value(&hi);
    if ((lf && rmax < al - tol) || (uf && rmin > au + tol)) {
      f0 = 1;
      continue;
    }
    const int li = !lf || rmin >= al - tol;
    const int ui = !uf || rmax <= au + tol;
    if (li && ui) {
      full[i] = 1;
      f1 = 1;
    } else if (lf && li) {
      dl[i] = 1;
      f2 = 1;
    } else if (uf && ui) {
      du[i] = 1;
      f3 = 1;
    }
  }
  if (f0) {
    plan->has_infeasible = 1;
  } else if (f1 || f2 || f3) {
    if (tape_on(plan, p) && p->record_postsolve_tape_cpu) {
      or_tape* t = (or_tape*)plan->tape;
      for (int32_t i = 0; i < m; ++i) {
        if (!full[i]) continue;
        const double vx[2] = {plan->new_AL[i], plan->new_AU[i]};
        or_tape_push(t, GP_RED_DELETED_ROW, GP_DUAL_MINIMAL, &i, 1, vx, 2);
      }
      for (int32_t i = 0; i < m; ++i) {
        const double al = plan->new_AL[i], au = plan->new_AU[i];
        if (dl[i]) {
          const double vx[4] = {al, au, -INFINITY, au};
          or_tape_push(t, GP_RED_ROW_LOWER_BOUND_CHANGE, GP_DUAL_MINIMAL, &i, 1, vx, 4);
        }
        if (du[i]) {
          const double vx[4] = {al, au, al, INFINITY};
          or_tape_push(t, GP_RED_ROW_UPPER_BOUND_CHANGE, GP_DUAL_MINIMAL, &i, 1, vx, 4);
        }
      }
    }
    for (int32_t i = 0; i < m; ++i) {
      if (full[i]) plan->keep_row_mask[i] = 0;
      if (dl[i]) plan->new_AL[i] = -INFINITY;
      if (du[i]) plan->new_AU[i] = INFINITY;
    }
    plan->has_row_action = 1;
    plan->has_redundant_row_deletion = plan->has_redundant_row_deletion || f1;
    plan->has_row_deletion = plan->has_row_deletion || f1;
    plan->has_change = 1;
  }
  free(full);
  free(dl);
  free(du);
  return 0;
}

/* ---- bound strengthening ------------------------------------------------------
 * rule_bound_strengthening.cu:349-1557 */
static double residual_min(double fin, int32_t ninf, double term) {
  if (isfinite(term)) return ninf == 0 ? fin - term : -INFINITY;
  if (term < 0.0) return ninf > 1 ? -INFINITY : fin;
  return fin;
}

static double residual_max(double fin, int32_t pinf, double term) {
  if (isfinite(term)) return pinf == 0 ? fin - term : INFINITY;
  if (term > 0.0) return pinf > 1 ? INFINITY : fin;
  return fin;
}

/* implied column bounds from one row entry (:405-445) */
static void implied_bounds(double a, double lj, double uj, double al, double au, double min_fin,
                           int32_t min_ninf, double max_fin, int32_t max_pinf, double* il,
                           double* iu) {
  const double tmin = a >= 0.0 ? a * lj : a * uj;
  const double tmax = a >= 0.0 ? a * uj : a * lj;
  const double rmin = residual_min(min_fin, min_ninf, tmin);
  const double rmax = residual_max(max_fin, max_pinf, tmax);
  *il = -INFINITY;
  *iu = INFINITY;
  if (a > 0.0) {
    if (isfinite(al) && isfinite(rmax)) *il = (al - rmax) / a;
    if (isfinite(au) && isfinite(rmin)) *iu = (au - rmin) / a;
  } else {
    if (isfinite(au) && isfinite(rmin)) *il = (au - rmin) / a;
    if (isfinite(al) && isfinite(rmax)) *iu = (al - rmax) / a;
  }
}

int or_rule_bound_strengthening(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                                const gp_params* p) {
  if (plan->has_infeasible || plan->has_unbounded) return 0;
  const int32_t m = lp->A.rows, n = lp->A.cols;
  if (m == 0 || n == 0) return 0;
  const uint8_t* keep_col = s->live_masks ? plan->keep_col_mask : NULL;
  const uint8_t* keep_row = plan->keep_row_mask;
  double* min_fin = (double*)or_xmalloc(sizeof(double) * (size_t)m);
  double* max_fin = (double*)or_xmalloc(sizeof(double) * (size_t)m);
  int32_t* min_ninf = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)m);
  int32_t* max_pinf = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)m);
  double* cl = (double*)or_xmalloc(sizeof(double) * (size_t)n);
  double* cu = (double*)or_xmalloc(sizeof(double) * (size_t)n);
  double* fl = (double*)or_xmalloc(sizeof(double) * (size_t)n);
  double* fu = (double*)or_xmalloc(sizeof(double) * (size_t)n);
  double* vmax = (double*)or_xcalloc((size_t)n, sizeof(double));
  uint8_t* lch = (uint8_t*)or_xcalloc((size_t)n, 1);
  uint8_t* uch = (uint8_t*)or_xcalloc((size_t)n, 1);
  uint8_t* fixed = (uint8_t*)or_xcalloc((size_t)n, 1);
  double* fval = (double*)or_xcalloc((size_t)n, sizeof(double));
  memcpy(cl, plan->new_l, sizeof(double) * (size_t)n);
  memcpy(cu, plan->new_u, sizeof(double) * (size_t)n);
  const double* l = plan->new_l;
  const double* u = plan->new_u;
  or_row_activity_summary(&lp->A, l, u, p->zero_tol, keep_row, keep_col, min_fin, max_fin, min_ninf,
                          max_pinf);
  /* candidates (:373-456) */
  for (int32_t i = 0; i < m; ++i) {
    if (!keep_row[i] || s->row_nnz[i] <= 1) continue;
    for (gp_offset q = lp->A.row_ptr[i]; q < lp->A.row_ptr[i + 1]; ++q) {
      const int32_t j = lp->A.col_idx[q];
      if (keep_col && !keep_col[j]) continue;
      const double a = lp->A.vals[q];
      if (fabs(a) <= p->zero_tol) continue;
      double il, iu;
      implied_bounds(a, l[j], u[j], plan->new_AL[i], plan->new_AU[i], min_fin[i], min_ninf[i],
                     max_fin[i], max_pinf[i], &il, &iu);
      double nl = l[j], nu = u[j];
      if (isfinite(il)) nl = or_cmax(nl, il);
      if (isfinite(iu)) nu = or_cmin(nu, iu);
      if (nl > l[j]) or_cand_max(&cl[j], nl);
      if (nu < u[j]) or_cand_min(&cu[j], nu);
    }
  }
  /* column max |a| over live rows (gpu_presolve_kernels.cu:150-167) */
  for (int32_t j = 0; j < n; ++j) {
    double v = 0.0;
    for (gp_offset q = lp->AT.row_ptr[j]; q < lp->AT.row_ptr[j + 1]; ++q) {
      if (s->live_masks && !keep_row[lp->AT.col_idx[q]]) continue;
      v = or_cmax(v, fabs(lp->AT.vals[q]));
    }
    vmax[j] = v;
  }
  /* finalize (:646-796) */
  const double* ref_l = plan->propagation_reference_l ? plan->propagation_reference_l
                                                      : (lp->l ? lp->l : plan->new_l);
  const double* ref_u = plan->propagation_reference_u ? plan->propagation_reference_u
                                                      : (lp->u ? lp->u : plan->new_u);
  const double ft = p->feasibility_tol;
  int f0 = 0, f1 = 0, f2 = 0;
  for (int32_t j = 0; j < n && !f0; ++j) {
    if (!plan->keep_col_mask[j]) {
      fl[j] = l[j];
      fu[j] = u[j];
      continue;
    }
    const double lj = l[j], uj = u[j];
    double cand_l = cl[j], cand_u = cu[j];
    double new_l = lj, new_u = uj;
    int lc = 0, uc = 0, fx = 0;
    double fat = 0.0;
    double el = lj, eu = uj;
    if (isfinite(cand_l)) el = or_cmax(el, cand_l);
    if (isfinite(cand_u)) eu = or_cmin(eu, cand_u);
    if (el > eu + ft) {
      f0 = 1;
      break;
    }
    const int was_free = !isfinite(ref_l[j]) && !isfinite(ref_u[j]);
    const int box = isfinite(cand_l) && isfinite(cand_u);
    const int zero_cost = fabs(plan->new_c[j]) <= p->zero_tol;
    const int nearly_fixed = box && (cand_u <= cand_l || (cand_u - cand_l) * vmax[j] <= ft);
    if (p->bound_strengthening_preserve_free_zero_cost && was_free && box && zero_cost &&
        !nearly_fixed) {
      cand_l = lj;
      cand_u = uj;
    }
    if (isfinite(cand_l) && cand_l > lj && isfinite(uj)) {
      if (cand_l >= uj + ft) {
        f0 = 1;
        break;
      }
      if (cand_l >= uj || (uj - cand_l) * vmax[j] <= ft) {
        fx = 1;
        fat = uj;
        new_l = uj;
        new_u = uj;
      }
    }
    if (!fx && isfinite(cand_l) && cand_l > lj) {
      const int tighten = !isfinite(lj) || ((cand_l - lj > ft * 1.0e4) && (cand_l - lj > 1.0e-2 * fabs(lj)));
      if (tighten) {
        if (cand_l != nearbyint(cand_l)) cand_l -= 0.5 * ft * fabs(cand_l);
        new_l = cand_l;
        lc = 1;
      }
    }
    if (!fx && isfinite(cand_u) && cand_u < uj && isfinite(new_l)) {
      if (cand_u <= new_l - ft) {
        f0 = 1;
        break;
      }
      if (cand_u <= new_l || (cand_u - new_l) * vmax[j] <= ft) {
        fx = 1;
        fat = new_l;
        new_u = new_l;
      }
    }
    if (!fx && isfinite(cand_u) && cand_u < uj) {
      const int tighten = !isfinite(uj) || ((uj - cand_u > ft * 1.0e4) && (uj - cand_u > 1.0e-2 * fabs(uj)));
      if (tighten) {
        if (cand_u != nearbyint(cand_u))
          cand_u += p->bound_strengthening_upper_relaxation_factor * ft * fabs(cand_u);
        new_u = cand_u;
        uc = 1;
      }
    }
    if (fx) {
      lc = uc = 0;
      f2 = 1;
    }
    if (lc || uc) f1 = 1;
    fl[j] = new_l;
    fu[j] = new_u;
    lch[j] = (uint8_t)lc;
    uch[j] = (uint8_t)uc;
    fixed[j] = (uint8_t)fx;
    fval[j] = fat;
  }
  if (f0) {
    plan->has_infeasible = 1;
    goto done;
  }
  if (f1 || f2) {
    if (tape_on(plan, p) && f1) {
      /* support rows (:813-903) and implied-bound records (:904-1081) */
      const double support_tol = or_cmax(10.0 * p->bound_tol, 1.0e-10);
      or_tape* t = (or_tape*)plan->tape;
      for (int32_t j = 0; j < n; ++j) {
        if (!lch[j] && !uch[j]) continue;
        int32_t sl = INT_MAX, su = INT_MAX;
        for (gp_offset q = lp->AT.row_ptr[j]; q < lp->AT.row_ptr[j + 1]; ++q) {
          const int32_t i = lp->AT.col_idx[q];
          if (!keep_row[i] || s->row_nnz[i] <= 1) continue;
          const double a = lp->AT.vals[q];
          if (fabs(a) <= p->zero_tol) continue;
          double il, iu;
          implied_bounds(a, l[j], u[j], plan->new_AL[i], plan->new_AU[i], min_fin[i], min_ninf[i],
                         max_fin[i], max_pinf[i], &il, &iu);
          if (lch[j] && fl[j] > l[j] + p->bound_tol && isfinite(il) && fabs(il - fl[j]) <= support_tol)
            sl = i < sl ? i : sl;
          if (uch[j] && fu[j] < u[j] - p->bound_tol && isfinite(iu) && fabs(iu - fu[j]) <= support_tol)
            su = i < su ? i : su;
        }
        double cur_l = l[j];
        const double cur_u = u[j];
        if (lch[j]) {
          const double vx[4] = {cur_l, cur_u, fl[j], cur_u};
          const int32_t ix[2] = {j, sl};
          if (sl == INT_MAX) or_tape_push(t, GP_RED_VARIABLE_BOUND_CHANGE, GP_DUAL_MINIMAL, ix, 1, vx, 4);
          else or_tape_push(t, GP_RED_ROW_IMPLIED_BOUND_CHANGE, GP_DUAL_MINIMAL, ix, 2, vx, 4);
          cur_l = fl[j];
        }
        if (uch[j]) {
          const double vx[4] = {cur_l, cur_u, cur_l, fu[j]};
          const int32_t ix[2] = {j, su};
          if (su == INT_MAX) or_tape_push(t, GP_RED_VARIABLE_BOUND_CHANGE, GP_DUAL_MINIMAL, ix, 1, vx, 4);
          else or_tape_push(t, GP_RED_ROW_IMPLIED_BOUND_CHANGE, GP_DUAL_MINIMAL, ix, 2, vx, 4);
        }
      }
    }
    /* apply (:797-812) */
    for (int32_t j = 0; j < n; ++j) {
      plan->new_l[j] = fl[j];
      plan->new_u[j] = fu[j];
      if (fixed[j]) plan->keep_col_mask[j] = 0;
    }
    int f3 = 0;
    if (f2) {
      /* fixed-column row shift, then drop rows left empty (:1082-1160) */
      for (int32_t i = 0; i < m; ++i) {
        if (!keep_row[i]) continue;
        or_rowacc acc;
        or_rowacc_init(&acc);
        int32_t live = 0;
        const gp_offset p0 = lp->A.row_ptr[i];
        for (gp_offset q = p0; q < lp->A.row_ptr[i + 1]; ++q) {
          or_rowacc_pos(&acc, q - p0);
          const int32_t j = lp->A.col_idx[q];
          if (fixed[j]) or_rowacc_add(&acc, lp->A.vals[q] * fval[j]);
          live += plan->keep_col_mask[j] != 0;
        }
        const double shift = or_rowacc_value(&acc);
        plan->new_AL[i] -= shift;
        plan->new_AU[i] -= shift;
        if (live == 0) {
          if (plan->new_AL[i] <= ft && plan->new_AU[i] >= -ft) plan->keep_row_mask[i] = 0;
          else f3 = 1;
        }
      }
    }
    if (f3) {
      plan->has_infeasible = 1;
      goto done;
    }
    double obj = 0.0;
    if (f2) {
      /* sequential long-double sum in column order (rule_bound_strengthening.cu:1537-1548,
       * fixed_variable_tape.cu:305-325) */
      obj = or_fixed_objective_ld(fixed, NULL, plan->new_c, fval, n);
      if (tape_on(plan, p)) {
        or_tape* t = (or_tape*)plan->tape;
        for (int32_t j = 0; j < n; ++j) {
          if (!fixed[j]) continue;
          const gp_offset a0 = lp->AT.row_ptr[j], a1 = lp->AT.row_ptr[j + 1];
          int32_t* ix = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(a1 - a0 + 1));
          double* vx = (double*)or_xmalloc(sizeof(double) * (size_t)(a1 - a0 + 2));
          int64_t ni = 0, nv = 0;
          ix[ni++] = j;
          vx[nv++] = fval[j];
          vx[nv++] = plan->new_c[j];
          for (gp_offset q = a0; q < a1; ++q) {
            const int32_t r = lp->AT.col_idx[q];
            if (!plan->keep_row_mask[r]) continue;
            ix[ni++] = r;
            vx[nv++] = lp->AT.vals[q];
          }
          or_tape_push(t, GP_RED_FIXED_VARIABLE, GP_DUAL_MINIMAL, ix, ni, vx, nv);
          free(ix);
          free(vx);
        }
      }
    }
    plan->obj_constant_delta += obj;
    plan->has_row_action = plan->has_row_action || f2;
    plan->has_col_action = plan->has_col_action || f2;
    plan->has_change = 1;
  }
done:
  free(min_fin);
  free(max_fin);
  free(min_ninf);
  free(max_pinf);
  free(cl);
  free(cu);
  free(fl);
  free(fu);
  free(vmax);
  free(lch);
  free(uch);
  free(fixed);
  free(fval);
  return 0;
}

/* ---- dispatch ---------------------------------------------------------------- */
int or_rule_apply(int32_t rule, gp_plan* plan, const gp_device_lp* lp, const gp_stats* stats,
                  const gp_params* params) {
  switch (rule) {
    case GP_RULE_FIXED_VARIABLES: return or_rule_fixed_variables(plan, lp, params);
    case GP_RULE_EMPTY_COLUMNS: return or_rule_empty_columns(plan, lp, stats, params);
    case GP_RULE_EMPTY_ROWS: return or_rule_empty_rows(plan, lp, stats, params);
    case GP_RULE_SINGLETON_ROWS: return or_rule_singleton_rows(plan, lp, stats, params);
    case GP_RULE_SINGLETON_COLUMNS: return or_rule_singleton_columns(plan, lp, stats, params);
    case GP_RULE_BOUND_STRENGTHENING: return or_rule_bound_strengthening(plan, lp, stats, params);
    case GP_RULE_REDUNDANT_ROWS: return or_rule_redundant_rows(plan, lp, stats, params);
    case GP_RULE_DUAL_FIX: return or_rule_dual_fix(plan, lp, stats, params);
    case GP_RULE_DOUBLETON_EQUATIONS: return or_rule_doubleton_equations(plan, lp, stats, params);
    case GP_RULE_PARALLEL_ROWS: return or_rule_parallel_rows(plan, lp, stats, params);
    case GP_RULE_PARALLEL_COLUMNS: return or_rule_parallel_columns(plan, lp, stats, params);
    default: return -1;
  }
}
