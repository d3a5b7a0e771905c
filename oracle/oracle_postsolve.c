/*
 * oracle_postsolve.c -- CPU restatement of the postsolve replay
 * (gpu_postsolve.cu:85-147,355-398,535-670 and postsolve/postsolve_<name>.cuh).
 * TEST INFRASTRUCTURE ONLY.
 */
#include "oracle_internal.h"

#include <stdlib.h>
#include <string.h>

typedef struct ps_state {
  double *x, *y, *z;
  uint8_t* retrieved;
  const gp_device_lp* model; /* original model or NULL */
  double tol;
} ps_state;

static int ps_finite(double x) { return x > -INFINITY && x < INFINITY; }
static int at_lower(double x, double l, double tol) { return ps_finite(l) && x <= l + tol; }
static int at_upper(double x, double u, double tol) { return ps_finite(u) && x >= u - tol; }
static double clamp_inf(double x, double lo, double hi) {
  double y = x;
  if (ps_finite(lo)) y = or_cmax(y, lo);
  if (ps_finite(hi)) y = or_cmin(y, hi);
  return y;
}

/* postsolve_fixed_variables.cuh:8-41 */
static void replay_fixed(ps_state* s, const int32_t* ix, int64_t i0, int64_t i1, const double* v,
                         int64_t v0, int primal, int dual) {
  const int32_t col = ix[i0];
  if (primal) s->x[col] = v[v0];
  if (dual) {
    double zj = v[v0 + 1];
    for (int64_t p = i0 + 1; p <= i1; ++p) zj -= v[v0 + 2 + (p - (i0 + 1))] * s->y[ix[p]];
    s->z[col] = zj;
    s->retrieved[col] = 1;
  }
}

/* postsolve_fixed_variables.cuh:42-90 */
static void replay_infinite(ps_state* s, const int32_t* ix, int64_t i0, const double* v, int64_t v0,
                            int primal, int dual) {
  const int pos = ix[i0] > 0;
  const int32_t col = ix[i0 + 1];
  const int32_t rows = ix[i0 + 2];
  double extreme = v[v0];
  int64_t ip = i0 + 3, vp = v0 + 1;
  if (primal) {
    for (int32_t r = 0; r < rows; ++r) {
      const int64_t len = ix[ip];
      double side = v[vp], coeff = 0.0;
      for (int64_t t = 0; t < len; ++t) {
        const int32_t c = ix[ip + 1 + t];
        const double a = v[vp + 1 + t];
        if (c == col) coeff = a;
        else side -= a * s->x[c];
      }
      const double cand = side / coeff;
      extreme = pos ? or_cmax(extreme, cand) : or_cmin(extreme, cand);
      ip += len + 1;
      vp += len + 1;
    }
    s->x[col] = extreme;
  }
  if (dual) {
    s->z[col] = 0.0;
    s->retrieved[col] = 1;
  }
}

/* postsolve_singleton_columns.cuh:8-51 */
static void replay_singleton_column(ps_state* s, const int32_t* ix, int64_t i0, const double* v,
                                    int64_t v0, int primal, int dual) {
  const int32_t ec = ix[i0], row = ix[i0 + 1], support = ix[i0 + 2];
  const double pivot = v[v0], rhs = v[v0 + 1], cost = v[v0 + 4];
  const int deleted = v[v0 + 5] != 0.0;
  double partial = 0.0;
  for (int32_t t = 0; t < support; ++t) partial += v[v0 + 6 + t] * s->x[ix[i0 + 3 + t]];
  if (primal) s->x[ec] = (rhs - partial) / pivot;
  if (dual && row >= 0) {
    if (deleted) s->y[row] = cost / pivot;
    else s->y[row] += cost / pivot;
  }
  if (dual) {
    s->z[ec] = deleted || row < 0 ? 0.0 : cost - pivot * s->y[row];
    s->retrieved[ec] = 1;
  }
}

/* postsolve_parallel_columns.cuh */
static double merged_source_value(double lo, double hi, double tol) {
  if (isfinite(lo) && isfinite(hi)) return 0.5 * (lo + hi);
  if (isfinite(lo)) {
    if (0.0 > lo + tol) return 0.0;
    return lo + or_cmax(1.0, fabs(lo) * or_cmax(tol, 1.0e-12));
  }
  if (isfinite(hi)) {
    if (0.0 < hi - tol) return 0.0;
    return hi - or_cmax(1.0, fabs(hi) * or_cmax(tol, 1.0e-12));
  }
  return 0.0;
}

static double agg_lower(double tl, double sl, double su, double r) {
  if (r > 0.0) {
    if (isfinite(tl) && isfinite(sl)) return tl + r * sl;
  } else if (isfinite(tl) && isfinite(su)) {
    return tl + r * su;
  }
  return -INFINITY;
}

static double agg_upper(double tu, double sl, double su, double r) {
  if (r > 0.0) {
    if (isfinite(tu) && isfinite(su)) return tu + r * su;
  } else if (isfinite(tu) && isfinite(sl)) {
    return tu + r * sl;
  }
  return INFINITY;
}

static void replay_parallel_column(ps_state* s, const int32_t* ix, int64_t i0, const double* v,
                                   int64_t v0, int primal, int dual) {
  const int32_t from = ix[i0], to = ix[i0 + 1];
  const double ratio = v[v0], lo0 = v[v0 + 1], hi0 = v[v0 + 2], tl = v[v0 + 3], tu = v[v0 + 4];
  const double tol = s->tol;
  if (primal) {
    const double yv = s->x[to];
    const double zto = s->z[to];
    double lo = lo0, hi = hi0;
    if (isfinite(tu)) {
      const double b = (yv - tu) / ratio;
      if (ratio > 0.0) lo = or_cmax(lo, b);
      else hi = or_cmin(hi, b);
    }
    if (isfinite(tl)) {
      const double b = (yv - tl) / ratio;
      if (ratio > 0.0) hi = or_cmin(hi, b);
      else lo = or_cmax(lo, b);
    }
    const double lower = agg_lower(tl, lo0, hi0, ratio);
    const double upper = agg_upper(tu, lo0, hi0, ratio);
    double xf;
    if (zto > tol && isfinite(lower) && fabs(yv - lower) <= tol) xf = ratio > 0.0 ? lo0 : hi0;
    else if (zto < -tol && isfinite(upper) && fabs(yv - upper) <= tol) xf = ratio > 0.0 ? hi0 : lo0;
    else xf = merged_source_value(lo, hi, tol);
    s->x[from] = xf;
    s->x[to] = yv - ratio * xf;
  }
  if (dual) {
    s->z[from] = ratio * s->z[to];
    s->retrieved[from] = 1;
  }
}

/* postsolve_doubleton_equations.cuh */
static void replay_doubleton(ps_state* s, const int32_t* ix, int64_t i0, int64_t i1, const double* v,
                             int64_t v0, int64_t v1, int primal, int dual) {
  const int32_t ec = ix[i0], kc = ix[i0 + 1], row = ix[i0 + 2];
  const double ecoef = v[v0], kcoef = v[v0 + 1], rhs = v[v0 + 2], eobj = v[v0 + 9];
  const double alpha = -kcoef / ecoef, beta = rhs / ecoef;
  const double tol = s->tol;
  if (primal) s->x[ec] = beta + alpha * s->x[kc];
  if (dual && s->model) {
    const double okl = v[v0 + 5], oku = v[v0 + 6], nkl = v[v0 + 7], nku = v[v0 + 8];
    const double lsrc = v1 >= v0 + 10 ? v[v0 + 10] : 0.0;
    const double usrc = v1 >= v0 + 11 ? v[v0 + 11] : 0.0;
    double zk = 0.0, ze = 0.0;
    const double xk = s->x[kc], zh = s->z[kc];
    if (fabs(alpha) <= tol) {
      zk = zh;
      ze = 0.0;
    } else {
      const int al = ps_finite(nkl) && xk <= nkl + tol;
      const int au = ps_finite(nku) && xk >= nku - tol;
      if (!al && !au) {
        zk = ze = 0.0;
      } else {
        double src = al ? lsrc : usrc;
        if (src == 0.0) {
          if (al) src = (ps_finite(okl) && fabs(nkl - okl) <= tol) ? 1.0 : 2.0;
          else src = (ps_finite(oku) && fabs(nku - oku) <= tol) ? 1.0 : 2.0;
        }
        if (src == 2.0) {
          zk = 0.0;
          ze = zh / alpha;
        } else {
          zk = zh;
          ze = 0.0;
        }
      }
    }
    const int tape_support = i1 >= i0 + 3 && v1 >= v0 + 11;
    double sum = 0.0;
    if (tape_support) {
      const int32_t sc = ix[i0 + 3];
      for (int32_t t = 0; t < sc; ++t) sum += v[v0 + 12 + t] * s->y[ix[i0 + 4 + t]];
      s->y[row] = (eobj - sum - ze) / ecoef;
    } else {
      const gp_csr* AT = &s->model->AT;
      for (gp_offset p = AT->row_ptr[ec]; p < AT->row_ptr[ec + 1]; ++p)
        if (AT->col_idx[p] != row) sum += AT->vals[p] * s->y[AT->col_idx[p]];
      s->y[row] = (s->model->c[ec] - sum - ze) / ecoef;
    }
    s->z[kc] = zk;
    s->z[ec] = ze;
    s->retrieved[kc] = 1;
    s->retrieved[ec] = 1;
  } else if (dual) {
    s->z[ec] = 0.0;
    s->retrieved[ec] = 1;
    if (row >= 0) s->y[row] = eobj / ecoef;
  }
}

/* postsolve_parallel_rows.cuh */
static int row_state(double act, double lo, double hi, double tol) {
  const int al = ps_finite(lo) && act <= lo + tol;
  const int au = ps_finite(hi) && act >= hi - tol;
  if (al && au) return 3;
  if (al) return 1;
  if (au) return 2;
  return 0;
}

static void state_interval(int st, double* lo, double* hi) {
  if (st == 3) { *lo = -INFINITY; *hi = INFINITY; }
  else if (st == 1) { *lo = 0.0; *hi = INFINITY; }
  else if (st == 2) { *lo = -INFINITY; *hi = 0.0; }
  else { *lo = 0.0; *hi = 0.0; }
}

static double row_activity(const gp_csr* A, int32_t row, const double* x) {
  double a = 0.0;
  for (gp_offset p = A->row_ptr[row]; p < A->row_ptr[row + 1]; ++p) a += A->vals[p] * x[A->col_idx[p]];
  return a;
}

static void replay_parallel_row(ps_state* s, const int32_t* ix, int64_t i0, const double* v,
                                int64_t v0) {
  const int32_t kept = ix[i0], del = ix[i0 + 1];
  const double ratio = v[v0], kal = v[v0 + 1], kau = v[v0 + 2], dal = v[v0 + 3], dau = v[v0 + 4];
  const double tol = s->tol;
  if (fabs(ratio) <= tol) return;
  const double rl = dal * ratio, ru = dau * ratio;
  const double sal = ratio >= 0.0 ? rl : ru, sau = ratio >= 0.0 ? ru : rl;
  if (sal <= kal + tol && kau <= sau + tol) {
    s->y[del] = 0.0;
    return;
  }
  if (!s->model) return;
  const double ap = row_activity(&s->model->A, kept, s->x);
  const double ad = row_activity(&s->model->A, del, s->x);
  const int kst = row_state(ap, kal, kau, tol);
  const int dst = row_state(ad, dal, dau, tol);
  double ydl, ydh, ypl, yph;
  state_interval(dst, &ydl, &ydh);
  const double pivot = ratio * s->y[kept];
  if (kst == 3) { ypl = -INFINITY; yph = INFINITY; }
  else if (kst == 0) { ypl = pivot; yph = pivot; }
  else if (kst == 1) {
    if (ratio > 0.0) { ypl = -INFINITY; yph = pivot; }
    else { ypl = pivot; yph = INFINITY; }
  } else {
    if (ratio > 0.0) { ypl = pivot; yph = INFINITY; }
    else { ypl = -INFINITY; yph = pivot; }
  }
  const double sl = or_cmax(ydl, ypl), sh = or_cmin(ydh, yph);
  if (sl <= sh + tol) {
    const double ydn = clamp_inf(0.0, sl, sh);
    s->y[kept] = s->y[kept] - ydn / ratio;
    s->y[del] = ydn;
    return;
  }
  /* documented dual split (parallel_row_doc_recovery = 1) */
  const double yb = s->y[kept];
  double yp = yb, yd = 0.0;
  int ok = 1;
  if (fabs(yb) <= tol) {
    yp = 0.0;
    yd = 0.0;
  } else {
    const double bl = or_cmax(kal, sal), bu = or_cmin(kau, sau);
    if (!(bl <= bu + tol)) {
      ok = 0;
    } else {
      const int lt = bl > kal + tol, ut = bu < kau - tol;
      if (lt && ut) ok = 0;
      else if (!lt && !ut) { yp = yb; yd = 0.0; }
      else if (yb > 0.0) {
        if (lt) { yp = 0.0; yd = ratio * yb; }
        else { yp = yb; yd = 0.0; }
      } else {
        if (lt) { yp = yb; yd = 0.0; }
        else { yp = 0.0; yd = ratio * yb; }
      }
    }
  }
  if (ok) {
    s->y[kept] = yp;
    s->y[del] = yd;
  }
}

/* postsolve_singleton_rows.cuh */
static void replay_singleton_row(ps_state* s, const int32_t* ix, int64_t i0, const double* v,
                                 int64_t v0) {
  const int32_t row = ix[i0], col = ix[i0 + 1];
  const double oal = v[v0], oau = v[v0 + 1], a = v[v0 + 2];
  const double tol = s->tol;
  if (fabs(a) <= tol) return;
  const gp_device_lp* M = s->model;
  double g = M->c[col];
  for (gp_offset p = M->AT.row_ptr[col]; p < M->AT.row_ptr[col + 1]; ++p)
    if (M->AT.col_idx[p] != row) g -= M->AT.vals[p] * s->y[M->AT.col_idx[p]];
  const double xj = s->x[col], zb = s->z[col];
  const double act = a * xj;
  const int rl = isfinite(oal) && act <= oal + tol;
  const int ru = isfinite(oau) && act >= oau - tol;
  double izl = -INFINITY, izh = INFINITY;
  if (!(rl && ru)) {
    if (rl) {
      if (a > 0.0) izh = g;
      else izl = g;
    } else if (ru) {
      if (a > 0.0) izl = g;
      else izh = g;
    } else {
      izl = g;
      izh = g;
    }
  }
  const int al = at_lower(xj, M->l[col], tol), au = at_upper(xj, M->u[col], tol);
  double zl = izl, zh = izh;
  if (!al && !au) {
    if ((isfinite(izl) && izl > tol) || (isfinite(izh) && izh < -tol)) {
      zl = g;
      zh = g;
    } else {
      zl = 0.0;
      zh = 0.0;
    }
  } else if (al && !au) {
    zl = or_cmax(zl, 0.0);
  } else if (!al && au) {
    zh = or_cmin(zh, 0.0);
  }
  if (isfinite(zl) && isfinite(zh) && zl > zh + tol) return;
  const double zn = clamp_inf(zb, zl, zh);
  s->z[col] = zn;
  s->y[row] = (g - zn) / a;
  s->retrieved[col] = 1;
}

/* postsolve_bound_strengthening.cuh */
static void replay_row_implied(ps_state* s, const int32_t* ix, int64_t i0, const double* v,
                               int64_t v0) {
  const int32_t col = ix[i0], row = ix[i0 + 1];
  const double tol = s->tol;
  if (!s->retrieved[col]) return;
  const double ol = v[v0], ou = v[v0 + 1], nl = v[v0 + 2], nu = v[v0 + 3];
  const int lch = nl > ol + tol, uch = nu < ou - tol;
  if (lch == uch) return;
  const double implied = lch ? nl : nu;
  const double other = lch ? ou : ol;
  const int other_lower = !lch;
  const double xj = s->x[col], zj = s->z[col];
  if (ps_finite(other) && fabs(xj - other) <= tol) {
    const int ok = other_lower ? (zj >= -tol) : (zj <= tol);
    if (ok) return;
  }
  if (fabs(xj - implied) > tol) return;
  const gp_csr* A = &s->model->A;
  double aij = 0.0;
  for (gp_offset p = A->row_ptr[row]; p < A->row_ptr[row + 1]; ++p)
    if (A->col_idx[p] == col) {
      aij = A->vals[p];
      break;
    }
  if (fabs(aij) <= tol) return;
  s->y[row] += zj / aij;
  for (gp_offset p = A->row_ptr[row]; p < A->row_ptr[row + 1]; ++p) {
    const int32_t k = A->col_idx[p];
    if (k == col || !s->retrieved[k]) continue;
    s->z[k] -= (A->vals[p] / aij) * zj;
  }
  s->z[col] = 0.0;
  s->retrieved[col] = 1;
}

/* postsolve_projected_variables.cuh:14-68 (FmeCol): indices K, column, K
 * rows, N, L, (orient, len) x L, N columns; values (constant, coefficients)
 * x L.  Largest lower side, else smallest upper side, else 0. */
static void replay_projected(ps_state* s, const int32_t* ix, int64_t i0, const double* v, int64_t v0,
                             int primal, int dual) {
  const int32_t K = ix[i0], col = ix[i0 + 1];
  const int32_t L = ix[i0 + 3 + K];
  if (primal) {
    int64_t meta = i0 + 4 + K, cp = meta + 2 * (int64_t)L, vp = v0;
    double lo = -INFINITY, hi = INFINITY;
    int has_lo = 0, has_hi = 0;
    for (int32_t side = 0; side < L; ++side) {
      const int32_t orient = ix[meta++];
      const int32_t len = ix[meta++];
      double cand = v[vp];
      for (int32_t t = 0; t < len; ++t) cand += v[vp + 1 + t] * s->x[ix[cp++]];
      if (orient > 0) {
        lo = or_cmax(lo, cand);
        has_lo = 1;
      } else {
        hi = or_cmin(hi, cand);
        has_hi = 1;
      }
      vp += 1 + len;
    }
    s->x[col] = has_lo ? lo : (has_hi ? hi : 0.0);
  }
  if (dual) {
    s->z[col] = 0.0;
    s->retrieved[col] = 1;
    for (int32_t t = 0; t < K; ++t) s->y[ix[i0 + 2 + t]] = 0.0;
  }
}

typedef struct ps_tape {
  const int32_t* types;
  const gp_offset* istarts;
  const gp_offset* vstarts;
  const int32_t* idx;
  const double* vals;
} ps_tape;

/* replay_postsolve_record dispatch (gpu_postsolve.cu:85-147) */
static void replay_record(ps_state* s, const ps_tape* t, int64_t k, int primal, int dual) {
  const int32_t type = t->types[k];
  const int64_t i0 = t->istarts[k], i1 = t->istarts[k + 1] - 1;
  const int64_t v0 = t->vstarts[k], v1 = t->vstarts[k + 1] - 1;
  const int32_t* ix = t->idx;
  const double* v = t->vals;
  const int orig = s->model != NULL;
  switch (type) {
    case GP_RED_PARALLEL_COLUMN: replay_parallel_column(s, ix, i0, v, v0, primal, dual); break;
    case GP_RED_DOUBLETON_EQUATION: replay_doubleton(s, ix, i0, i1, v, v0, v1, primal, dual); break;
    case GP_RED_SINGLETON_COLUMN_SUBSTITUTION: replay_singleton_column(s, ix, i0, v, v0, primal, dual); break;
    case GP_RED_FIXED_VARIABLE: replay_fixed(s, ix, i0, i1, v, v0, primal, dual); break;
    case GP_RED_INFINITE_BOUND_DUAL_FIX: replay_infinite(s, ix, i0, v, v0, primal, dual); break;
    case GP_RED_FME_COL: replay_projected(s, ix, i0, v, v0, primal, dual); break;
    case GP_RED_ROW_DUAL_SHIFT:
      if (dual) s->y[ix[i0]] += v[v0];
      break;
    case GP_RED_PARALLEL_ROW:
      if (dual && i1 >= i0 + 1 && v1 >= v0 + 4) replay_parallel_row(s, ix, i0, v, v0);
      break;
    case GP_RED_DELETED_ROW:
      if (dual && orig && i1 >= i0 + 1 && v1 >= v0 + 2) replay_singleton_row(s, ix, i0, v, v0);
      break;
    case GP_RED_ROW_IMPLIED_BOUND_CHANGE:
      if (dual && orig && i1 >= i0 + 1) replay_row_implied(s, ix, i0, v, v0);
      break;
    default: break;
  }
}

/* _postsolve_project_single_column_dual_from_original (postsolve_common.cuh) */
static void project_column(ps_state* s, int32_t col) {
  const gp_device_lp* M = s->model;
  double aty = 0.0;
  for (gp_offset p = M->AT.row_ptr[col]; p < M->AT.row_ptr[col + 1]; ++p)
    aty += M->AT.vals[p] * s->y[M->AT.col_idx[p]];
  const double r = M->c[col] - aty;
  const double xj = s->x[col], lj = M->l[col], uj = M->u[col];
  const int al = at_lower(xj, lj, s->tol), au = at_upper(xj, uj, s->tol);
  double zj;
  if (al && au) zj = r;
  else if (!al && !au) zj = 0.0;
  else if (al) zj = or_cmax(r, 0.0);
  else zj = or_cmin(r, 0.0);
  const double eps = or_cmin(s->tol, 1.0e-12);
  if (zj > eps && !isfinite(lj)) zj = 0.0;
  else if (zj < -eps && !isfinite(uj)) zj = 0.0;
  s->z[col] = zj;
}

/* Structural primal recoveries: absolute-value steps
 * (postsolve_piecewise_linear_reformulation_absolute_value.cu:9-103,199-205)
 * and variable-pair steps (postsolve_piecewise_linear_reformulation_variable_pair.cu:9-41). */
static void restore_abs_value(double* x, const gp_abs_value_step* st, const double* lower) {
  for (int32_t i = 0; i < st->split_count; ++i) {
    const double t = x[st->split_t_col[i]], e = x[st->split_e_col[i]];
    x[st->split_t_col[i]] = t + e;
    x[st->split_e_col[i]] = (t - e) / st->split_rho[i];
  }
  for (int32_t i = 0; i < st->pair_count; ++i) x[st->pair_free_col[i]] = x[st->pair_bound_col[i]];
  for (int32_t i = 0; i < st->linked_count; ++i)
    x[st->linked_slack_col[i]] = st->linked_factor[i] * x[st->linked_t_col[i]];
  for (int32_t i = 0; i < st->max_count; ++i) {
    if (st->max_starts[i + 1] == st->max_starts[i]) continue;
    const int32_t sc = st->max_slack_col[i];
    double v = lower ? lower[sc] : 0.0;
    if (!isfinite(v)) v = 0.0;
    for (gp_offset q = st->max_starts[i]; q < st->max_starts[i + 1]; ++q)
      v = or_cmax(v, st->max_factors[q] * x[st->max_t_cols[q]]);
    x[sc] = v;
  }
}

static void restore_var_pair(double* x, const gp_var_pair_step* st) {
  double* cp = (double*)or_xcalloc((size_t)(st->count > 0 ? st->count : 1), sizeof(double));
  double* cm = (double*)or_xcalloc((size_t)(st->count > 0 ? st->count : 1), sizeof(double));
  for (int32_t i = 0; i < st->count; ++i) {
    cp[i] = x[st->root_plus_cols[i]];
    cm[i] = x[st->root_minus_cols[i]];
  }
  for (int32_t i = 0; i < st->count; ++i) {
    x[st->plus_cols[i]] = st->plus_shift[i] + cp[i];
    x[st->minus_cols[i]] = st->minus_shift[i] + cm[i];
  }
  free(cp);
  free(cm);
}

static int bad_col(int32_t c, int32_t n0) { return c < 0 || c >= n0; }

/* validate_postsolve_tape (gpu_postsolve.cu:400-437) */
static int validate_tape(const gp_host_record* r) {
  const int64_t count = r->record_count;
  if (count < 0 || r->index_count < 0 || r->value_count < 0) return -1;
  if (count == 0)
    return ((r->index_starts && r->index_starts[0] != 0) || (r->value_starts && r->value_starts[0] != 0) ||
            r->index_count != 0 || r->value_count != 0) ? -1 : 0;
  if (!r->types || !r->index_starts || !r->value_starts || !r->dual_modes) return -1;
  if (r->index_starts[0] != 0 || r->value_starts[0] != 0) return -1;
  for (int64_t k = 0; k < count; ++k) {
    if (r->index_starts[k] < 0 || r->index_starts[k] > r->index_starts[k + 1] || r->value_starts[k] < 0 ||
        r->value_starts[k] > r->value_starts[k + 1])
      return -1;
    if (r->types[k] < GP_RED_FIXED_VARIABLE || r->types[k] > GP_RED_FME_COL) return -1;
    if (r->dual_modes[k] > GP_DUAL_MINIMAL) return -1;
  }
  if (r->index_starts[count] != r->index_count || r->value_starts[count] != r->value_count) return -1;
  return 0;
}

/* validate_primal_recovery_timeline (gpu_postsolve.cu:439-503) */
static int validate_timeline(const gp_host_record* r) {
  const int32_t na = r->abs_value_count > 0 ? r->abs_value_count : 0;
  const int32_t nv = r->var_pair_count > 0 ? r->var_pair_count : 0;
  uint8_t* seen_a = (uint8_t*)or_xcalloc((size_t)(na ? na : 1), 1);
  uint8_t* seen_v = (uint8_t*)or_xcalloc((size_t)(nv ? nv : 1), 1);
  int bad = 0;
  int64_t prev = 0;
  for (int32_t c = 0; c < r->checkpoint_count && !bad; ++c) {
    const gp_checkpoint* cp = &r->checkpoints[c];
    if (cp->tape_position < prev || cp->tape_position > r->record_count) {
      bad = 1;
      break;
    }
    prev = cp->tape_position;
    if (cp->kind == GP_RECOVERY_ABS_VALUE) {
      if (cp->payload_index < 0 || cp->payload_index >= na || seen_a[cp->payload_index]) {
        bad = 1;
        break;
      }
      seen_a[cp->payload_index] = 1;
      const gp_abs_value_step* st = &r->abs_value[cp->payload_index];
      for (int32_t i = 0; i < st->split_count; ++i)
        bad |= bad_col(st->split_t_col[i], r->n0) || bad_col(st->split_e_col[i], r->n0);
      for (int32_t i = 0; i < st->pair_count; ++i)
        bad |= bad_col(st->pair_bound_col[i], r->n0) || bad_col(st->pair_free_col[i], r->n0);
      for (int32_t i = 0; i < st->linked_count; ++i)
        bad |= bad_col(st->linked_slack_col[i], r->n0) || bad_col(st->linked_t_col[i], r->n0);
      for (int32_t i = 0; i < st->max_count && !bad; ++i) {
        bad |= st->max_starts[i + 1] < st->max_starts[i] || bad_col(st->max_slack_col[i], r->n0);
        for (gp_offset q = st->max_starts[i]; q < st->max_starts[i + 1] && !bad; ++q)
          bad |= bad_col(st->max_t_cols[q], r->n0);
      }
    } else if (cp->kind == GP_RECOVERY_VAR_PAIR) {
      if (cp->payload_index < 0 || cp->payload_index >= nv || seen_v[cp->payload_index]) {
        bad = 1;
        break;
      }
      seen_v[cp->payload_index] = 1;
      const gp_var_pair_step* st = &r->var_pair[cp->payload_index];
      for (int32_t i = 0; i < st->count; ++i)
        bad |= bad_col(st->plus_cols[i], r->n0) || bad_col(st->minus_cols[i], r->n0) ||
               bad_col(st->root_plus_cols[i], r->n0) || bad_col(st->root_minus_cols[i], r->n0);
    } else {
      bad = 1;
    }
  }
  for (int32_t i = 0; i < na && !bad; ++i) bad |= !seen_a[i];
  for (int32_t i = 0; i < nv && !bad; ++i) bad |= !seen_v[i];
  free(seen_a);
  free(seen_v);
  return bad ? -1 : 0;
}

static void replay_range(ps_state* s, const ps_tape* t, int64_t begin, int64_t end, int primal) {
  for (int64_t k = end - 1; k >= begin; --k) replay_record(s, t, k, primal, !primal);
}

/* postsolve_gpu (gpu_postsolve.cu:535-670), serial: scatter, fixed restore,
 * primal replay between checkpoints in reverse with their structural
 * payloads, dual replay, projection (needs x, y, z and the model's AT, c, l,
 * u; the reference has no y without rows). */
int or_postsolve_record(const gp_host_record* r, const gp_device_lp* original, const double* original_l,
                        const double* x, const double* y, const double* z, double* x_out, double* y_out,
                        double* z_out) {
  if (validate_tape(r) || validate_timeline(r)) return -1;
  const int32_t m0 = r->m0, n0 = r->n0;
  if (!original_l && original) original_l = original->l;
  ps_state s;
  s.x = x_out;
  s.y = y_out;
  s.z = z_out;
  s.model = original;
  s.tol = 1e-7;
  s.retrieved = (uint8_t*)or_xcalloc((size_t)(n0 ? n0 : 1), 1);
  memset(x_out, 0, sizeof(double) * (size_t)n0);
  memset(z_out, 0, sizeof(double) * (size_t)n0);
  memset(y_out, 0, sizeof(double) * (size_t)m0);
  for (int32_t j = 0; j < r->n1; ++j) {
    if (x) x_out[r->col_red2org[j]] = x[j];
    if (z) z_out[r->col_red2org[j]] = z[j];
    s.retrieved[r->col_red2org[j]] = 1;
  }
  for (int32_t i = 0; i < r->m1; ++i)
    if (y) y_out[r->row_red2org[i]] = y[i];
  for (int32_t k = 0; k < r->fixed_count; ++k) x_out[r->fixed_idx[k]] = r->fixed_val[k];
  const ps_tape t = {r->types, r->index_starts, r->value_starts, r->indices, r->vals};
  int64_t range_end = r->record_count;
  for (int32_t c = r->checkpoint_count - 1; c >= 0; --c) {
    const gp_checkpoint* cp = &r->checkpoints[c];
    replay_range(&s, &t, cp->tape_position, range_end, 1);
    if (cp->kind == GP_RECOVERY_ABS_VALUE) restore_abs_value(x_out, &r->abs_value[cp->payload_index], original_l);
    else restore_var_pair(x_out, &r->var_pair[cp->payload_index]);
    range_end = cp->tape_position;
  }
  replay_range(&s, &t, 0, range_end, 1);
  replay_range(&s, &t, 0, r->record_count, 0);
  if (original && n0 > 0 && m0 > 0 && original->AT.rows > 0 && original->AT.row_ptr && original->c &&
      original->l && original->u) {
    uint8_t* protect = (uint8_t*)or_xcalloc((size_t)n0, 1);
    for (int64_t k = 0; k < r->record_count; ++k) {
      const int32_t type = r->types[k];
      const int64_t i0 = r->index_starts[k];
      int32_t a = -1, b = -1;
      if (type == GP_RED_SINGLETON_COLUMN_SUBSTITUTION || type == GP_RED_PARALLEL_COLUMN) a = r->indices[i0];
      if (type == GP_RED_DOUBLETON_EQUATION) {
        a = r->indices[i0];
        b = r->indices[i0 + 1];
      }
      if (type == GP_RED_INFINITE_BOUND_DUAL_FIX) a = r->indices[i0 + 1];
      if (a >= 0 && a < n0) protect[a] = 1;
      if (b >= 0 && b < n0) protect[b] = 1;
    }
    for (int32_t j = 0; j < n0; ++j)
      if (!protect[j]) project_column(&s, j);
    free(protect);
  }
  free(s.retrieved);
  return 0;
}

int or_postsolve(const or_result* r, const gp_device_lp* original, const double* x, const double* y,
                 const double* z, double* x_out, double* y_out, double* z_out) {
  gp_host_record h;
  memset(&h, 0, sizeof h);
  h.m0 = r->m0;
  h.n0 = r->n0;
  h.m1 = r->m1;
  h.n1 = r->n1;
  h.row_red2org = r->row_red2org;
  h.col_red2org = r->col_red2org;
  h.fixed_count = r->fixed_count;
  h.fixed_idx = r->fixed_idx;
  h.fixed_val = r->fixed_val;
  h.record_count = (gp_offset)r->tape.count;
  h.types = r->tape.types;
  h.index_starts = r->tape.istarts;
  h.value_starts = r->tape.vstarts;
  h.dual_modes = r->tape.modes;
  h.index_count = (gp_offset)r->tape.icount;
  h.indices = r->tape.idx;
  h.value_count = (gp_offset)r->tape.vcount;
  h.vals = r->tape.vals;
  return or_postsolve_record(&h, original, NULL, x, y, z, x_out, y_out, z_out);
}
