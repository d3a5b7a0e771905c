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

/* replay_postsolve_record dispatch (gpu_postsolve.cu:85-147) */
static void replay_record(ps_state* s, const or_tape* t, int64_t k, int primal, int dual) {
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

int or_postsolve(const or_result* r, const gp_device_lp* original, const double* x, const double* y,
                 const double* z, double* x_out, double* y_out, double* z_out) {
  const int32_t m0 = r->m0, n0 = r->n0;
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
  for (int64_t k = r->tape.count - 1; k >= 0; --k) replay_record(&s, &r->tape, k, 1, 0);
  for (int64_t k = r->tape.count - 1; k >= 0; --k) replay_record(&s, &r->tape, k, 0, 1);
  if (original && n0 > 0) {
    uint8_t* protect = (uint8_t*)or_xcalloc((size_t)n0, 1);
    for (int64_t k = 0; k < r->tape.count; ++k) {
      const int32_t type = r->tape.types[k];
      const int64_t i0 = r->tape.istarts[k];
      int32_t a = -1, b = -1;
      if (type == GP_RED_SINGLETON_COLUMN_SUBSTITUTION || type == GP_RED_PARALLEL_COLUMN) a = r->tape.idx[i0];
      if (type == GP_RED_DOUBLETON_EQUATION) {
        a = r->tape.idx[i0];
        b = r->tape.idx[i0 + 1];
      }
      if (type == GP_RED_INFINITE_BOUND_DUAL_FIX) a = r->tape.idx[i0 + 1];
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
