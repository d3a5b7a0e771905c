/*
 * oracle_core.c -- tape, numerics, live statistics, activities, plans and the
 * matrix primitives of the CPU restatement.  TEST INFRASTRUCTURE ONLY: the
 * parity checker for the HIP engine, never linked into it.
 */
#include "oracle_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void* or_xmalloc(size_t bytes) {
  void* p = malloc(bytes ? bytes : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu bytes)\n", bytes);
    abort();
  }
  return p;
}

void* or_xcalloc(size_t count, size_t size) {
  void* p = calloc(count ? count : 1, size ? size : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

void* or_xrealloc(void* p, size_t bytes) {
  void* q = realloc(p, bytes ? bytes : 1);
  if (!q) {
    fprintf(stderr, "oracle: out of memory (%zu bytes)\n", bytes);
    abort();
  }
  return q;
}

/* ---- tape ---------------------------------------------------------------- */

void or_tape_init(or_tape* t) {
  memset(t, 0, sizeof(*t));
  t->cap = 16;
  t->icap = 64;
  t->vcap = 64;
  t->types = (int32_t*)or_xmalloc(sizeof(int32_t) * t->cap);
  t->modes = (uint8_t*)or_xmalloc(t->cap);
  t->istarts = (gp_offset*)or_xmalloc(sizeof(gp_offset) * (t->cap + 1));
  t->vstarts = (gp_offset*)or_xmalloc(sizeof(gp_offset) * (t->cap + 1));
  t->idx = (int32_t*)or_xmalloc(sizeof(int32_t) * t->icap);
  t->vals = (double*)or_xmalloc(sizeof(double) * t->vcap);
  t->istarts[0] = 0;
  t->vstarts[0] = 0;
}

void or_tape_free(or_tape* t) {
  free(t->types);
  free(t->modes);
  free(t->istarts);
  free(t->vstarts);
  free(t->idx);
  free(t->vals);
  memset(t, 0, sizeof(*t));
}

void or_tape_push(or_tape* t, int32_t type, uint8_t mode, const int32_t* idx, int64_t ni,
                  const double* vals, int64_t nv) {
  if (t->count + 1 > t->cap) {
    t->cap = 2 * t->cap + 1;
    t->types = (int32_t*)or_xrealloc(t->types, sizeof(int32_t) * t->cap);
    t->modes = (uint8_t*)or_xrealloc(t->modes, t->cap);
    t->istarts = (gp_offset*)or_xrealloc(t->istarts, sizeof(gp_offset) * (t->cap + 1));
    t->vstarts = (gp_offset*)or_xrealloc(t->vstarts, sizeof(gp_offset) * (t->cap + 1));
  }
  while (t->icount + ni > t->icap) {
    t->icap = 2 * t->icap + 1;
    t->idx = (int32_t*)or_xrealloc(t->idx, sizeof(int32_t) * t->icap);
  }
  while (t->vcount + nv > t->vcap) {
    t->vcap = 2 * t->vcap + 1;
    t->vals = (double*)or_xrealloc(t->vals, sizeof(double) * t->vcap);
  }
  t->types[t->count] = type;
  t->modes[t->count] = mode;
  if (ni) memcpy(t->idx + t->icount, idx, sizeof(int32_t) * ni);
  if (nv) memcpy(t->vals + t->vcount, vals, sizeof(double) * nv);
  t->icount += ni;
  t->vcount += nv;
  t->count += 1;
  t->istarts[t->count] = (gp_offset)t->icount;
  t->vstarts[t->count] = (gp_offset)t->vcount;
}

/* Append tape `src` to `dst`, mapping local indices through the current
 * physical->original maps (globalize, device_recovery.cu:62-185). */
void or_tape_append_global(or_tape* dst, const or_tape* src, const int32_t* rows, int32_t m,
                           const int32_t* cols, int32_t n) {
  for (int64_t k = 0; k < src->count; ++k) {
    const int64_t i0 = src->istarts[k], i1 = src->istarts[k + 1];
    const int64_t v0 = src->vstarts[k], v1 = src->vstarts[k + 1];
    const int64_t ni = i1 - i0;
    int32_t* ix = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(ni ? ni : 1));
    memcpy(ix, src->idx + i0, sizeof(int32_t) * (size_t)ni);
    or_globalize_record(src->types[k], ix, ni, rows, m, cols, n);
    or_tape_push(dst, src->types[k], src->modes[k], ix, ni, src->vals + v0, v1 - v0);
    free(ix);
  }
}

static int32_t map_col(int32_t x, const int32_t* cols, int32_t n) {
  return x >= 0 && x < n ? cols[x] : x;
}
static int32_t map_row(int32_t x, const int32_t* rows, int32_t m) {
  return x >= 0 && x < m ? rows[x] : x;
}

/* Record-type specific index globalisation (device_recovery.cu:62-185). */
void or_globalize_record(int32_t type, int32_t* ix, int64_t size, const int32_t* rows, int32_t m,
                         const int32_t* cols, int32_t n) {
  switch (type) {
    case GP_RED_FIXED_VARIABLE:
      if (size > 0) {
        ix[0] = map_col(ix[0], cols, n);
        for (int64_t p = 1; p < size; ++p) ix[p] = map_row(ix[p], rows, m);
      }
      break;
    case GP_RED_INFINITE_BOUND_DUAL_FIX:
      if (size >= 3) {
        ix[1] = map_col(ix[1], cols, n);
        int64_t p = 3;
        while (p < size) {
          const int64_t len = ix[p++];
          for (int64_t t = 0; t < len && p < size; ++t, ++p) ix[p] = map_col(ix[p], cols, n);
        }
      }
      break;
    case GP_RED_SINGLETON_COLUMN_SUBSTITUTION:
      if (size >= 3) {
        ix[0] = map_col(ix[0], cols, n);
        ix[1] = map_row(ix[1], rows, m);
        const int32_t support = ix[2];
        for (int32_t t = 0; t < support && 3 + t < size; ++t) ix[3 + t] = map_col(ix[3 + t], cols, n);
      }
      break;
    case GP_RED_PARALLEL_COLUMN:
      if (size >= 2) {
        ix[0] = map_col(ix[0], cols, n);
        ix[1] = map_col(ix[1], cols, n);
      }
      break;
    case GP_RED_PARALLEL_ROW:
      if (size >= 2) {
        ix[0] = map_row(ix[0], rows, m);
        ix[1] = map_row(ix[1], rows, m);
      }
      break;
    case GP_RED_DELETED_ROW:
      if (size >= 1) {
        ix[0] = map_row(ix[0], rows, m);
        if (size >= 2) ix[1] = map_col(ix[1], cols, n);
      }
      break;
    case GP_RED_ROW_LOWER_BOUND_CHANGE:
    case GP_RED_ROW_UPPER_BOUND_CHANGE:
    case GP_RED_ROW_DUAL_SHIFT:
      if (size >= 1) ix[0] = map_row(ix[0], rows, m);
      break;
    case GP_RED_VARIABLE_BOUND_CHANGE:
      if (size >= 1) ix[0] = map_col(ix[0], cols, n);
      break;
    case GP_RED_ROW_IMPLIED_BOUND_CHANGE:
      if (size >= 2) {
        ix[0] = map_col(ix[0], cols, n);
        ix[1] = map_row(ix[1], rows, m);
      }
      break;
    case GP_RED_DOUBLETON_EQUATION:
      if (size >= 4) {
        ix[0] = map_col(ix[0], cols, n);
        ix[1] = map_col(ix[1], cols, n);
        ix[2] = map_row(ix[2], rows, m);
        const int32_t support = ix[3];
        for (int32_t t = 0; t < support && 4 + t < size; ++t) ix[4 + t] = map_row(ix[4 + t], rows, m);
      }
      break;
    case GP_RED_FME_COL:
      /* [K, column, K rows, N, L, (orient, len) x L, N columns] */
      if (size >= 4) {
        const int32_t K = ix[0];
        ix[1] = map_col(ix[1], cols, n);
        for (int32_t t = 0; t < K && 2 + t < size; ++t) ix[2 + t] = map_row(ix[2 + t], rows, m);
        const int64_t sp = 2 + (int64_t)K;
        if (sp + 1 < size) {
          const int32_t N = ix[sp], L = ix[sp + 1];
          const int64_t first = sp + 2 + 2 * (int64_t)L;
          for (int32_t t = 0; t < N && first + t < size; ++t) ix[first + t] = map_col(ix[first + t], cols, n);
        }
      }
      break;
    default:
      break;
  }
}

/* ---- numerics -------------------------------------------------------------
 * Row sums: Neumaier-compensated, sequential in storage order inside blocks of
 * OR_SUM_BLOCK consecutive stored entries; block states are merged in block
 * order.  Global sums: fixed pairwise tree over chunks of 1024 (stride-halving
 * order), applied recursively to the chunk results. */

double or_row_sum_finite(const double* terms, int64_t n) {
  or_neu total = {0.0, 0.0}, blk = {0.0, 0.0};
  for (int64_t k = 0; k < n; ++k) {
    if (k > 0 && k % OR_SUM_BLOCK == 0) {
      or_neu_merge(&total, blk);
      blk.s = blk.c = 0.0;
    }
    or_neu_add(&blk, terms[k]);
  }
  or_neu_merge(&total, blk);
  return or_neu_value(total);
}

double or_det_sum(const double* x, int64_t n) {
  if (n <= 0) return 0.0;
  const int64_t chunks = (n + 1023) / 1024;
  double* partial = (double*)or_xmalloc(sizeof(double) * (size_t)chunks);
  double y[1024];
  for (int64_t c = 0; c < chunks; ++c) {
    for (int t = 0; t < 1024; ++t) {
      const int64_t i = c * 1024 + t;
      y[t] = i < n ? x[i] : 0.0;
    }
    for (int s = 512; s > 0; s >>= 1)
      for (int t = 0; t < s; ++t) y[t] = y[t] + y[t + s];
    partial[c] = y[0];
  }
  double result = chunks == 1 ? partial[0] : or_det_sum(partial, chunks);
  free(partial);
  return result;
}

/* ---- activities (gpu_presolve_kernels.cu:96-148) ------------------------- */

void or_row_activity_summary(const gp_csr* A, const double* l, const double* u, double zero_tol,
                             const uint8_t* keep_row, const uint8_t* keep_col,
                             double* min_fin, double* max_fin, int32_t* min_ninf,
                             int32_t* max_pinf) {
  for (int32_t i = 0; i < A->rows; ++i) {
    or_neu tmin = {0, 0}, tmax = {0, 0}, bmin = {0, 0}, bmax = {0, 0};
    int32_t ni = 0, pi = 0;
    const gp_offset p0 = A->row_ptr[i], p1 = A->row_ptr[i + 1];
    const int live_row = !keep_row || keep_row[i];
    for (gp_offset p = p0; p < p1; ++p) {
      const int64_t k = p - p0;
      if (k > 0 && k % OR_SUM_BLOCK == 0) {
        or_neu_merge(&tmin, bmin);
        or_neu_merge(&tmax, bmax);
        bmin.s = bmin.c = bmax.s = bmax.c = 0.0;
      }
      if (!live_row) continue;
      const int32_t j = A->col_idx[p];
      if (keep_col && !keep_col[j]) continue;
      const double a = A->vals[p];
      if (fabs(a) <= zero_tol) continue;
      const double lo = a >= 0.0 ? a * l[j] : a * u[j];
      const double hi = a >= 0.0 ? a * u[j] : a * l[j];
      if (isfinite(lo)) or_neu_add(&bmin, lo);
      else if (lo < 0.0) ++ni;
      if (isfinite(hi)) or_neu_add(&bmax, hi);
      else if (hi > 0.0) ++pi;
    }
    or_neu_merge(&tmin, bmin);
    or_neu_merge(&tmax, bmax);
    min_fin[i] = or_neu_value(tmin);
    max_fin[i] = or_neu_value(tmax);
    min_ninf[i] = ni;
    max_pinf[i] = pi;
  }
}

/* ---- live statistics (gpu_presolve_kernels.cu:422-517) ------------------- */

int or_stats_create(int32_t m, int32_t n, gp_stats* s) {
  memset(s, 0, sizeof(*s));
  s->row_nnz = (int32_t*)or_xcalloc((size_t)m, sizeof(int32_t));
  s->empty_row_mask = (uint8_t*)or_xcalloc((size_t)m, 1);
  s->singleton_row_mask = (uint8_t*)or_xcalloc((size_t)m, 1);
  s->singleton_row_col = (int32_t*)or_xcalloc((size_t)m, sizeof(int32_t));
  s->singleton_row_val = (double*)or_xcalloc((size_t)m, sizeof(double));
  s->col_nnz = (int32_t*)or_xcalloc((size_t)n, sizeof(int32_t));
  s->empty_col_mask = (uint8_t*)or_xcalloc((size_t)n, 1);
  s->singleton_column_mask = (uint8_t*)or_xcalloc((size_t)n, 1);
  s->singleton_column_row = (int32_t*)or_xcalloc((size_t)n, sizeof(int32_t));
  s->singleton_column_val = (double*)or_xcalloc((size_t)n, sizeof(double));
  return 0;
}

void or_stats_destroy(gp_stats* s) {
  free(s->row_nnz);
  free(s->empty_row_mask);
  free(s->singleton_row_mask);
  free(s->singleton_row_col);
  free(s->singleton_row_val);
  free(s->col_nnz);
  free(s->empty_col_mask);
  free(s->singleton_column_mask);
  free(s->singleton_column_row);
  free(s->singleton_column_val);
  memset(s, 0, sizeof(*s));
}

static int64_t live_stats_side(const gp_csr* M, const uint8_t* keep_outer, const uint8_t* keep_inner,
                               int32_t* degree, uint8_t* empty, uint8_t* singleton,
                               int32_t* support, double* coefficient) {
  int64_t total = 0;
  for (int32_t r = 0; r < M->rows; ++r) {
    const int active = !keep_outer || keep_outer[r];
    int32_t count = 0;
    gp_offset position = -1;
    if (active) {
      for (gp_offset p = M->row_ptr[r]; p < M->row_ptr[r + 1]; ++p) {
        if (keep_inner && !keep_inner[M->col_idx[p]]) continue;
        ++count;
        position = p;
      }
    }
    degree[r] = count;
    empty[r] = (uint8_t)(active && count == 0);
    singleton[r] = (uint8_t)(active && count == 1);
    support[r] = count == 1 ? M->col_idx[position] : -1;
    coefficient[r] = count == 1 ? M->vals[position] : 0.0;
    total += count;
  }
  return total;
}

void or_refresh_live_stats(gp_stats* s, const gp_device_lp* lp, const gp_plan* plan) {
  s->live_nnz = live_stats_side(&lp->A, plan->keep_row_mask, plan->keep_col_mask, s->row_nnz,
                                s->empty_row_mask, s->singleton_row_mask, s->singleton_row_col,
                                s->singleton_row_val);
  live_stats_side(&lp->AT, plan->keep_col_mask, plan->keep_row_mask, s->col_nnz, s->empty_col_mask,
                  s->singleton_column_mask, s->singleton_column_row, s->singleton_column_val);
  s->live_masks = 1;
}

/* ---- plan (gpu_presolve.cu:567-706) -------------------------------------- */

int or_plan_create(const gp_device_lp* lp, gp_plan* plan) {
  const int32_t m = lp->A.rows, n = lp->A.cols;
  memset(plan, 0, sizeof(*plan));
  plan->keep_row_mask = (uint8_t*)or_xmalloc((size_t)m);
  plan->keep_col_mask = (uint8_t*)or_xmalloc((size_t)n);
  memset(plan->keep_row_mask, 1, (size_t)m);
  memset(plan->keep_col_mask, 1, (size_t)n);
  plan->new_c = (double*)or_xmalloc(sizeof(double) * (size_t)n);
  plan->new_l = (double*)or_xmalloc(sizeof(double) * (size_t)n);
  plan->new_u = (double*)or_xmalloc(sizeof(double) * (size_t)n);
  plan->new_AL = (double*)or_xmalloc(sizeof(double) * (size_t)m);
  plan->new_AU = (double*)or_xmalloc(sizeof(double) * (size_t)m);
  if (n) {
    memcpy(plan->new_c, lp->c, sizeof(double) * (size_t)n);
    memcpy(plan->new_l, lp->l, sizeof(double) * (size_t)n);
    memcpy(plan->new_u, lp->u, sizeof(double) * (size_t)n);
  }
  if (m) {
    memcpy(plan->new_AL, lp->AL, sizeof(double) * (size_t)m);
    memcpy(plan->new_AU, lp->AU, sizeof(double) * (size_t)m);
  }
  return 0;
}

void or_plan_enable_tape(gp_plan* plan) {
  if (plan->tape) return;
  or_tape* t = (or_tape*)or_xmalloc(sizeof(or_tape));
  or_tape_init(t);
  plan->tape = t;
}

void or_plan_destroy(gp_plan* plan) {
  free(plan->keep_row_mask);
  free(plan->keep_col_mask);
  free(plan->new_c);
  free(plan->new_l);
  free(plan->new_u);
  free(plan->new_AL);
  free(plan->new_AU);
  if (plan->has_new_A) or_csr_free(&plan->new_A);
  if (plan->tape) {
    or_tape_free((or_tape*)plan->tape);
    free(plan->tape);
  }
  memset(plan, 0, sizeof(*plan));
}

/* ---- matrix primitives ----------------------------------------------------- */

/* CSR of A^T; each output row lists its entries in increasing input-row order
 * (the unique csr2csc result for a sorted input, transpose64.cu:54-67). */
void or_transpose(const gp_csr* A, gp_csr* AT) {
  const int32_t m = A->rows, n = A->cols;
  const gp_offset nnz = A->nnz;
  AT->rows = n;
  AT->cols = m;
  AT->nnz = nnz;
  AT->row_ptr = (gp_offset*)or_xcalloc((size_t)n + 1, sizeof(gp_offset));
  AT->col_idx = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)nnz);
  AT->vals = (double*)or_xmalloc(sizeof(double) * (size_t)nnz);
  for (gp_offset p = 0; p < nnz; ++p) AT->row_ptr[A->col_idx[p] + 1]++;
  for (int32_t j = 0; j < n; ++j) AT->row_ptr[j + 1] += AT->row_ptr[j];
  gp_offset* next = (gp_offset*)or_xmalloc(sizeof(gp_offset) * (size_t)(n ? n : 1));
  if (n) memcpy(next, AT->row_ptr, sizeof(gp_offset) * (size_t)n);
  for (int32_t i = 0; i < m; ++i)
    for (gp_offset p = A->row_ptr[i]; p < A->row_ptr[i + 1]; ++p) {
      const int32_t j = A->col_idx[p];
      const gp_offset q = next[j]++;
      AT->col_idx[q] = i;
      AT->vals[q] = A->vals[p];
    }
  free(next);
}

/* Stable compaction by keep masks (gpu_presolve.cu:1793-1930). */
void or_compact(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, gp_csr* out) {
  int32_t m_new = 0, n_new = 0;
  int32_t* col_map = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(A->cols ? A->cols : 1));
  for (int32_t j = 0; j < A->cols; ++j) col_map[j] = (!keep_col || keep_col[j]) ? n_new++ : -1;
  for (int32_t i = 0; i < A->rows; ++i) m_new += (!keep_row || keep_row[i]) ? 1 : 0;
  out->rows = m_new;
  out->cols = n_new;
  out->row_ptr = (gp_offset*)or_xcalloc((size_t)m_new + 1, sizeof(gp_offset));
  gp_offset nnz = 0;
  for (int32_t i = 0; i < A->rows; ++i) {
    if (keep_row && !keep_row[i]) continue;
    for (gp_offset p = A->row_ptr[i]; p < A->row_ptr[i + 1]; ++p) nnz += col_map[A->col_idx[p]] >= 0;
  }
  out->nnz = nnz;
  out->col_idx = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)nnz);
  out->vals = (double*)or_xmalloc(sizeof(double) * (size_t)nnz);
  gp_offset q = 0;
  int32_t r = 0;
  for (int32_t i = 0; i < A->rows; ++i) {
    if (keep_row && !keep_row[i]) continue;
    for (gp_offset p = A->row_ptr[i]; p < A->row_ptr[i + 1]; ++p) {
      const int32_t j = col_map[A->col_idx[p]];
      if (j < 0) continue;
      out->col_idx[q] = j;
      out->vals[q] = A->vals[p];
      ++q;
    }
    out->row_ptr[++r] = q;
  }
  free(col_map);
}

void or_csr_free(gp_csr* A) {
  free(A->row_ptr);
  free(A->col_idx);
  free(A->vals);
  memset(A, 0, sizeof(*A));
}

void or_lp_free(gp_device_lp* lp) {
  or_csr_free(&lp->A);
  or_csr_free(&lp->AT);
  free(lp->c);
  free(lp->AL);
  free(lp->AU);
  free(lp->l);
  free(lp->u);
  memset(lp, 0, sizeof(*lp));
}

static void* dup_bytes(const void* src, size_t bytes) {
  void* p = or_xmalloc(bytes);
  if (bytes) memcpy(p, src, bytes);
  return p;
}

static void csr_copy(const gp_csr* s, gp_csr* d) {
  *d = *s;
  d->row_ptr = (gp_offset*)dup_bytes(s->row_ptr, sizeof(gp_offset) * ((size_t)s->rows + 1));
  d->col_idx = (int32_t*)dup_bytes(s->col_idx, sizeof(int32_t) * (size_t)s->nnz);
  d->vals = (double*)dup_bytes(s->vals, sizeof(double) * (size_t)s->nnz);
}

int or_lp_copy(const gp_device_lp* s, gp_device_lp* d) {
  const size_t m = (size_t)s->A.rows, n = (size_t)s->A.cols;
  csr_copy(&s->A, &d->A);
  csr_copy(&s->AT, &d->AT);
  d->c = (double*)dup_bytes(s->c, sizeof(double) * n);
  d->l = (double*)dup_bytes(s->l, sizeof(double) * n);
  d->u = (double*)dup_bytes(s->u, sizeof(double) * n);
  d->AL = (double*)dup_bytes(s->AL, sizeof(double) * m);
  d->AU = (double*)dup_bytes(s->AU, sizeof(double) * m);
  d->obj_constant = s->obj_constant;
  return 0;
}
