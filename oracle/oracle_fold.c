/*
 * oracle_fold.c -- CPU restatement of the reference's LP folding
 * (cpp/src/folding/: folding.cu:69-136, folding_refinement.cu:227-361,
 * folding_kernels.cu:46-186, folding_reduce.cu:239-565, folding_map.cu).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker of the engine's folding
 * stage (gp_fold, and gp_preprocess with enable_folding).
 *
 * Host pointers in the gpupre.h structs.  Sums run left to right in the
 * orders stated below, which the HIP kernels reproduce, so classes, scales
 * and the folded LP compare bit for bit.  The class signatures use the C
 * library's log / cos / sqrt (the device uses its own): the value that
 * matters, the signature rounded to float, agrees unless a signature lies
 * within an ulp of a float rounding boundary.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint64_t splitmix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

/* color_seed (folding_kernels.cu:52-54) */
static uint64_t round_seed(int32_t round, uint64_t side) {
  return splitmix((uint64_t)round ^ (side * 0x9e3779b97f4a7c15ULL));
}

/* unit_from_mix (folding_kernels.cu:40-44) */
static double unit53(uint64_t x) {
  const double u = (double)(splitmix(x) >> 11) * 0x1.0p-53;
  return u > 0.0 ? u : 0x1.0p-53;
}

/* kernel_write_color_sig (folding_kernels.cu:103-114): Box-Muller normal */
static double signature(uint64_t seed, int32_t cls) {
  const uint64_t c = (uint64_t)cls;
  const double u1 = unit53(seed ^ (c * 0xbf58476d1ce4e5b9ULL));
  const double u2 = unit53((seed + 0x9e3779b97f4a7c15ULL) ^ (c * 0x94d049bb133111ebULL));
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

typedef struct {
  int32_t cls;
  float value;
  int32_t idx;
} class_item;

static int class_item_less(const void* a, const void* b) {
  const class_item* x = (const class_item*)a;
  const class_item* y = (const class_item*)b;
  if (x->cls != y->cls) return x->cls < y->cls ? -1 : 1;
  if (x->value != y->value) return x->value < y->value ? -1 : 1;
  return x->idx < y->idx ? -1 : x->idx > y->idx; /* ties: any order gives the same classes */
}

/* update_color_id (folding_kernels.cu:162-204): sort (class, float value),
 * a new class where the class changes or the value exceeds its predecessor
 * by more than tol, ids = ranks; optionally the new signatures. */
static int32_t refine(int32_t n, int32_t* cls, const double* value, double tol, double* sig, uint64_t seed) {
  if (n <= 0) return 0;
  class_item* it = (class_item*)malloc(sizeof(class_item) * (size_t)n);
  for (int32_t i = 0; i < n; ++i) {
    it[i].cls = cls[i];
    it[i].value = (float)value[i];
    it[i].idx = i;
  }
  qsort(it, (size_t)n, sizeof(class_item), class_item_less);
  int32_t id = 0;
  for (int32_t k = 0; k < n; ++k) {
    const int start = k == 0 || it[k].cls != it[k - 1].cls || (double)it[k].value > (double)it[k - 1].value + tol;
    id += start;
    cls[it[k].idx] = id;
    if (sig) sig[it[k].idx] = signature(seed, id);
  }
  free(it);
  return id;
}

/* one line of M times a vector, left to right from 0 */
static void line_products(const gp_csr* M, const double* v, double* out) {
  for (int32_t i = 0; i < M->rows; ++i) {
    double s = 0.0;
    for (gp_offset p = M->row_ptr[i]; p < M->row_ptr[i + 1]; ++p) s += M->vals[p] * v[M->col_idx[p]];
    out[i] = s;
  }
}

static int too_little(int32_t mc, int32_t nc, int32_t m, int32_t n) {
  return (double)mc > 0.8 * m && (double)nc > 0.8 * n;
}

typedef struct {
  uint64_t key;
  int64_t pos;
  double val;
} fold_entry;

static int fold_entry_less(const void* a, const void* b) {
  const fold_entry* x = (const fold_entry*)a;
  const fold_entry* y = (const fold_entry*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->pos < y->pos ? -1 : x->pos > y->pos;
}

/* kernel_reduce_color_values_constant / _bounds_constant
 * (folding_reduce.cu:154-226): the sum over the members in increasing index
 * from 0 (the exact-constant shortcut gives the same bits), divided by the
 * class size for averages */
static void class_sums(int32_t n, const int32_t* cls, int32_t classes, const int32_t* count, const double* v,
                       int average, double* out) {
  for (int32_t k = 0; k < classes; ++k) out[k] = 0.0;
  for (int32_t i = 0; i < n; ++i) out[cls[i] - 1] += v[i];
  if (average)
    for (int32_t k = 0; k < classes; ++k) out[k] /= (double)count[k];
}

static void build_folded(const gp_device_lp* lp, double tol, or_fold* f) {
  const int32_t m = lp->A.rows, n = lp->A.cols, mf = f->mf, nf = f->nf;
  int32_t* rcount = (int32_t*)calloc((size_t)mf, sizeof(int32_t));
  int32_t* ccount = (int32_t*)calloc((size_t)nf, sizeof(int32_t));
  for (int32_t i = 0; i < m; ++i) rcount[f->row_class[i] - 1]++;
  for (int32_t j = 0; j < n; ++j) ccount[f->col_class[j] - 1]++;
  gp_device_lp* out = &f->lp;
  memset(out, 0, sizeof(*out));
  out->c = (double*)malloc(sizeof(double) * (size_t)(nf ? nf : 1));
  out->l = (double*)malloc(sizeof(double) * (size_t)(nf ? nf : 1));
  out->u = (double*)malloc(sizeof(double) * (size_t)(nf ? nf : 1));
  out->AL = (double*)malloc(sizeof(double) * (size_t)(mf ? mf : 1));
  out->AU = (double*)malloc(sizeof(double) * (size_t)(mf ? mf : 1));
  out->obj_constant = lp->obj_constant;
  class_sums(n, f->col_class, nf, ccount, lp->c, 0, out->c);
  class_sums(n, f->col_class, nf, ccount, lp->l, 1, out->l);
  class_sums(n, f->col_class, nf, ccount, lp->u, 1, out->u);
  class_sums(m, f->row_class, mf, rcount, lp->AL, 1, out->AL);
  class_sums(m, f->row_class, mf, rcount, lp->AU, 1, out->AU);
  f->row_scale = (double*)malloc(sizeof(double) * (size_t)(m ? m : 1));
  f->col_scale = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
  for (int32_t i = 0; i < m; ++i) f->row_scale[i] = 1.0 / (double)rcount[f->row_class[i] - 1];
  for (int32_t j = 0; j < n; ++j) f->col_scale[j] = 1.0 / (double)ccount[f->col_class[j] - 1];

  /* kernel_emit_triplet_warp (folding_reduce.cu:54-77): key (R, C), value
   * a * (1 / |R|); a stable sort by key and a sum per key in sorted order
   * (folding_reduce.cu:343-360); entries with |sum| <= tol dropped
   * (kernel_mark_nonzero, :79-86) */
  const int64_t nnz = lp->A.nnz;
  fold_entry* e = (fold_entry*)malloc(sizeof(fold_entry) * (size_t)(nnz ? nnz : 1));
  for (int32_t i = 0; i < m; ++i) {
    const int32_t r = f->row_class[i] - 1;
    const double w = 1.0 / (double)rcount[r];
    for (gp_offset p = lp->A.row_ptr[i]; p < lp->A.row_ptr[i + 1]; ++p) {
      e[p].key = (uint64_t)r * (uint64_t)nf + (uint64_t)(f->col_class[lp->A.col_idx[p]] - 1);
      e[p].pos = p;
      e[p].val = lp->A.vals[p] * w;
    }
  }
  qsort(e, (size_t)nnz, sizeof(fold_entry), fold_entry_less);
  gp_offset* rp = (gp_offset*)calloc((size_t)mf + 1, sizeof(gp_offset));
  int32_t* ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
  double* cv = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  int64_t kept = 0;
  for (int64_t k = 0; k < nnz;) {
    double s = 0.0;
    int64_t t = k;
    for (; t < nnz && e[t].key == e[k].key; ++t) s += e[t].val;
    if (fabs(s) > tol) {
      const int32_t r = (int32_t)(e[k].key / (uint64_t)nf);
      ci[kept] = (int32_t)(e[k].key % (uint64_t)nf);
      cv[kept] = s;
      rp[r + 1]++;
      ++kept;
    }
    k = t;
  }
  for (int32_t r = 0; r < mf; ++r) rp[r + 1] += rp[r];
  out->A.rows = mf;
  out->A.cols = nf;
  out->A.nnz = (gp_offset)kept;
  out->A.row_ptr = rp;
  out->A.col_idx = ci;
  out->A.vals = cv;
  /* A^T by a counting pass (rows of each column in increasing order) */
  gp_offset* tp = (gp_offset*)calloc((size_t)nf + 1, sizeof(gp_offset));
  int32_t* ti = (int32_t*)malloc(sizeof(int32_t) * (size_t)(kept ? kept : 1));
  double* tv = (double*)malloc(sizeof(double) * (size_t)(kept ? kept : 1));
  for (int64_t q = 0; q < kept; ++q) tp[ci[q] + 1]++;
  for (int32_t c = 0; c < nf; ++c) tp[c + 1] += tp[c];
  gp_offset* next = (gp_offset*)malloc(sizeof(gp_offset) * (size_t)(nf ? nf : 1));
  for (int32_t c = 0; c < nf; ++c) next[c] = tp[c];
  for (int32_t r = 0; r < mf; ++r)
    for (gp_offset q = rp[r]; q < rp[r + 1]; ++q) {
      const gp_offset d = next[ci[q]]++;
      ti[d] = r;
      tv[d] = cv[q];
    }
  out->AT.rows = nf;
  out->AT.cols = mf;
  out->AT.nnz = (gp_offset)kept;
  out->AT.row_ptr = tp;
  out->AT.col_idx = ti;
  out->AT.vals = tv;
  free(next);
  free(e);
  free(rcount);
  free(ccount);
}

void or_fold_lp(const gp_device_lp* lp, double tol, or_fold* f) {
  memset(f, 0, sizeof(*f));
  const int32_t m = lp->A.rows, n = lp->A.cols;
  f->m = f->mf = m;
  f->n = f->nf = n;
  f->refinement_ok = 1;
  if (m == 0 || n == 0) return;
  f->row_class = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
  f->col_class = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  for (int32_t i = 0; i < m; ++i) f->row_class[i] = 1;
  for (int32_t j = 0; j < n; ++j) f->col_class[j] = 1;
  double* rsig = (double*)malloc(sizeof(double) * (size_t)m);
  double* csig = (double*)malloc(sizeof(double) * (size_t)n);
  double* prod_c = (double*)malloc(sizeof(double) * (size_t)n);
  double* prod_r = (double*)malloc(sizeof(double) * (size_t)m);
  /* init_color (folding_refinement.cu:235-262) */
  int32_t mc = refine(m, f->row_class, lp->AL, tol, NULL, 0);
  mc = refine(m, f->row_class, lp->AU, tol, NULL, 0);
  int32_t nc = refine(n, f->col_class, lp->c, tol, NULL, 0);
  nc = refine(n, f->col_class, lp->l, tol, NULL, 0);
  nc = refine(n, f->col_class, lp->u, tol, NULL, 0);
  for (int32_t i = 0; i < m; ++i) rsig[i] = signature(round_seed(0, 2), f->row_class[i]);
  for (int32_t j = 0; j < n; ++j) csig[j] = signature(round_seed(0, 5), f->col_class[j]);
  f->mf = mc;
  f->nf = nc;
  int ok = !too_little(mc, nc, m, n);
  /* refine_color rounds (folding_refinement.cu:288-359) */
  const int32_t max_round = m + n > 1 ? m + n : 1;
  for (int32_t round = 1; ok && round <= max_round; ++round) {
    f->rounds = round;
    line_products(&lp->AT, rsig, prod_c);
    const int32_t nc2 = refine(n, f->col_class, prod_c, tol, csig, round_seed(round, 1));
    line_products(&lp->A, csig, prod_r);
    const int32_t mc2 = refine(m, f->row_class, prod_r, tol, rsig, round_seed(round, 2));
    const int changed = nc2 > nc || mc2 > mc;
    nc = f->nf = nc2;
    mc = f->mf = mc2;
    if (too_little(mc, nc, m, n)) ok = 0;
    else if (!changed) break;
  }
  f->refinement_ok = ok;
  free(rsig);
  free(csig);
  free(prod_c);
  free(prod_r);
  if (!ok || (mc == m && nc == n)) return;
  build_folded(lp, tol, f);
  f->applied = 1;
}

void or_fold_free(or_fold* f) {
  if (!f) return;
  free(f->row_class);
  free(f->col_class);
  free(f->row_scale);
  free(f->col_scale);
  if (f->applied) {
    free(f->lp.A.row_ptr);
    free(f->lp.A.col_idx);
    free(f->lp.A.vals);
    free(f->lp.AT.row_ptr);
    free(f->lp.AT.col_idx);
    free(f->lp.AT.vals);
    free(f->lp.c);
    free(f->lp.l);
    free(f->lp.u);
    free(f->lp.AL);
    free(f->lp.AU);
  }
  memset(f, 0, sizeof(*f));
}

/* unfold_solution (folding_map.cu:285-307) */
void or_unfold(const or_fold* f, const double* xr, const double* yr, const double* zr, double* x, double* y,
               double* z) {
  for (int32_t j = 0; j < f->n; ++j) {
    const int32_t c = f->col_class[j] - 1;
    if (x && xr) x[j] = xr[c];
    if (z && zr) z[j] = zr[c] * f->col_scale[j];
  }
  for (int32_t i = 0; i < f->m; ++i)
    if (y && yr) y[i] = yr[f->row_class[i] - 1] * f->row_scale[i];
}
