/*
 * lpgen.c -- seeded synthetic LP generator for the presolve workloads.
 *
 *   min c'x  s.t.  AL <= Ax <= AU,  l <= x <= u
 *
 * A planted point x0 satisfies every row and bound, so the LP is feasible;
 * free columns receive c_j = (A'y0)_j for a sign-consistent y0, so the dual is
 * feasible and the LP is bounded.  Planted structure (fractions of rows /
 * columns): singleton rows, doubleton equalities, parallel rows (scaled copies
 * of another row, factor U[0.5,4]), duplicate columns (exact copies of A_j,
 * c_j, l_j, u_j), fixed columns (l = u) and empty columns.
 *
 * Column degrees are either uniform in {d-2..d+2} or Zipf(s) truncated at a
 * cap and stochastically rescaled to a target mean.  Row positions are uniform
 * and distinct per column.  Values are N(0,1) (Box-Muller) resampled while
 * |a| < 1e-3.  Everything is driven by one splitmix64 seed; no library RNG.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lpgen.h"

typedef struct rng {
  uint64_t s;
} rng;

static uint64_t next_u64(rng* r) {
  uint64_t z = (r->s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static double unif(rng* r) { return (double)(next_u64(r) >> 11) * (1.0 / 9007199254740992.0); }
static double unif_pos(rng* r) { return ((double)(next_u64(r) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
static double unif_ab(rng* r, double a, double b) { return a + (b - a) * unif(r); }
static uint32_t below(rng* r, uint32_t n) { return (uint32_t)(((next_u64(r) >> 32) * (uint64_t)n) >> 32); }

static double normal(rng* r) {
  const double u1 = unif_pos(r), u2 = unif(r);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static double coeff(rng* r, int integral) {
  if (integral) {
    const double mag = (double)(below(r, 4) + 1);
    return unif(r) < 0.5 ? -mag : mag;
  }
  double a;
  do {
    a = normal(r);
  } while (fabs(a) < 1e-3);
  return a;
}

/* uniform on [a, b], or a uniform integer in [ceil(a), floor(b)] */
static double draw(rng* r, double a, double b, int integral) {
  if (!integral) return a + (b - a) * unif(r);
  const double lo = ceil(a), hi = floor(b);
  if (hi <= lo) return lo;
  double v = lo + floor((hi - lo + 1.0) * unif(r));
  return v > hi ? hi : v;
}

static void* xmalloc(size_t bytes) {
  void* p = malloc(bytes ? bytes : 1);
  if (!p) {
    fprintf(stderr, "lpgen: out of memory\n");
    abort();
  }
  return p;
}

/* ---- Zipf degree law ------------------------------------------------------ */
typedef struct zipf {
  int cap;
  double* cdf; /* cdf[k-1] = P(D <= k) */
  double scale;
} zipf;

static double zipf_mean(const zipf* z, double f) {
  double mean = 0.0, prev = 0.0;
  for (int k = 1; k <= z->cap; ++k) {
    const double pk = z->cdf[k - 1] - prev;
    prev = z->cdf[k - 1];
    const double x = f * k;
    mean += pk * (x >= 1.0 ? x : 1.0);
  }
  return mean;
}

static void zipf_init(zipf* z, double s, int cap, double target_mean) {
  z->cap = cap;
  z->cdf = (double*)xmalloc(sizeof(double) * (size_t)cap);
  double total = 0.0;
  for (int k = 1; k <= cap; ++k) total += pow((double)k, -s);
  double acc = 0.0;
  for (int k = 1; k <= cap; ++k) {
    acc += pow((double)k, -s) / total;
    z->cdf[k - 1] = acc;
  }
  z->cdf[cap - 1] = 1.0;
  /* E[max(1, sround(f*D))] = sum_k P(D=k) * max(f*k, 1); bisection on f */
  double lo = 0.0, hi = 1.0;
  while (zipf_mean(z, hi) < target_mean && hi < 1e6) hi *= 2.0;
  for (int it = 0; it < 100; ++it) {
    const double f = 0.5 * (lo + hi);
    if (zipf_mean(z, f) < target_mean) lo = f;
    else hi = f;
  }
  z->scale = 0.5 * (lo + hi);
}

static int zipf_draw(zipf* z, rng* r) {
  const double u = unif(r);
  int lo = 0, hi = z->cap - 1;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (z->cdf[mid] >= u) hi = mid;
    else lo = mid + 1;
  }
  const double x = z->scale * (double)(lo + 1);
  const double fl = floor(x);
  int d = (int)fl + (unif(r) < x - fl ? 1 : 0);
  return d < 1 ? 1 : d;
}

/* ---- generator --------------------------------------------------------------- */

void lpgen_default_spec(lpgen_spec* s, int32_t m, int32_t n, uint64_t seed) {
  memset(s, 0, sizeof(*s));
  s->m = m;
  s->n = n;
  s->seed = seed;
  s->degree_mode = LPGEN_DEGREE_UNIFORM;
  s->mean_degree = 6.0;
  s->zipf_s = 1.5;
  s->zipf_cap = 500;
  s->frac_equality = 0.40;
  s->frac_free = 0.10;
  s->frac_singleton_rows = 0.05;
  s->frac_doubleton_rows = 0.05;
  s->frac_parallel_rows = 0.02;
  s->frac_duplicate_cols = 0.02;
  s->frac_fixed_cols = 0.03;
  s->frac_empty_cols = 0.0;
}

enum { COL_REGULAR = 0, COL_DUPLICATE = 1, COL_EMPTY = 2 };
enum { ROW_REGULAR = 0, ROW_SINGLETON = 1, ROW_DOUBLETON = 2, ROW_PARALLEL = 3 };

typedef struct coo {
  int64_t count, cap;
  int32_t* row;
  int32_t* col;
  double* val;
} coo;

static void coo_push(coo* c, int32_t i, int32_t j, double v) {
  if (c->count == c->cap) {
    c->cap = c->cap ? 2 * c->cap : 1024;
    c->row = (int32_t*)realloc(c->row, sizeof(int32_t) * (size_t)c->cap);
    c->col = (int32_t*)realloc(c->col, sizeof(int32_t) * (size_t)c->cap);
    c->val = (double*)realloc(c->val, sizeof(double) * (size_t)c->cap);
    if (!c->row || !c->col || !c->val) {
      fprintf(stderr, "lpgen: out of memory\n");
      abort();
    }
  }
  c->row[c->count] = i;
  c->col[c->count] = j;
  c->val[c->count] = v;
  c->count++;
}

int lpgen_generate(const lpgen_spec* spec, lpgen_lp* out) {
  memset(out, 0, sizeof(*out));
  const int32_t m = spec->m, n = spec->n;
  if (m <= 0 || n <= 0) return 1;
  rng r = {spec->seed * 0x632be59bd9b4e019ULL + 0x1234567ULL};

  uint8_t* ctype = (uint8_t*)xmalloc((size_t)n);
  uint8_t* rtype = (uint8_t*)xmalloc((size_t)m);
  double* l = (double*)xmalloc(sizeof(double) * (size_t)n);
  double* u = (double*)xmalloc(sizeof(double) * (size_t)n);
  double* c = (double*)xmalloc(sizeof(double) * (size_t)n);
  double* x0 = (double*)xmalloc(sizeof(double) * (size_t)n);
  int32_t* dup_src = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)n);
  uint8_t* is_free = (uint8_t*)xmalloc((size_t)n);

  /* column structure and bounds */
  for (int32_t j = 0; j < n; ++j) {
    const double t = unif(&r);
    ctype[j] = t < spec->frac_empty_cols ? COL_EMPTY
               : t < spec->frac_empty_cols + spec->frac_duplicate_cols ? COL_DUPLICATE
                                                                         : COL_REGULAR;
    const double b = unif(&r);
    const double ub = draw(&r, 1.0, 100.0, spec->integral);
    c[j] = unif_ab(&r, -1.0, 1.0);
    is_free[j] = 0;
    if (b < spec->frac_fixed_cols) {
      x0[j] = draw(&r, 0.0, ub, spec->integral);
      l[j] = u[j] = x0[j];
    } else if (b < spec->frac_fixed_cols + spec->frac_free) {
      l[j] = -INFINITY;
      u[j] = INFINITY;
      x0[j] = draw(&r, -10.0, 10.0, spec->integral);
      is_free[j] = 1;
    } else {
      l[j] = 0.0;
      u[j] = ub;
      x0[j] = draw(&r, 0.0, ub, spec->integral);
    }
  }
  /* a regular source for every duplicate column */
  int32_t n_regular = 0;
  int32_t* regular_cols = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)n);
  for (int32_t j = 0; j < n; ++j)
    if (ctype[j] == COL_REGULAR) regular_cols[n_regular++] = j;
  if (n_regular == 0) {
    ctype[0] = COL_REGULAR;
    regular_cols[n_regular++] = 0;
  }
  for (int32_t j = 0; j < n; ++j) {
    dup_src[j] = -1;
    if (ctype[j] == COL_DUPLICATE) {
      const int32_t s = regular_cols[below(&r, (uint32_t)n_regular)];
      dup_src[j] = s;
    }
  }
  /* row types */
  int32_t n_reg_rows = 0;
  int32_t* reg_rows = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)m);
  for (int32_t i = 0; i < m; ++i) {
    const double t = unif(&r);
    rtype[i] = t < spec->frac_singleton_rows ? ROW_SINGLETON
               : t < spec->frac_singleton_rows + spec->frac_doubleton_rows ? ROW_DOUBLETON
               : t < spec->frac_singleton_rows + spec->frac_doubleton_rows + spec->frac_parallel_rows
                   ? ROW_PARALLEL
                   : ROW_REGULAR;
    if (rtype[i] == ROW_REGULAR) reg_rows[n_reg_rows++] = i;
  }
  if (n_reg_rows == 0) {
    rtype[0] = ROW_REGULAR;
    reg_rows[n_reg_rows++] = 0;
  }

  /* regular column entries into regular rows */
  zipf z;
  memset(&z, 0, sizeof(z));
  if (spec->degree_mode == LPGEN_DEGREE_ZIPF) zipf_init(&z, spec->zipf_s, spec->zipf_cap, spec->mean_degree);
  coo e;
  memset(&e, 0, sizeof(e));
  int32_t* picked = (int32_t*)xmalloc(sizeof(int32_t) * 1024);
  int64_t* col_first = (int64_t*)xmalloc(sizeof(int64_t) * (size_t)n);
  int32_t* col_count = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)n);
  const int md = (int)floor(spec->mean_degree + 0.5);
  for (int32_t j = 0; j < n; ++j) {
    col_first[j] = e.count;
    col_count[j] = 0;
    if (ctype[j] != COL_REGULAR) continue;
    int d;
    if (spec->degree_mode == LPGEN_DEGREE_ZIPF) d = zipf_draw(&z, &r);
    else d = md - 2 + (int)below(&r, 5);
    if (d < 1) d = 1;
    if (d > 1000) d = 1000;
    if (d > n_reg_rows) d = n_reg_rows;
    int got = 0;
    while (got < d) {
      const int32_t i = reg_rows[below(&r, (uint32_t)n_reg_rows)];
      int dup = 0;
      for (int k = 0; k < got; ++k)
        if (picked[k] == i) {
          dup = 1;
          break;
        }
      if (dup) continue;
      picked[got++] = i;
      coo_push(&e, i, j, coeff(&r, spec->integral));
    }
    col_count[j] = d;
  }
  /* duplicate columns copy entries and c, l, u of their source */
  for (int32_t j = 0; j < n; ++j) {
    if (ctype[j] != COL_DUPLICATE) continue;
    const int32_t s = dup_src[j];
    const int64_t f = col_first[s];
    for (int32_t k = 0; k < col_count[s]; ++k) coo_push(&e, e.row[f + k], j, e.val[f + k]);
    c[j] = c[s];
    l[j] = l[s];
    u[j] = u[s];
    is_free[j] = is_free[s];
    x0[j] = isfinite(l[j]) ? (l[j] == u[j] ? l[j] : draw(&r, l[j], u[j], spec->integral))
                           : draw(&r, -10.0, 10.0, spec->integral);
  }
  /* non-empty columns for planted rows */
  int32_t n_live = 0;
  int32_t* live_cols = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)n);
  for (int32_t j = 0; j < n; ++j)
    if (ctype[j] != COL_EMPTY) live_cols[n_live++] = j;
  for (int32_t i = 0; i < m; ++i) {
    if (rtype[i] == ROW_SINGLETON) {
      coo_push(&e, i, live_cols[below(&r, (uint32_t)n_live)], coeff(&r, spec->integral));
    } else if (rtype[i] == ROW_DOUBLETON) {
      int32_t a = live_cols[below(&r, (uint32_t)n_live)], b;
      do {
        b = live_cols[below(&r, (uint32_t)n_live)];
      } while (b == a && n_live > 1);
      if (a > b) {
        const int32_t t = a;
        a = b;
        b = t;
      }
      coo_push(&e, i, a, coeff(&r, spec->integral));
      if (b != a) coo_push(&e, i, b, coeff(&r, spec->integral));
    }
  }

  /* CSR of everything generated so far (stable counting sort by row) */
  int64_t* rp = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));
  for (int64_t k = 0; k < e.count; ++k) rp[e.row[k] + 1]++;
  for (int32_t i = 0; i < m; ++i) rp[i + 1] += rp[i];
  int64_t nnz0 = rp[m];
  int32_t* ci0 = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)nnz0);
  double* va0 = (double*)xmalloc(sizeof(double) * (size_t)nnz0);
  {
    int64_t* nx = (int64_t*)xmalloc(sizeof(int64_t) * (size_t)m);
    memcpy(nx, rp, sizeof(int64_t) * (size_t)m);
    for (int64_t k = 0; k < e.count; ++k) {
      const int64_t q = nx[e.row[k]]++;
      ci0[q] = e.col[k];
      va0[q] = e.val[k];
    }
    free(nx);
  }
  free(e.row);
  free(e.col);
  free(e.val);

  /* parallel rows: scaled copies of a non-empty regular row */
  int64_t extra = 0;
  int32_t* par_src = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)m);
  double* par_f = (double*)xmalloc(sizeof(double) * (size_t)m);
  for (int32_t i = 0; i < m; ++i) {
    par_src[i] = -1;
    if (rtype[i] != ROW_PARALLEL) continue;
    int32_t s = -1;
    for (int tries = 0; tries < 64; ++tries) {
      const int32_t t = reg_rows[below(&r, (uint32_t)n_reg_rows)];
      if (rp[t + 1] > rp[t]) {
        s = t;
        break;
      }
    }
    par_src[i] = s;
    par_f[i] = spec->integral ? (double)(below(&r, 3) + 1) * (unif(&r) < 0.25 ? -1.0 : 1.0)
                              : unif_ab(&r, 0.5, 4.0);
    if (s >= 0) extra += rp[s + 1] - rp[s];
  }
  const int64_t nnz = nnz0 + extra;
  if (nnz > INT32_MAX) {
    fprintf(stderr, "lpgen: nnz exceeds int32 offsets\n");
    abort();
  }
  int32_t* row_ptr = (int32_t*)xmalloc(sizeof(int32_t) * ((size_t)m + 1));
  int32_t* col_idx = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)nnz);
  double* vals = (double*)xmalloc(sizeof(double) * (size_t)nnz);
  int64_t q = 0;
  row_ptr[0] = 0;
  for (int32_t i = 0; i < m; ++i) {
    const int64_t start = q;
    if (par_src[i] >= 0) {
      const int32_t s = par_src[i];
      for (int64_t p = rp[s]; p < rp[s + 1]; ++p) {
        col_idx[q] = ci0[p];
        vals[q] = va0[p] * par_f[i];
        ++q;
      }
    } else {
      for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
        col_idx[q] = ci0[p];
        vals[q] = va0[p];
        ++q;
      }
    }
    /* sort the row by column (insertion sort; rows are short) */
    for (int64_t a = start + 1; a < q; ++a) {
      const int32_t cj = col_idx[a];
      const double cv = vals[a];
      int64_t b = a - 1;
      while (b >= start && col_idx[b] > cj) {
        col_idx[b + 1] = col_idx[b];
        vals[b + 1] = vals[b];
        --b;
      }
      col_idx[b + 1] = cj;
      vals[b + 1] = cv;
    }
    row_ptr[i + 1] = (int32_t)q;
  }
  free(rp);
  free(ci0);
  free(va0);

  /* row sides from the planted point; sign-consistent dual point y0 */
  double* AL = (double*)xmalloc(sizeof(double) * (size_t)m);
  double* AU = (double*)xmalloc(sizeof(double) * (size_t)m);
  double* y0 = (double*)xmalloc(sizeof(double) * (size_t)m);
  for (int32_t i = 0; i < m; ++i) {
    double act = 0.0;
    for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) act += vals[p] * x0[col_idx[p]];
    const double t = unif(&r);
    const double slack = draw(&r, 0.0, 10.0, spec->integral);
    const double yv = unif_ab(&r, 0.0, 1.0);
    if (rtype[i] == ROW_DOUBLETON || t < spec->frac_equality) {
      AL[i] = AU[i] = act;
      y0[i] = 2.0 * yv - 1.0;
    } else if (t < spec->frac_equality + 0.5 * (1.0 - spec->frac_equality)) {
      AL[i] = -INFINITY;
      AU[i] = act + slack;
      y0[i] = -yv; /* <= row: y <= 0 under c - A'y - z = 0 */
    } else {
      AL[i] = act - slack;
      AU[i] = INFINITY;
      y0[i] = yv; /* >= row: y >= 0 */
    }
  }
  /* free columns: c_j = (A'y0)_j (z_j = 0), so y0 is dual feasible */
  double* aty = (double*)calloc((size_t)n, sizeof(double));
  for (int32_t i = 0; i < m; ++i)
    for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) aty[col_idx[p]] += vals[p] * y0[i];
  for (int32_t j = 0; j < n; ++j)
    if (is_free[j]) c[j] = aty[j];
  /* bounded duplicate columns keep the exact cost of their source; a free
   * duplicate keeps c = (A'y0)_j, since planted rows may have made it differ
   * from its source and any other cost would leave the LP dual infeasible */
  for (int32_t j = 0; j < n; ++j)
    if (ctype[j] == COL_DUPLICATE && !is_free[j]) c[j] = c[dup_src[j]];
  free(aty);

  out->m = m;
  out->n = n;
  out->nnz = (int32_t)nnz;
  out->row_ptr = row_ptr;
  out->col_idx = col_idx;
  out->vals = vals;
  out->c = c;
  out->l = l;
  out->u = u;
  out->AL = AL;
  out->AU = AU;
  out->x0 = x0;
  out->y0 = y0;
  out->obj_constant = 0.0;
  out->zipf_scale = z.scale;

  free(ctype);
  free(rtype);
  free(dup_src);
  free(is_free);
  free(regular_cols);
  free(reg_rows);
  free(picked);
  free(col_first);
  free(col_count);
  free(live_cols);
  free(par_src);
  free(par_f);
  free(z.cdf);
  return 0;
}

/* ---- Adams-Johnson QAP LP ------------------------------------------------- */
/* lpgen_qap restates the LP that scripts/generate_qap20.jl:25-109 writes as MPS
 * (same row order, column order, signs, sides and free bounds):
 *   rows [0, n^4)                 -y_q <= 0                       (explicit bound rows)
 *   rows n^4 + [0, n)             sum_j x[i,j] = 1
 *   rows n^4 + n + [0, n)         sum_i x[i,j] = 1
 *   rows n^4 + 2n + idx3(t,i,j)   sum_a y[a,t,i,j] - x[i,j] = 0
 *   rows ... + n^3 + idx3(a,b,c)  sum_b' y[a,b',b,c] - x[b,c] = 0
 *   rows ... + 2n^3 + q           y_q - y_rq = 0 at min(q,rq), its negation at
 *                                 max(q,rq); empty (= 0) when q == rq
 *   rows ... + n^4 + idx2(i,j)    0 <= x[i,j] <= 1                (L row with range 1)
 * with idx2/idx3/idx4 the column-major indices of the script (0-based here),
 * columns x[i,j] at idx2(i,j) and y[i,j,k,l] at n^2 + idx4(i,j,k,l), all free,
 * and cost flow[i][k] * distance[j][l] on y[i,j,k,l].  flow and distance are
 * seeded integers U{0..100} with zero diagonals; each off-diagonal flow is
 * zero with probability 1/2.  x0 = 1/n, y0 = 1/n^2 is a feasible point. */
int lpgen_qap(int32_t nq, uint64_t seed, lpgen_lp* out) {
  memset(out, 0, sizeof(*out));
  if (nq < 2 || nq > 100) return 1;
  const int64_t n = nq, n2 = n * n, n3 = n2 * n, n4 = n3 * n;
  const int64_t m64 = 2 * n4 + 2 * n3 + n2 + 2 * n, ncols64 = n4 + n2;
  const int64_t nnz64 = n2 * (2 + 2 * n + 1) + 3 * n4 + 2 * (n4 - n2);
  if (m64 > INT32_MAX || nnz64 > INT32_MAX) return 1;
  rng r = {seed * 0x2545f4914f6cdd1dULL + 0x51ed270b27a1f3c5ULL};
  int32_t* flow = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)n2);
  int32_t* dist = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)n2);
  for (int64_t a = 0; a < n; ++a)
    for (int64_t b = 0; b < n; ++b) {
      const int zero_flow = unif(&r) < 0.5;
      const int32_t f = (int32_t)below(&r, 101), d = (int32_t)below(&r, 101);
      flow[a * n + b] = (a == b || zero_flow) ? 0 : f;
      dist[a * n + b] = a == b ? 0 : d;
    }
  const int32_t m = (int32_t)m64, ncols = (int32_t)ncols64;
  const int64_t eq0 = n4, sym0 = n4 + 2 * n + 2 * n3, rng0 = n4 + 2 * n + 2 * n3 + n4;
  /* column-major entry lists, turned into CSR by a counting pass over rows */
  int32_t* cnt = (int32_t*)calloc((size_t)m + 1, sizeof(int32_t));
  int32_t* crow = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)nnz64);
  double* cval = (double*)xmalloc(sizeof(double) * (size_t)nnz64);
  int32_t* cptr = (int32_t*)xmalloc(sizeof(int32_t) * ((size_t)ncols + 1));
  int64_t e = 0;
#define QAP_PUT(row, v) (crow[e] = (int32_t)(row), cval[e] = (v), cnt[(row) + 1]++, ++e)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i) {
      cptr[i + n * j] = (int32_t)e;
      QAP_PUT(eq0 + i, 1.0);
      QAP_PUT(eq0 + n + j, 1.0);
      for (int64_t t = 0; t < n; ++t) QAP_PUT(eq0 + 2 * n + (t + n * i + n2 * j), -1.0);
      for (int64_t t = 0; t < n; ++t) QAP_PUT(eq0 + 2 * n + n3 + (t + n * i + n2 * j), -1.0);
      QAP_PUT(rng0 + (i + n * j), 1.0);
    }
  double* c = (double*)calloc((size_t)ncols, sizeof(double));
  for (int64_t l = 0; l < n; ++l)
    for (int64_t k = 0; k < n; ++k)
      for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
          const int64_t q = i + n * j + n2 * k + n3 * l, rq = k + n * l + n2 * i + n3 * j;
          cptr[n2 + q] = (int32_t)e;
          QAP_PUT(q, -1.0);
          QAP_PUT(eq0 + 2 * n + (j + n * k + n2 * l), 1.0);
          QAP_PUT(eq0 + 2 * n + n3 + (i + n * k + n2 * l), 1.0);
          if (q != rq) {
            const int64_t first = q < rq ? q : rq, second = q < rq ? rq : q;
            const double sg = q < rq ? 1.0 : -1.0;
            QAP_PUT(sym0 + first, sg);
            QAP_PUT(sym0 + second, -sg);
          }
          c[n2 + q] = (double)flow[i * n + k] * (double)dist[j * n + l];
        }
#undef QAP_PUT
  cptr[ncols] = (int32_t)e;
  for (int32_t i = 0; i < m; ++i) cnt[i + 1] += cnt[i];
  int32_t* row_ptr = (int32_t*)xmalloc(sizeof(int32_t) * ((size_t)m + 1));
  memcpy(row_ptr, cnt, sizeof(int32_t) * ((size_t)m + 1));
  int32_t* col_idx = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)e);
  double* vals = (double*)xmalloc(sizeof(double) * (size_t)e);
  for (int32_t col = 0; col < ncols; ++col) /* columns in order: rows come out sorted */
    for (int32_t p = cptr[col]; p < cptr[col + 1]; ++p) {
      const int32_t pos = cnt[crow[p]]++;
      col_idx[pos] = col;
      vals[pos] = cval[p];
    }
  double* AL = (double*)xmalloc(sizeof(double) * (size_t)m);
  double* AU = (double*)xmalloc(sizeof(double) * (size_t)m);
  for (int64_t i = 0; i < m; ++i) {
    if (i < eq0) {
      AL[i] = -INFINITY;
      AU[i] = 0.0;
    } else if (i < rng0) {
      AL[i] = AU[i] = (i < eq0 + 2 * n) ? 1.0 : 0.0;
    } else {
      AL[i] = 0.0;
      AU[i] = 1.0;
    }
  }
  double* l = (double*)xmalloc(sizeof(double) * (size_t)ncols);
  double* u = (double*)xmalloc(sizeof(double) * (size_t)ncols);
  double* x0 = (double*)xmalloc(sizeof(double) * (size_t)ncols);
  for (int64_t j = 0; j < ncols; ++j) {
    l[j] = -INFINITY;
    u[j] = INFINITY;
    x0[j] = j < n2 ? 1.0 / (double)n : 1.0 / (double)n2;
  }
  out->m = m;
  out->n = ncols;
  out->nnz = (int32_t)e;
  out->row_ptr = row_ptr;
  out->col_idx = col_idx;
  out->vals = vals;
  out->c = c;
  out->l = l;
  out->u = u;
  out->AL = AL;
  out->AU = AU;
  out->x0 = x0;
  out->y0 = (double*)calloc((size_t)m, sizeof(double));
  out->obj_constant = 0.0;
  out->zipf_scale = 0.0;
  free(flow);
  free(dist);
  free(cnt);
  free(crow);
  free(cval);
  free(cptr);
  return 0;
}

void lpgen_free(lpgen_lp* lp) {
  free(lp->row_ptr);
  free(lp->col_idx);
  free(lp->vals);
  free(lp->c);
  free(lp->l);
  free(lp->u);
  free(lp->AL);
  free(lp->AU);
  free(lp->x0);
  free(lp->y0);
  memset(lp, 0, sizeof(*lp));
}
