/* Internal helpers of the CPU restatement.  TEST INFRASTRUCTURE ONLY. */
#ifndef GP_ORACLE_INTERNAL_H
#define GP_ORACLE_INTERNAL_H

#include "oracle.h"

#include <math.h>
#include <stddef.h>

#define OR_SUM_BLOCK 256

typedef struct or_neu {
  double s, c;
} or_neu;

/* Neumaier step: s <- s + x with the rounding error accumulated in c. */
static inline void or_neu_add(or_neu* a, double x) {
  const double t = a->s + x;
  if (fabs(a->s) >= fabs(x))
    a->c += (a->s - t) + x;
  else
    a->c += (x - t) + a->s;
  a->s = t;
}

static inline void or_neu_merge(or_neu* a, or_neu b) {
  or_neu_add(a, b.s);
  a->c += b.c;
}

static inline double or_neu_value(or_neu a) { return a.s + a.c; }

/* fmax / fmin with C-library tie semantics (ties return the first argument),
 * spelled out so CPU and GPU agree on signed zeros. */
static inline double or_cmax(double x, double y) { return (x >= y || y != y) ? x : y; }
static inline double or_cmin(double x, double y) { return (x <= y || y != y) ? x : y; }

/* Blocked row accumulator: or_rowacc_pos(k) must be called for every stored
 * position k = 0, 1, ... of the row (skipped entries included) before adding
 * that position's term. */
typedef struct or_rowacc {
  or_neu total, blk;
} or_rowacc;

static inline void or_rowacc_init(or_rowacc* r) {
  r->total.s = r->total.c = r->blk.s = r->blk.c = 0.0;
}
static inline void or_rowacc_pos(or_rowacc* r, int64_t k) {
  if (k > 0 && k % OR_SUM_BLOCK == 0) {
    or_neu_merge(&r->total, r->blk);
    r->blk.s = r->blk.c = 0.0;
  }
}
static inline void or_rowacc_add(or_rowacc* r, double x) { or_neu_add(&r->blk, x); }
static inline double or_rowacc_value(or_rowacc* r) {
  or_neu t = r->total;
  or_neu_merge(&t, r->blk);
  return or_neu_value(t);
}

/* Max / min under the IEEE total order on non-NaN values (-0 < +0), so that
 * concurrent candidate updates have one order-independent result. */
static inline int or_total_greater(double a, double b) {
  return a > b || (a == b && signbit(b) && !signbit(a));
}
static inline void or_cand_max(double* slot, double v) {
  if (or_total_greater(v, *slot)) *slot = v;
}
static inline void or_cand_min(double* slot, double v) {
  if (or_total_greater(*slot, v)) *slot = v;
}

/* Objective of fixed columns: sequential long-double (x87 extended) sum of
 * c[j] * v[j] over selected (fixed and not excluded) columns in increasing
 * order, rounded to double once (fixed_variable_tape.cu:305-325). */
double or_fixed_objective_ld(const uint8_t* fixed, const uint8_t* excluded, const double* c, const double* v,
                             int32_t n);

void* or_xmalloc(size_t bytes);
void* or_xcalloc(size_t count, size_t size);
void* or_xrealloc(void* p, size_t bytes);

void or_globalize_record(int32_t type, int32_t* ix, int64_t size, const int32_t* rows, int32_t m,
                         const int32_t* cols, int32_t n);
void or_tape_append_global(or_tape* dst, const or_tape* src, const int32_t* rows, int32_t m,
                           const int32_t* cols, int32_t n);

/* Individual rules (cpp/src/presolve/rules/rule_<name>.cu). */
int or_rule_fixed_variables(gp_plan* plan, const gp_device_lp* lp, const gp_params* p);
int or_rule_empty_columns(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                          const gp_params* p);
int or_rule_empty_rows(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s, const gp_params* p);
int or_rule_singleton_rows(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                           const gp_params* p);
int or_rule_redundant_rows(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                           const gp_params* p);
int or_rule_bound_strengthening(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                                const gp_params* p);
int or_rule_singleton_columns(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                              const gp_params* p);
int or_rule_dual_fix(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s, const gp_params* p);
int or_rule_doubleton_equations(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                                const gp_params* p);
int or_rule_parallel_rows(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                          const gp_params* p);
int or_rule_parallel_columns(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                             const gp_params* p);

#endif
