/* lpgen.h -- seeded synthetic LP generator (host memory, CSR). */
#ifndef GP_LPGEN_H
#define GP_LPGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { LPGEN_DEGREE_UNIFORM = 0, LPGEN_DEGREE_ZIPF = 1 };

typedef struct lpgen_spec {
  int32_t m, n;
  uint64_t seed;
  int32_t degree_mode;
  double mean_degree;
  double zipf_s;
  int32_t zipf_cap;
  double frac_equality;
  double frac_free;
  double frac_singleton_rows;
  double frac_doubleton_rows;
  double frac_parallel_rows;
  double frac_duplicate_cols;
  double frac_fixed_cols;
  double frac_empty_cols;
  /* 1: integer coefficients, planted point, bounds, slacks and parallel-row
   * factors, so every activity and row shift of the LP is exact in binary64 */
  int32_t integral;
} lpgen_spec;

typedef struct lpgen_lp {
  int32_t m, n, nnz;
  int32_t* row_ptr;
  int32_t* col_idx;
  double* vals;
  double *c, *l, *u, *AL, *AU;
  double* x0; /* planted primal point (feasible)        */
  double* y0; /* planted dual point (sign-consistent)   */
  double obj_constant;
  double zipf_scale;
} lpgen_lp;

void lpgen_default_spec(lpgen_spec* spec, int32_t m, int32_t n, uint64_t seed);
int lpgen_generate(const lpgen_spec* spec, lpgen_lp* out);
/* Adams-Johnson LP linearisation of a seeded QAP of size n (2..40). */
int lpgen_qap(int32_t n, uint64_t seed, lpgen_lp* out);
void lpgen_free(lpgen_lp* lp);

#ifdef __cplusplus
}
#endif

#endif
