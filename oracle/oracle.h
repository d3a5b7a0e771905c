/*
 * oracle.h -- CPU restatement of the reference presolve / postsolve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * engine: only tests, the smoke check and the CPU-baseline leg of the bench
 * may load it.  The product path never calls it.
 *
 * It uses the gpupre.h structs with HOST pointers.  Every function names the
 * reference file:line whose behaviour it restates.  Floating-point order is
 * defined exactly (see or_row_sum / or_det_sum) and is the order the HIP
 * kernels reproduce, so masks, counts, maps and tapes compare bit-exactly.
 */
#ifndef GP_ORACLE_H
#define GP_ORACLE_H

#include "../include/gpupre.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Growable host tape: PostsolveTape (presolve_structs.hpp:320-332). */
typedef struct or_tape {
  int64_t count, icount, vcount;
  int64_t cap, icap, vcap;
  int32_t* types;
  gp_offset* istarts; /* count+1 */
  gp_offset* vstarts; /* count+1 */
  uint8_t* modes;
  int32_t* idx;
  double* vals;
} or_tape;

void or_tape_init(or_tape* t);
void or_tape_free(or_tape* t);
void or_tape_push(or_tape* t, int32_t type, uint8_t mode, const int32_t* idx, int64_t ni,
                  const double* vals, int64_t nv);

/* Numerics shared with the HIP kernels. */
double or_row_sum_finite(const double* terms, int64_t n); /* test hook */
double or_det_sum(const double* x, int64_t n);

/* Row activity summary over live entries (gpu_presolve_kernels.cu:96-148). */
void or_row_activity_summary(const gp_csr* A, const double* l, const double* u, double zero_tol,
                             const uint8_t* keep_row, const uint8_t* keep_col,
                             double* min_fin, double* max_fin, int32_t* min_ninf, int32_t* max_pinf);

/* Live statistics (gpu_presolve_kernels.cu:422-517). */
int or_stats_create(int32_t m, int32_t n, gp_stats* s);
void or_stats_destroy(gp_stats* s);
void or_refresh_live_stats(gp_stats* s, const gp_device_lp* lp, const gp_plan* plan);

/* Plan (gpu_presolve.cu:567-706). */
int or_plan_create(const gp_device_lp* lp, gp_plan* plan);
void or_plan_destroy(gp_plan* plan);
void or_plan_enable_tape(gp_plan* plan);

/* FixedVariable records + long-double objective (fixed_variable_tape.cu:141-327). */
double or_fixed_variable_tape(or_tape* t, const uint8_t* fixed, const uint8_t* excluded, const double* val,
                              const uint8_t* keep_row, const double* c, const gp_csr* AT, int32_t n, int objective);

/* compute_parallel_row_hashes (rule_parallel_rows.cu:141-196,583-627). */
void or_parallel_row_hashes(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, double tol,
                            uint64_t* out);

/* Rules. */
/* Structural Startup rules (oracle_startup.c): gp_startup_rule on a host
 * plan.  Returns 1 when applicable, 0 when not, -1 on a bad argument. */
int or_startup_rule(int32_t rule, const gp_device_lp* lp, const gp_params* p, gp_plan* plan, int pending,
                    int32_t* count);

int or_rule_apply(int32_t rule, gp_plan* plan, const gp_device_lp* lp, const gp_stats* stats,
                  const gp_params* params);

/* Matrix primitives (transpose64.cu:69-117, gpu_presolve.cu:1793-1930). */
void or_transpose(const gp_csr* A, gp_csr* AT);
void or_compact(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, gp_csr* out);
void or_csr_free(gp_csr* A);
void or_lp_free(gp_device_lp* lp);
int or_lp_copy(const gp_device_lp* src, gp_device_lp* dst);

/* Full pipeline (gpu_presolve.cu:2081-2645 with folding off). */
typedef struct or_result {
  gp_device_lp reduced; /* host arrays, owned */
  int32_t m0, n0, m1, n1;
  int32_t* row_red2org;
  int32_t* col_red2org;
  int32_t fixed_count, fixed_cap;
  int32_t* fixed_idx;
  double* fixed_val;
  or_tape tape;
  int32_t iterations;
  int32_t stopped_at_iteration_limit;
  int32_t has_change;
  int32_t has_infeasible;
  int32_t has_unbounded;
  double obj_constant_delta;
  int32_t trace_count, trace_cap;
  gp_round_trace* trace;
  int32_t primal_only; /* GP_PRIMAL_ONLY_* bits */
} or_result;

int or_presolve(const gp_device_lp* lp, const gp_params* params, or_result* out);
void or_result_free(or_result* r);

/* Postsolve (gpu_postsolve.cu:535-670). x,y,z in reduced order; outputs in
 * original order.  `original` may be NULL (no original-model handlers). */
/* postsolve_gpu over a caller-built record (gpu_postsolve.cu:400-670):
 * returns -1 when the tape or the structural timeline is malformed. */
int or_postsolve_record(const gp_host_record* r, const gp_device_lp* original, const double* original_l,
                        const double* x, const double* y, const double* z, double* x_out, double* y_out,
                        double* z_out);
int or_postsolve(const or_result* r, const gp_device_lp* original, const double* x,
                 const double* y, const double* z, double* x_out, double* y_out, double* z_out);

/* Learned pass scheduler (oracle_mlp.c; no reference counterpart). */
int or_rule_group(int32_t rule);
void or_schedule_features(const gp_round_trace* hist, int32_t done, int32_t phase, int64_t nnz_now,
                          int64_t nnz0, int32_t m0, int32_t n0, int32_t bound_only, int32_t limit,
                          float* features);
void or_mlp_forward(const float* params, const float* x, int32_t batch, float* logits);
void or_mlp_loss_grad(const float* params, const float* x, const float* y, const float* mask, int32_t batch,
                      double* loss, float* grad);
void or_mlp_adam_step(float* params, const float* grad, float* m, float* v, int32_t t, float lr, float beta1,
                      float beta2, float eps);

/* LP folding by colour refinement (cpp/src/folding/, oracle_fold.c). */
typedef struct or_fold {
  int32_t applied, refinement_ok, rounds;
  int32_t m, n, mf, nf;
  int32_t* row_class; /* [m] 1-based */
  int32_t* col_class; /* [n] 1-based */
  double* row_scale;  /* [m] when applied */
  double* col_scale;  /* [n] when applied */
  gp_device_lp lp;    /* folded LP (host arrays) when applied */
} or_fold;
void or_fold_lp(const gp_device_lp* lp, double tol, or_fold* out);
void or_fold_free(or_fold* f);
void or_unfold(const or_fold* f, const double* x_red, const double* y_red, const double* z_red, double* x,
               double* y, double* z);

#ifdef __cplusplus
}
#endif

#endif
