/*
 * gpupre.h -- C ABI of the MI355X LP presolve / postsolve engine.
 *
 * Every entry point is extern "C" with plain pointers and sizes.  Matrices and
 * vectors passed as `gp_device_lp` live in device (HBM) memory and stay owned
 * by the caller; the library never modifies them.  Errors are integer status
 * codes; gp_last_error() returns a thread-local message for the last failure.
 * Infeasible / unbounded models are flags in gp_summary, not errors.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   gp_preprocess      <- gpu_preprocessor::preprocess()
 *                         cpp/include/gpu_preprocessor/gpu_preprocessor.hpp:22-24
 *   gp_postprocess     <- gpu_preprocessor::postprocess()
 *                         cpp/include/gpu_preprocessor/gpu_preprocessor.hpp:30-32
 *   gp_free_summary    <- gpu_preprocessor::free_summary()
 *                         cpp/include/gpu_preprocessor/gpu_preprocessor.hpp:35-39
 *   gp_device_lp       <- presolve::LPInfoGpu / DeviceCsrMatrix
 *                         cpp/include/gpu_preprocessor/presolve/presolve_structs.hpp:160-180
 *   gp_params          <- presolve::PresolveParams   (same field names, same defaults)
 *                         cpp/include/gpu_preprocessor/presolve/presolve_structs.hpp:28-158
 *   gp_summary         <- folding::FoldingPipelineSummary + presolve::GpuPresolveSummary
 *                         cpp/include/gpu_preprocessor/folding/folding_structs.hpp:94-110,
 *                         cpp/include/gpu_preprocessor/presolve/gpu_presolve.hpp:14-30
 *   gp_plan / gp_stats <- presolve::PresolvePlanGpu / PresolveStatsGpu
 *                         cpp/include/gpu_preprocessor/presolve/presolve_structs.hpp:182-205,378-439
 *   gp_rule_*          <- apply_rule_*(PresolvePlanGpu&, const LPInfoGpu&, const PresolveStatsGpu&,
 *                         const PresolveParams&) in cpp/src/presolve/rules/rule_<name>.cu
 *   gp_refresh_live_stats <- refresh_live_stats  cpp/src/presolve/gpu_presolve_kernels.cu:493-517
 *   gp_transpose_csr   <- transpose_csr64        cpp/src/presolve/transpose64.cu:69-117
 *   gp_compact_csr     <- compact_csr_by_masks_gpu cpp/src/presolve/gpu_presolve.cu:1793-1930
 */
#ifndef GPUPRE_H
#define GPUPRE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP_ABI_VERSION 1

/* NnzOffset of the reference build (int32 default, offset.hpp:7-19). */
typedef int32_t gp_offset;

/* Status codes. */
enum {
  GP_OK = 0,
  GP_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference       */
  GP_ERR_LENGTH = 2,           /* std::length_error (offset overflow)           */
  GP_ERR_RUNTIME = 3,          /* std::runtime_error (HIP failure, bad tape)    */
  GP_ERR_NO_DEVICE = 4,
  GP_ERR_ALLOCATION = 5
};

/* Postsolve reduction types (PostsolveReductionType, presolve_structs.hpp:268-282). */
enum {
  GP_RED_FIXED_VARIABLE = 0,
  GP_RED_INFINITE_BOUND_DUAL_FIX = 1,
  GP_RED_SINGLETON_COLUMN_SUBSTITUTION = 2,
  GP_RED_PARALLEL_COLUMN = 3,
  GP_RED_PARALLEL_ROW = 4,
  GP_RED_DELETED_ROW = 5,
  GP_RED_ROW_LOWER_BOUND_CHANGE = 8,
  GP_RED_ROW_UPPER_BOUND_CHANGE = 9,
  GP_RED_ROW_DUAL_SHIFT = 10,
  GP_RED_VARIABLE_BOUND_CHANGE = 11,
  GP_RED_ROW_IMPLIED_BOUND_CHANGE = 12,
  GP_RED_DOUBLETON_EQUATION = 13,
  GP_RED_FME_COL = 14
};

/* PostsolveDualMode (presolve_structs.hpp:284-288). */
enum { GP_DUAL_NONE = 0, GP_DUAL_EXACT = 1, GP_DUAL_MINIMAL = 2 };

/* Rule identifiers used by gp_rule_apply, the per-round trace and the learned
 * pass scheduler's output layer. */
enum {
  GP_RULE_FIXED_VARIABLES = 0,
  GP_RULE_EMPTY_COLUMNS = 1,
  GP_RULE_EMPTY_ROWS = 2,
  GP_RULE_SINGLETON_ROWS = 3,
  GP_RULE_SINGLETON_COLUMNS = 4,
  GP_RULE_BOUND_STRENGTHENING = 5,
  GP_RULE_REDUNDANT_ROWS = 6,
  GP_RULE_DUAL_FIX = 7,
  GP_RULE_DOUBLETON_EQUATIONS = 8,
  GP_RULE_PARALLEL_ROWS = 9,
  GP_RULE_PARALLEL_COLUMNS = 10,
  GP_RULE_COUNT = 11
};

/* Scheduler phases (gpu_presolve.cu:2466-2475). */
enum { GP_PHASE_CLEANUP = 0, GP_PHASE_PROPAGATION = 1, GP_PHASE_AGGREGATION = 2,
       GP_PHASE_PARALLEL = 3, GP_PHASE_END = 4 };

/* DeviceCsrMatrix: rows x cols, row_ptr[rows+1], col_idx/vals[nnz]. */
typedef struct gp_csr {
  int32_t rows;
  int32_t cols;
  gp_offset nnz;
  gp_offset* row_ptr;
  int32_t* col_idx;
  double* vals;
} gp_csr;

/* LPInfoGpu: min c'x + obj_constant  s.t.  AL <= A x <= AU,  l <= x <= u.
 * AT is A transposed, stored as CSR (rows = columns of A). */
typedef struct gp_device_lp {
  gp_csr A;
  gp_csr AT;
  double* c;
  double* AL;
  double* AU;
  double* l;
  double* u;
  double obj_constant;
} gp_device_lp;

/* PresolveParams with the reference's field names and defaults.  Booleans are
 * int32 (0/1).  Fields of rules this engine does not build (the
 * structure-specific Startup rules) are accepted and ignored. */
typedef struct gp_params {
  int32_t struct_size; /* sizeof(gp_params); set by gp_default_params */
  int32_t max_iters;
  double max_time;
  int32_t verbose;
  int32_t record_postsolve_tape;
  int32_t record_postsolve_tape_cpu;
  int32_t enable_presolve;
  int32_t enable_folding;
  double folding_tolerance;
  double feasibility_tol;
  double bound_tol;
  double zero_tol;
  double bound_strengthening_upper_relaxation_factor;
  int32_t bound_strengthening_max_rounds;
  int32_t bound_strengthening_max_bound_only_rounds;
  int32_t bound_strengthening_preserve_free_zero_cost;
  int32_t doubleton_equations_max_fill_in_proxy;
  int32_t doubleton_equations_scan;
  int32_t doubleton_equations_min_selected_per_batch;
  double doubleton_equations_min_selected_ratio;
  int32_t doubleton_equations_max_batch_rounds;
  double doubleton_equations_max_time;
  int32_t enable_infeasible_and_fixed_variables;
  int32_t enable_empty_rows;
  int32_t enable_singleton_rows;
  int32_t enable_infeasible_and_redundant_rows;
  int32_t enable_bound_strengthening;
  int32_t enable_parallel_rows;
  int32_t enable_empty_columns;
  int32_t enable_singleton_columns;
  int32_t enable_doubleton_equations;
  int32_t enable_dual_fix;
  int32_t dual_fix_max_nonzero_cost_col_degree;
  int32_t enable_parallel_columns;
  double parallel_columns_max_abs_ratio;
  int32_t parallel_columns_scale_guard_max_group_size;
  int32_t parallel_columns_scale_guard_min_groups;
  int32_t enable_structure_specific_rules;
  /* Startup-rule settings (presolve_structs.hpp:103-146): accepted, not built;
   * gp_summary.unsupported_stages reports a request for these rules. */
  int32_t enable_two_row_variable_elimination;
  int32_t two_row_variable_elimination_max_row_nnz;
  double two_row_variable_elimination_safety_tol;
  int32_t two_row_variable_elimination_min_candidates;
  double two_row_variable_elimination_min_candidate_ratio;
  int32_t enable_zero_cost_product_variables;
  int32_t zero_cost_product_variables_min_candidates;
  double zero_cost_product_variables_min_candidate_ratio;
  int32_t enable_piecewise_linear_reformulation_variable_pair;
  int32_t piecewise_linear_reformulation_variable_pair_epigraph_probe_pairs;
  int32_t piecewise_linear_reformulation_variable_pair_epigraph_min_pairs;
  double piecewise_linear_reformulation_variable_pair_epigraph_min_potential_gain;
  double piecewise_linear_reformulation_variable_pair_epigraph_min_actual_gain;
  int32_t piecewise_linear_reformulation_variable_pair_epigraph_max_tail_cols;
  int32_t piecewise_linear_reformulation_variable_pair_epigraph_max_quotient_terms_per_row;
  int32_t piecewise_linear_reformulation_variable_pair_shifted_antipodal_probe_pairs;
  int32_t piecewise_linear_reformulation_variable_pair_shifted_antipodal_min_pairs;
  double piecewise_linear_reformulation_variable_pair_shifted_antipodal_min_potential_gain;
  double piecewise_linear_reformulation_variable_pair_shifted_antipodal_min_actual_gain;
  int32_t piecewise_linear_reformulation_variable_pair_shifted_antipodal_max_quotient_terms_per_row;
  int32_t enable_column_domination;
  int32_t column_domination_probe_rows;
  int32_t column_domination_probe_cols;
  int32_t column_domination_probe_nnz;
  int32_t column_domination_min_fixed_cols;
  double column_domination_min_candidate_ratio;
  int32_t enable_piecewise_linear_reformulation_absolute_value;
  int32_t piecewise_linear_reformulation_absolute_value_residual_bound_mode;
  double piecewise_linear_reformulation_absolute_value_residual_bound_as_free_min;
  int32_t enable_redundant_variable_bounds;
  int32_t enable_startup;
  int64_t bound_strengthening_bound_only_nnz_round_budget;
  /* Engine extensions (no reference counterpart). */
  int32_t collect_trace;     /* record one gp_round_trace per scheduler step    */
  int32_t learned_schedule;  /* 1: gate rule groups with the learned pass scheduler */
  const float* scheduler_weights; /* host blob of GP_MLP_BLOB floats (learned_schedule=1) */
  int32_t kernel_timing;     /* 1: time every kernel family with HIP events     */
  int32_t incremental_live_stats; /* 1 (default): update live statistics from the
                                     deleted lines instead of re-sweeping A and AT
                                     (identical results; the reference's runtime
                                     option of the same name, runtime_options.hpp) */
} gp_params;

/* One scheduler execution step (the reference prints these as
 * ">>> [schedule-progress]" lines, gpu_presolve.cu:2500-2592). */
typedef struct gp_round_trace {
  int32_t phase;
  int32_t next_phase;
  int64_t nnz_before;
  int64_t nnz_after;
  int32_t rows_live;
  int32_t cols_live;
  int32_t fired_mask;    /* bit r set when rule r changed the plan */
  int32_t skipped_mask;  /* bit r set when the learned scheduler skipped rule r */
  double seconds;
  int32_t bound_only_rounds; /* bound-only propagation rounds in a row, at round start */
  int32_t bound_only_limit;  /* the run's bound-only round limit                      */
  int32_t final_sweep;       /* 1: round of the plain fixpoint sweep after a learned phase */
  int32_t reserved;
} gp_round_trace;

/* Per-kernel-family device time accumulated when params.kernel_timing=1. */
enum {
  GP_KT_LIVE_STATS = 0,       /* full live-statistics sweep of A and AT        */
  GP_KT_ACTIVITY = 1,         /* row activity summary (row fold)               */
  GP_KT_BOUND_CANDIDATES = 2, /* bound-strengthening candidates                */
  GP_KT_COMPACT = 3,          /* CSR compaction at materialize                 */
  GP_KT_TRANSPOSE = 4,        /* CSR transpose                                 */
  GP_KT_RULES_OTHER = 5,      /* rule work outside the families below          */
  GP_KT_HASH_SORT = 6,        /* parallel rows / columns hashing and sorts     */
  GP_KT_SC_ACTIVITY = 7,      /* singleton-column row activity cache (row fold)*/
  GP_KT_REDUNDANT = 8,        /* redundant-row activity test (row fold)        */
  GP_KT_COL_MAX = 9,          /* column max |a|                                */
  GP_KT_LIVE_STATS_INC = 10,  /* live statistics from deleted lines            */
  GP_KT_CLEANUP = 11,         /* fixed variables, empty rows / columns, singleton rows */
  GP_KT_SINGLETON_COLUMNS = 12, /* singleton columns outside the activity cache */
  GP_KT_BOUNDS = 13,          /* bound strengthening / redundant rows outside the folds */
  GP_KT_DUAL_FIX = 14,        /* dual fix                                      */
  GP_KT_DOUBLETON = 15,       /* doubleton equations                           */
  GP_KT_PARALLEL = 16,        /* parallel rows / columns outside the hashing   */
  GP_KT_VECTORS = 17,         /* vector compaction at materialize              */
  GP_KT_ENTRY_BOUNDS = 18,    /* entry-bounds store: changed columns, carry through materialize */
  GP_KT_COUNT = 19
};

/* ---- learned pass scheduler (no reference counterpart) --------------------
 * An MLP 32-64-64-7 in FP32 (ReLU hidden layers, FP32 accumulation) reads the
 * per-round features of gp_schedule_features and scores the 7 rule groups
 * below; a group runs in a round only when its logit reaches the group's
 * threshold.  When the gated loop stops, the plain schedule runs on from that
 * state as the fixpoint check and fallback.  Parameter blob (FP32, row-major
 * out x in): W1[64][32] b1[64] W2[64][64] b2[64] W3[7][64] b3[7], then the 7
 * logit thresholds. */
#define GP_MLP_IN 32
#define GP_MLP_H1 64
#define GP_MLP_H2 64
#define GP_MLP_OUT 7
#define GP_MLP_PARAMS (GP_MLP_H1 * GP_MLP_IN + GP_MLP_H1 + GP_MLP_H2 * GP_MLP_H1 + GP_MLP_H2 + \
                       GP_MLP_OUT * GP_MLP_H2 + GP_MLP_OUT)
#define GP_MLP_BLOB (GP_MLP_PARAMS + GP_MLP_OUT)

/* Rule groups scored by the network. */
enum {
  GP_GROUP_FIXED_EMPTY = 0,      /* fixed variables, empty columns, empty rows */
  GP_GROUP_SINGLETON_ROWS = 1,
  GP_GROUP_SINGLETON_COLUMNS = 2,
  GP_GROUP_BOUNDS_REDUNDANT = 3, /* bound strengthening, redundant rows        */
  GP_GROUP_DUAL_FIX = 4,
  GP_GROUP_DOUBLETON = 5,
  GP_GROUP_PARALLEL = 6          /* parallel rows, parallel columns            */
};

/* Stages of the reference's preprocess() this engine does not build.  (Bit 1,
 * folding, is built: it is never set.) */
enum {
  GP_STAGE_FOLDING = 1,
  GP_STAGE_STRUCTURE_SPECIFIC_RULES = 2  /* enable_structure_specific_rules with one of
                                            its sub-rules on (gpu_presolve.cu:2361-2454) */
};

/* FoldingPipelineSummary + GpuPresolveSummary, flattened.  Value semantics:
 * gp_preprocess fills a caller-provided struct; gp_free_summary releases its
 * device memory and zeroes it (idempotent). */
/* Folding (FoldingRunSummary, cpp/include/gpu_preprocessor/folding/folding_structs.hpp:
 * 64-92; run_folding, cpp/src/folding/folding.cu:69-136): the LP folded by
 * colour refinement -- one row per row class and one column per column
 * class -- and the map back.  Arrays are device memory owned by the result
 * (gp_free_fold) or, inside a gp_summary, by the summary. */
typedef struct gp_fold_result {
  int32_t applied;
  int32_t original_rows;
  int32_t original_cols;
  int32_t folded_rows;
  int32_t folded_cols;
  int32_t rounds;           /* refinement rounds run */
  int32_t refinement_ok;    /* 0: stopped because both sides kept > 80% of their lines */
  int32_t final_row_colors;
  int32_t final_col_colors;
  int32_t reserved;
  double seconds;
  gp_device_lp folded_lp;   /* valid when applied */
  int32_t* row_color_id;    /* [original_rows] 1-based class of each row (when applied) */
  int32_t* col_color_id;    /* [original_cols] */
  double* row_scale;        /* [original_rows] 1 / class size */
  double* col_scale;        /* [original_cols] */
  void* state;              /* owner handle (NULL inside a gp_summary) */
} gp_fold_result;

typedef struct gp_summary {
  double preprocessing_seconds;
  double workspace_setup_seconds;
  double workspace_release_seconds;
  int32_t presolve_enabled;
  int32_t folding_applied;
  int32_t original_rows;
  int32_t original_cols;
  int32_t reduced_rows;
  int32_t reduced_cols;
  gp_offset reduced_nnz;
  int32_t iterations;
  int32_t stopped_at_iteration_limit;
  int32_t has_change;
  double elapsed_seconds;
  double obj_constant_delta;
  int32_t has_infeasible;
  int32_t has_unbounded;
  gp_device_lp reduced_lp;
  int32_t owns_reduced_lp;
  void* record; /* opaque recovery record (device resident) */
  int32_t trace_count;
  gp_round_trace* trace; /* host array, trace_count entries */
  double kernel_ms[GP_KT_COUNT];
  int64_t kernel_launches[GP_KT_COUNT];
  double kernel_bytes[GP_KT_COUNT]; /* algorithmic bytes per family */
  int32_t unsupported_stages; /* GP_STAGE_* bits: stages the parameters requested
                                 that this engine does not run (the reduction
                                 is then the one with those stages off) */
  gp_fold_result folding;     /* the folding stage (enable_folding); presolve ran
                                 on folding.folded_lp when folding.applied */
  int32_t primal_only_reductions; /* GP_PRIMAL_ONLY_* bits: the record's
                                 has_primal_only_*_reduction flags
                                 (presolve_structs.hpp:362-365) */
} gp_summary;

/* Reductions whose postsolve restores the primal point only. */
enum {
  GP_PRIMAL_ONLY_COLUMN_DOMINATION = 1,
  GP_PRIMAL_ONLY_PROJECTED_AUXILIARY = 2,
  GP_PRIMAL_ONLY_VARIABLE_PAIR = 4
};

/* Batched presolve (no reference counterpart; the reference presolves one LP
 * per call).  The instances are presolved together as their block-diagonal
 * direct sum: every round sweeps all instances in one launch per kernel.
 * Instance k occupies rows [row_offsets[k], row_offsets[k+1]) and columns
 * [col_offsets[k], col_offsets[k+1]) of the direct sum, and the same-numbered
 * reduced_* ranges of whole.reduced_lp.  Scheduling decisions are taken for
 * the batch.  An instance found infeasible or unbounded is retired (status[],
 * restarts); whole.has_infeasible / has_unbounded flag the batch only when a
 * stage that marks no instance ends it. */
typedef struct gp_batch_summary {
  int32_t count;
  gp_summary whole;            /* presolve of the direct sum (record, reduced LP) */
  gp_device_lp assembled;      /* the direct sum (library-owned device memory)    */
  int32_t* row_offsets;        /* host [count+1] */
  int32_t* col_offsets;        /* host [count+1] */
  int32_t* reduced_row_offsets;  /* host [count+1] */
  int32_t* reduced_col_offsets;  /* host [count+1] */
  gp_offset* reduced_nnz_offsets; /* host [count+1]: A entries of each reduced instance */
  double* obj_constant;        /* host [count]: objective constant of each reduced instance */
  double assemble_seconds;
  double presolve_seconds;
  double total_seconds;
  int32_t* status;             /* host [count]: 0 presolved in the batch, 1 infeasible, 2 unbounded */
  int32_t restarts;            /* times the batch started over without retired instances */
} gp_batch_summary;

/* Host view of a recovery record. */
typedef struct gp_record_sizes {
  int32_t m0, n0, m1, n1;
  int32_t fixed_count;
  int64_t record_count;
  int64_t index_count;
  int64_t value_count;
} gp_record_sizes;

/* ---- public pipeline ---------------------------------------------------- */

const char* gp_last_error(void);
int gp_abi_version(void);
void gp_default_params(gp_params* params);

/* preprocess(): presolve the device LP `input` (caller-owned, unmodified).
 * On a run that changes nothing the reduced LP borrows the input arrays. */
int gp_preprocess(const gp_device_lp* input, const gp_params* params, gp_summary* out);

/* postprocess(): x, y, z are device arrays in reduced order (n1, m1, n1);
 * x_out (n0), y_out (m0), z_out (n0) are host arrays in original order.
 * Convention c - A'y - z = 0.  `original` may be NULL (no dual projection). */
int gp_postprocess(const double* x, const double* y, const double* z,
                   const gp_summary* summary, const gp_device_lp* original,
                   double* x_out, double* y_out, double* z_out);

/* free_summary(): safe on NULL and on repeated calls. */
void gp_free_summary(gp_summary* summary);

/* Recovery record export (tests, diagnostics).  Any output pointer may be NULL. */
int gp_record_get_sizes(const gp_summary* summary, gp_record_sizes* sizes);
int gp_record_copy(const gp_summary* summary,
                   int32_t* row_red2org, int32_t* col_red2org,
                   int32_t* fixed_idx, double* fixed_val,
                   int32_t* types, gp_offset* index_starts, gp_offset* value_starts,
                   uint8_t* dual_modes, int32_t* indices, double* vals);

/* ---- postsolve of a caller-built recovery record -------------------------
 * PresolveRecordGpu (presolve_structs.hpp:338-368) as plain host arrays, for
 * postsolve_gpu(x_red, y_red, z_red, record, original_model, original_l)
 * (gpu_postsolve.cu:504-670, declared in gpu_postsolve.hpp).  The primal-only
 * structural recoveries of the reference's Startup rules are carried by the
 * same chronological timeline: a checkpoint names a payload and the number of
 * tape records that preceded it. */

/* PiecewiseLinearReformulationAbsoluteValuePrimalRecoveryStep
 * (presolve_structs.hpp:207-236): splits x_t <- x_t + x_e,
 * x_e <- (x_t - x_e) / rho; outer pairs x_free <- x_bound; linked slacks
 * x_slack <- factor * x_t; max slacks x_slack <- max(l_slack or 0,
 * max_k factors[k] * x_{t_cols[k]}) over max_t_cols[max_starts[i] ..
 * max_starts[i+1]). */
typedef struct gp_abs_value_step {
  int32_t split_count;
  const int32_t* split_t_col;
  const int32_t* split_e_col;
  const double* split_rho;
  int32_t pair_count;
  const int32_t* pair_bound_col;
  const int32_t* pair_free_col;
  int32_t linked_count;
  const int32_t* linked_slack_col;
  const int32_t* linked_t_col;
  const double* linked_factor;
  int32_t max_count;
  const int32_t* max_slack_col;
  const gp_offset* max_starts; /* max_count + 1 */
  const int32_t* max_t_cols;
  const double* max_factors;
} gp_abs_value_step;

/* PiecewiseLinearReformulationVariablePairRecoveryStep
 * (presolve_structs.hpp:238-250): x_plus <- plus_shift + x_root_plus,
 * x_minus <- minus_shift + x_root_minus, roots read before any write. */
typedef struct gp_var_pair_step {
  int32_t count;
  const int32_t* plus_cols;
  const int32_t* minus_cols;
  const int32_t* root_plus_cols;
  const int32_t* root_minus_cols;
  const double* plus_shift;
  const double* minus_shift;
} gp_var_pair_step;

/* PrimalRecoveryKind / PrimalRecoveryCheckpoint (presolve_structs.hpp:252-266). */
enum { GP_RECOVERY_ABS_VALUE = 0, GP_RECOVERY_VAR_PAIR = 1 };
typedef struct gp_checkpoint {
  int32_t kind;
  gp_offset payload_index;
  gp_offset tape_position;
} gp_checkpoint;

typedef struct gp_host_record {
  int32_t m0, n0, m1, n1;
  const int32_t* row_red2org; /* m1 */
  const int32_t* col_red2org; /* n1 */
  int32_t fixed_count;
  const int32_t* fixed_idx;
  const double* fixed_val;
  gp_offset record_count;
  const int32_t* types;
  const gp_offset* index_starts; /* record_count + 1 (may be NULL when empty) */
  const gp_offset* value_starts; /* record_count + 1 (may be NULL when empty) */
  const uint8_t* dual_modes;
  gp_offset index_count;
  const int32_t* indices;
  gp_offset value_count;
  const double* vals;
  int32_t abs_value_count;
  const gp_abs_value_step* abs_value;
  int32_t var_pair_count;
  const gp_var_pair_step* var_pair;
  int32_t checkpoint_count;
  const gp_checkpoint* checkpoints;
} gp_host_record;

/* postsolve_gpu() over a host record: validates the tape and the timeline
 * (validate_postsolve_tape / validate_primal_recovery_timeline,
 * gpu_postsolve.cu:400-503; GP_ERR_RUNTIME when malformed), scatters the
 * reduced vectors, restores fixed columns, replays primal ranges between
 * checkpoints in reverse with their structural payloads, then the dual pass,
 * then projects column duals from `original` when it carries AT, c, l and u
 * and the LP has rows.  x_red, y_red, z_red are device arrays in reduced
 * order and may be NULL; `original` (device LP, fields may be NULL) and
 * `original_l` (device, n0; defaults to original->l) may be NULL.  Outputs
 * are host arrays x_out (n0), y_out (m0), z_out (n0). */
int gp_postsolve_record(const double* x_red, const double* y_red, const double* z_red,
                        const gp_host_record* record, const gp_device_lp* original,
                        const double* original_l, double* x_out, double* y_out, double* z_out);

/* ---- launch-schedule autotuner (no reference counterpart) ---------------- */

/* Result-invariant launch schedule: lanes per row of the grouped sweeps
 * (0, 2, 4, 8, 16 or 64; 64 / lanes rows per wavefront), entries per lane of
 * the compaction tiles (0, 4, 8 or 16; 1,024 lanes per tile) and rows per
 * workgroup of the row folds (0, 128, 256 or 512).  0 = derived from the
 * matrix's mean row length; for lanes_col_max, 0 = entry-parallel tiles (no
 * lanes per column); for compact_per_lane, 0 = 8.  Process-wide. */
typedef struct gp_tuning {
  int32_t lanes_candidates;
  int32_t lanes_col_max;
  int32_t compact_per_lane;
  int32_t lanes_live_stats;
  int32_t fold_rows;
} gp_tuning;

int gp_get_tuning(gp_tuning* tuning);
int gp_set_tuning(const gp_tuning* tuning);
/* Measure each candidate schedule on `lp` (best of `reps` presolves each) and
 * return the fastest per knob in *best -- or all zeros when the tuned
 * schedule is not faster overall.  The current schedule is left unchanged.
 * tuned_ms / default_ms (may be NULL): presolve core ms with either. */
int gp_autotune(const gp_device_lp* lp, const gp_params* params, int32_t reps, gp_tuning* best,
                double* tuned_ms, double* default_ms);

/* ---- batched presolve (no reference counterpart) ------------------------- */

/* Presolve `count` device LPs (caller-owned, unmodified) as one batch.  An
 * instance that a rule finds infeasible or unbounded retires alone: the batch
 * starts over without it, it becomes an empty block of the direct sum (the
 * offsets describe the direct sum actually presolved) and status[k] says why. */
int gp_preprocess_batch(const gp_device_lp* instances, int32_t count, const gp_params* params,
                        gp_batch_summary* out);
/* Postsolve of the whole batch: x, y, z are device arrays over the reduced
 * direct sum; outputs are host arrays over the original direct sum. */
int gp_postprocess_batch(const double* x, const double* y, const double* z,
                         const gp_batch_summary* summary, double* x_out, double* y_out, double* z_out);
/* Reduced instance k as its own device LP (library-allocated; gp_lp_free). */
int gp_batch_instance(const gp_batch_summary* summary, int32_t k, gp_device_lp* out);
/* Safe on NULL and on repeated calls. */
void gp_free_batch_summary(gp_batch_summary* summary);

/* ---- rule-level operator API (PresolvePlanGpu / PresolveStatsGpu) -------- */

/* Deferred plan: masks and replacement vectors over the physical LP. */
typedef struct gp_plan {
  uint8_t* keep_row_mask;
  uint8_t* keep_col_mask;
  double* new_c;
  double* new_l;
  double* new_u;
  double* new_AL;
  double* new_AU;
  const uint8_t* rule_entry_keep_col;
  const double* propagation_reference_l;
  const double* propagation_reference_u;
  double obj_constant_delta;
  int32_t has_change;
  int32_t has_row_action;
  int32_t has_col_action;
  int32_t has_redundant_row_deletion;
  int32_t has_row_deletion;
  int32_t has_infeasible;
  int32_t has_unbounded;
  int32_t has_new_A;
  gp_csr new_A;   /* replacement matrix built by doubleton substitution */
  void* tape;     /* opaque device tape (local indices); NULL = do not record */
  void* rule_cache; /* opaque per-plan caches of the rules (library-owned; NULL at creation) */
} gp_plan;

typedef struct gp_stats {
  int32_t live_masks;
  int64_t live_nnz;
  int32_t* row_nnz;
  uint8_t* empty_row_mask;
  uint8_t* singleton_row_mask;
  int32_t* singleton_row_col;
  double* singleton_row_val;
  int32_t* col_nnz;
  uint8_t* empty_col_mask;
  uint8_t* singleton_column_mask;
  int32_t* singleton_column_row;
  double* singleton_column_val;
} gp_stats;

/* allocate_plan(): masks = 1, new_* = copies of lp vectors (device). */
int gp_plan_create(const gp_device_lp* lp, gp_plan* plan);
void gp_plan_destroy(gp_plan* plan);
int gp_stats_create(int32_t m, int32_t n, gp_stats* stats);
void gp_stats_destroy(gp_stats* stats);
int gp_refresh_live_stats(gp_stats* stats, const gp_device_lp* lp, const gp_plan* plan);

/* Apply one general rule (GP_RULE_*) to the plan. */
/* update_live_stats_from_deletions (gpu_presolve_kernels.cu:520-682): bring
 * statistics computed for the masks (old_keep_row, old_keep_col) to the
 * plan's current masks by visiting only the rows and columns deleted since;
 * the result equals gp_refresh_live_stats bit for bit.  *accepted = 0 (the
 * statistics then need a full refresh, live_nnz is left as it was) when a row
 * or column came back to life.  stats->live_nnz must hold the old total. */
int gp_update_live_stats(gp_stats* stats, const gp_device_lp* lp, const gp_plan* plan,
                         const uint8_t* old_keep_row, const uint8_t* old_keep_col, int32_t* accepted);
int gp_rule_apply(int32_t rule, gp_plan* plan, const gp_device_lp* lp,
                  const gp_stats* stats, const gp_params* params);

/* Structural Startup rules that share one deletion-only plan, in the order
 * Startup runs them (gpu_presolve.cu:2361-2410):
 *   GP_STARTUP_COLUMN_DOMINATION   analyze_column_domination +
 *     build_column_domination_plan (rule_column_domination.cu:244-377):
 *     on a unit set-covering LP (rows 1 <= a x, entries 1, [0, 1] columns
 *     with positive integer costs, a cost-one column in every row), columns
 *     with c_j > nnz(A[:, j]) are fixed to 0 (keep_col 0, new_l = new_u = 0);
 *   GP_STARTUP_TWO_ROW_ELIMINATION analyze_ / build_two_row_variable_elimination
 *     (rule_two_row_variable_elimination.cu:520-641): a zero-cost bounded
 *     column in exactly two short one-sided rows whose envelopes certify it
 *     is projected out with both rows (FmeCol record);
 *   GP_STARTUP_ZERO_COST_PRODUCTS  analyze_ / build_zero_cost_product_variables
 *     (rule_zero_cost_product_variables.cu:436-567): a zero-cost [0, 1]
 *     column z in exactly the rows z >= x + y - 1, z <= x, z <= y of binary
 *     boxes x, y is projected out with them (FmeCol record).
 * With pending = 1 the analysis reads the plan's masks and vectors (the
 * `pending` plan of the reference); with pending = 0 it reads `lp` alone.
 * When the rule is applicable and `plan` is given (device plan), the records
 * go to the plan's tape (if any) and the deletions into its masks.
 * *candidate_count and *applicable receive the analysis result (either may
 * be NULL).  The rules run only when their enable_* field is set. */
enum {
  GP_STARTUP_COLUMN_DOMINATION = 0,
  GP_STARTUP_TWO_ROW_ELIMINATION = 1,
  GP_STARTUP_ZERO_COST_PRODUCTS = 2
};
int gp_startup_rule(int32_t rule, const gp_device_lp* lp, const gp_params* params, gp_plan* plan,
                    int32_t pending, int32_t* candidate_count, int32_t* applicable);
/* Export the plan tape (local indices). */
/* append_compact_fixed_variable_tape_from_device (fixed_variable_tape.cu:141-327):
 * FixedVariable records (idx [col, kept rows...], vals [v, c, coefficients...],
 * Minimal dual mode) of every column with fixed[j] != 0 and, when `excluded`
 * is given, excluded[j] == 0, in column order, appended to the plan's tape
 * (no records when `plan` is NULL or has no tape).  With compute_objective
 * the objective delta sum c_j * v_j is formed as the reference forms it: a
 * sequential sum in x86 extended precision, rounded to binary64 once.
 * Device arrays: fixed, excluded (n or NULL), fixed_val, c (n), keep_row
 * (AT->cols); AT is the transpose with AT->rows == n. */
int gp_fixed_variable_tape(gp_plan* plan, const uint8_t* fixed, const uint8_t* excluded,
                           const double* fixed_val, const uint8_t* keep_row, const double* c,
                           const gp_csr* AT, int32_t n, int32_t compute_objective, double* objective);
int gp_plan_tape_sizes(const gp_plan* plan, int64_t* records, int64_t* indices, int64_t* values);
int gp_plan_tape_copy(const gp_plan* plan, int32_t* types, gp_offset* index_starts,
                      gp_offset* value_starts, uint8_t* dual_modes, int32_t* indices,
                      double* vals);
int gp_plan_enable_tape(gp_plan* plan);

/* Row activity summary (compute_row_activity_summary): finite min/max sums
 * (compensated, fixed order) and counts of infinite contributions. */
/* gp_row_activity_summary without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * The sweep's scratch (interleaved column bounds and packed column-mask bits)
 * is a fixed device array of the library, 32.25 MiB per device, so A->cols is
 * limited to 2097152 (GP_ERR_INVALID_ARGUMENT beyond).  Calls that overlap on
 * one device share that array and race.  Same arguments and results as
 * gp_row_activity_summary. */
int gp_row_activity_summary_noalloc(const gp_csr* A, const double* l, const double* u, double zero_tol,
                                    const uint8_t* keep_row, const uint8_t* keep_col, double* min_fin,
                                    double* max_fin, int32_t* min_ninf, int32_t* max_pinf);

/* Stable sort of (key, value) pairs by key bits [0, end_bit), in place on the
 * device; the engine's replacement for the reference's library sorts
 * (cub::DeviceRadixSort in postsolve/parallel_replay.cuh, thrust::sort_by_key
 * in rule_doubleton_equations.cu:126-148 and rule_parallel_columns.cu:746-766,
 * the thrust sorts of rule_parallel_rows.cu:643-888). */
int gp_radix_sort_pairs(uint64_t* keys, int32_t* vals, int64_t n, int32_t end_bit);

/* compute_parallel_row_hashes (rule_parallel_rows.cu:141-196,583-627): per row
 * the FNV-1a chain over (column, bits of a / first live a) of its live
 * entries, XOR-aggregated per entry when nnz > 32 rows; `row` for rows not
 * kept, mix(offset, row) for rows without live entries or with a pivot at or
 * below zero_tol.  Device arrays; keep_col may be NULL (every column live). */
int gp_parallel_row_hashes(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, double zero_tol,
                           uint64_t* hash);
int gp_row_activity_summary(const gp_csr* A, const double* l, const double* u, double zero_tol,
                            const uint8_t* keep_row, const uint8_t* keep_col,
                            double* min_fin, double* max_fin, int32_t* min_ninf, int32_t* max_pinf);

/* ---- learned pass scheduler ----------------------------------------------- */

/* Group of a rule (GP_RULE_* -> GP_GROUP_*). */
int gp_rule_group(int32_t rule);
/* Features of the round about to run (host memory, no device call).
 * history[0..rounds_done) are the rounds already run (trace records), phase the
 * phase of the next round, nnz_now its live nnz, nnz0/m0/n0 the input's size,
 * bound_only_rounds/limit the scheduler's bound-only counter and limit. */
void gp_schedule_features(const gp_round_trace* history, int32_t rounds_done, int32_t phase,
                          int64_t nnz_now, int64_t nnz0, int32_t m0, int32_t n0,
                          int32_t bound_only_rounds, int32_t bound_only_limit, float* features);
/* Training and inference kernels; every pointer is device memory.
 * forward: logits[batch][7] of x[batch][32]. */
int gp_mlp_forward(const float* params, const float* x, int32_t batch, float* logits);
/* Masked binary cross-entropy with logits, mean over the mask's ones:
 * loss[1] and grad[GP_MLP_PARAMS] of params at (x, y, mask), each [batch][7]
 * for y and mask. */
int gp_mlp_loss_grad(const float* params, const float* x, const float* y, const float* mask,
                     int32_t batch, float* loss, float* grad);
/* One Adam step (bias-corrected, step t >= 1) over GP_MLP_PARAMS values. */
int gp_mlp_adam_step(float* params, const float* grad, float* m, float* v, int32_t t, float lr,
                     float beta1, float beta2, float eps);

/* ---- matrix primitives --------------------------------------------------- */

/* AT = transpose(A); AT arrays are allocated by the library (gp_csr_free). */
int gp_transpose_csr(const gp_csr* A, gp_csr* AT);
/* Stable compaction by keep masks (NULL = keep all). */
int gp_compact_csr(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, gp_csr* out);
void gp_csr_free(gp_csr* matrix);

/* ---- device memory helpers (for harnesses without a GPU runtime) ---------- */
int gp_device_count(int32_t* count);
int gp_malloc(void** ptr, size_t bytes);
int gp_free(void* ptr);
int gp_memcpy_h2d(void* dst, const void* src, size_t bytes);
int gp_memcpy_d2h(void* dst, const void* src, size_t bytes);
int gp_memcpy_d2d(void* dst, const void* src, size_t bytes);
int gp_synchronize(void);
/* Upload a host CSR LP; AT is built on the device by gp_transpose_csr. */
int gp_lp_upload(int32_t m, int32_t n, gp_offset nnz, const gp_offset* row_ptr,
                 const int32_t* col_idx, const double* vals, const double* c,
                 const double* AL, const double* AU, const double* l, const double* u,
                 double obj_constant, gp_device_lp* out);
void gp_lp_free(gp_device_lp* lp);

/* ---- folding ----------------------------------------------------------------
 * run_folding (cpp/include/gpu_preprocessor/folding/folding.hpp:16-18) with
 * the folding tolerance; gp_unfold is unfold_solution_device_to_host
 * (folding.hpp:38-41): reduced x, y, z device arrays in, original-order host
 * arrays out (x copied, y and z scaled by 1 / class size); gp_free_fold is
 * free_folded_lp. */
int gp_fold(const gp_device_lp* lp, double tolerance, gp_fold_result* out);
int gp_unfold(const gp_fold_result* fold, const double* x_red, const double* y_red, const double* z_red,
              double* x, double* y, double* z);
void gp_free_fold(gp_fold_result* fold);

/* ---- MPS ingest -------------------------------------------------------------
 * An MPS file as mps_reader::read_mps reads it
 * (cpp/mps_reader/include/mps_reader/mpsreader.hpp:14-19, LpModel of
 * lp_model.hpp:12-81): rows x cols, costs, row sides AL/AU, column bounds
 * l/u and the objective constant, with the matrix as its entries in file
 * order (duplicates of one (row, column) not yet summed).  Host memory owned
 * by the library: gp_mps_free. */
typedef struct gp_mps_model {
  int32_t rows;
  int32_t cols;
  int64_t entries;
  int32_t* entry_row;
  int32_t* entry_col;
  double* entry_val;
  double* c;
  double* AL;
  double* AU;
  double* l;
  double* u;
  double obj_constant;
  char* name; /* NAME card, NUL-terminated ("" when absent) */
} gp_mps_model;

/* read_mps (mpsreader.hpp:19; cpp/mps_reader/src/io/mpsreader.cpp:686-709) on
 * an in-memory text: free format, then fixed format; GP_ERR_INVALID_ARGUMENT
 * with the reference's message (the free-format one) when both fail.  Host
 * only: no GPU call. */
int gp_mps_parse(const char* text, size_t length, gp_mps_model* out);
/* The same on a file (read_mps_file, cpp/tools/gpu_preprocessor_tool_common.hpp:143-154;
 * GP_ERR_RUNTIME "failed to open <path>" when it cannot be opened). */
int gp_mps_read(const char* path, gp_mps_model* out);
void gp_mps_free(gp_mps_model* model);
/* upload_lp (cpp/tools/gpu_preprocessor_tool_common.hpp:108-141): the model in
 * HBM as a device LP.  Duplicates are summed on the device in file order
 * (the reference's per-column stable sort and left-to-right sum,
 * mpsreader.cpp:625-648), A^T comes out with increasing row indices per
 * column and A with increasing column indices per row (CscMatrix::to_csr,
 * sparse_matrix.cpp:111-135), explicit zeros kept.  Free with gp_lp_free. */
int gp_mps_upload(const gp_mps_model* model, gp_device_lp* out);
/* The device assembly of gp_mps_upload for any caller holding matrix entries
 * (row, column, value) in its own order, duplicates allowed: same result as
 * building the CSC on the host and calling upload_lp.  GP_ERR_INVALID_ARGUMENT
 * when an entry lies outside m x n; GP_ERR_LENGTH beyond 2^31 - 1 entries. */
int gp_lp_upload_entries(int32_t m, int32_t n, int64_t entries, const int32_t* entry_row,
                         const int32_t* entry_col, const double* entry_val, const double* c,
                         const double* AL, const double* AU, const double* l, const double* u,
                         double obj_constant, gp_device_lp* out);

/* ---- FP64 first-order LP solver ---------------------------------------------
 * The solver the reference hands the reduced LP to: HPR-LP, the restarted
 * Halpern Peaceman-Rachford method (third_party/hprlp: HPRLP_main_solve /
 * solve, include/HPRLP.h; parameters and results of include/api/structs.h:
 * 256-291, 334-358), called from solver_interface/integration.cpp:44-119.
 * Tolerance, iteration and time limits keep HPR-LP's meaning and defaults
 * (stop_tol 1e-6, time_limit 1000 s); termination when the relative primal
 * residual, dual residual and duality gap are all <= tolerance. */
enum {
  GP_LP_OPTIMAL = 0,
  GP_LP_ITERATION_LIMIT = 1,
  GP_LP_TIME_LIMIT = 2,
  GP_LP_NUMERICAL_ERROR = 3,
  GP_LP_UNBOUNDED = 4,  /* only reported for an LP without matrix entries */
  GP_LP_INFEASIBLE = 5  /* only reported for an LP without matrix entries */
};

typedef struct gp_lp_solve_params {
  int32_t struct_size;      /* set by gp_lp_solve_default_params */
  int32_t max_iterations;   /* max_iter */
  double tolerance;         /* stop_tol: relative KKT tolerance */
  double time_limit;        /* seconds */
  int32_t check_every;      /* check_iter: iterations between termination / restart checks */
  int32_t ruiz_iterations;  /* use_Ruiz_scaling: equilibration passes (0 = off) */
  int32_t pock_chambolle;   /* use_Pock_Chambolle_scaling */
  int32_t bound_cost_scaling; /* use_bc_scaling */
  int32_t power_iterations; /* limit of the ||A||^2 power iteration */
  int32_t reserved;
  double power_tolerance;
  /* restarts: when the fixed-point residual falls to `restart_sufficient` of
   * its value at the last restart, or below `restart_necessary` of it while
   * rising, or after `restart_artificial` of all iterations so far */
  double restart_sufficient;
  double restart_necessary;
  double restart_artificial;
  /* at a restart sigma = (primal move / dual move) / ||A||, then, when the
   * relative primal and dual residuals differ by more than the factor
   * residual_balance_band, times clamp((primal / dual)^residual_balance,
   * 0.1, 10): a larger sigma (primal step) when the dual residual lags */
  double residual_balance;
  double residual_balance_band;
} gp_lp_solve_params;

typedef struct gp_lp_solve_info {
  int32_t status;           /* GP_LP_* */
  int32_t iterations;
  int32_t restarts;
  int32_t reserved;
  double primal_objective;  /* c'x + obj_constant */
  double dual_objective;
  double primal_residual;   /* ||Ax - proj_[AL,AU](Ax)|| / (1 + ||b||) */
  double dual_residual;     /* ||c - A'y - z|| / (1 + ||c||) */
  double gap;               /* |primal - dual| / (1 + |primal| + |dual|) */
  double seconds;
  double sigma;             /* final primal step */
  double lambda;            /* estimate of ||A||_2^2 of the scaled matrix */
  double bytes_per_iteration; /* algorithmic HBM bytes of one iteration */
} gp_lp_solve_info;

void gp_lp_solve_default_params(gp_lp_solve_params* params);
/* Solve the device LP; x (cols), y (rows), z (cols) are caller-owned device
 * arrays that receive the solution (any may be NULL).  Convention
 * c - A'y - z = 0, as postprocess expects. */
int gp_lp_solve(const gp_device_lp* lp, const gp_lp_solve_params* params, double* x, double* y, double* z,
                gp_lp_solve_info* info);

#ifdef __cplusplus
}
#endif

#endif /* GPUPRE_H */
