/*
 * oracle_sched.c -- CPU restatement of the presolve scheduler, deferred plans,
 * materialization and recovery record (gpu_presolve.cu:1953-2645,
 * device_recovery.cu:233-276,881-1077).  TEST INFRASTRUCTURE ONLY.
 */
#include "oracle_internal.h"

#include <stdio.h>

#include <limits.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct or_working {
  gp_device_lp lp;
  int owns;
} or_working;

typedef struct or_deferred {
  gp_plan plan;
  gp_stats stats;
  uint8_t* entry_cols;
  double* reference_bounds;
  int active, pending, stats_dirty, nnz_dirty;
} or_deferred;

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void record_fixed_push(or_result* r, int32_t idx, double val) {
  if (r->fixed_count + 1 > r->fixed_cap) {
    r->fixed_cap = 2 * r->fixed_cap + 16;
    r->fixed_idx = (int32_t*)or_xrealloc(r->fixed_idx, sizeof(int32_t) * (size_t)r->fixed_cap);
    r->fixed_val = (double*)or_xrealloc(r->fixed_val, sizeof(double) * (size_t)r->fixed_cap);
  }
  r->fixed_idx[r->fixed_count] = idx;
  r->fixed_val[r->fixed_count] = val;
  r->fixed_count++;
}

/* append_fixed (device_recovery.cu:233-276): columns dropped from `keep`
 * (and still present in `prior`, when given) whose bounds are fixed. */
static void append_fixed(or_result* r, const gp_plan* plan, const gp_device_lp* lp,
                         const uint8_t* keep, const uint8_t* prior) {
  const int32_t n = lp->A.cols;
  const double* l = plan->new_l ? plan->new_l : lp->l;
  const double* u = plan->new_u ? plan->new_u : lp->u;
  for (int32_t j = 0; j < n; ++j) {
    if ((!prior || prior[j]) && !keep[j] && isfinite(l[j]) && isfinite(u[j]) &&
        fabs(l[j] - u[j]) <= 1e-9)
      record_fixed_push(r, r->col_red2org[j], 0.5 * (l[j] + u[j]));
  }
}

static void flush_plan_tape(or_result* r, gp_plan* plan) {
  or_tape* t = (or_tape*)plan->tape;
  if (!t || t->count == 0) return;
  or_tape_append_global(&r->tape, t, r->row_red2org, r->m1, r->col_red2org, r->n1);
  or_tape_free(t);
  or_tape_init(t);
}

/* capture_cleanup_recovery (device_recovery.cu:1011-1027). */
static void capture_cleanup_recovery(or_result* r, gp_plan* plan, const gp_device_lp* lp,
                                     const uint8_t* entry_cols) {
  flush_plan_tape(r, plan);
  if (plan->has_col_action) append_fixed(r, plan, lp, plan->keep_col_mask, entry_cols);
}

static int plan_keeps_all(const gp_plan* plan, int32_t m, int32_t n) {
  for (int32_t i = 0; i < m; ++i)
    if (!plan->keep_row_mask[i]) return 0;
  for (int32_t j = 0; j < n; ++j)
    if (!plan->keep_col_mask[j]) return 0;
  return 1;
}

static double* compact_vec(const double* src, const uint8_t* keep, int32_t n_old, int32_t n_new) {
  double* out = (double*)or_xmalloc(sizeof(double) * (size_t)(n_new ? n_new : 1));
  int32_t k = 0;
  for (int32_t j = 0; j < n_old; ++j)
    if (keep[j]) out[k++] = src[j];
  return out;
}

/* apply_plan_to_working_lp (gpu_presolve.cu:1953-2072) with the device
 * recovery backend (update_device_recovery, device_recovery.cu:892-975). */
static void apply_plan(or_working* cur, gp_plan* plan, or_result* r, int captured) {
  const int32_t m = cur->lp.A.rows, n = cur->lp.A.cols;
  if (cur->owns && !plan->has_new_A && plan_keeps_all(plan, m, n)) {
    flush_plan_tape(r, plan);
    if (n) {
      memcpy(cur->lp.c, plan->new_c, sizeof(double) * (size_t)n);
      memcpy(cur->lp.l, plan->new_l, sizeof(double) * (size_t)n);
      memcpy(cur->lp.u, plan->new_u, sizeof(double) * (size_t)n);
    }
    if (m) {
      memcpy(cur->lp.AL, plan->new_AL, sizeof(double) * (size_t)m);
      memcpy(cur->lp.AU, plan->new_AU, sizeof(double) * (size_t)m);
    }
    cur->lp.obj_constant += plan->obj_constant_delta;
    return;
  }
  int32_t m_new = 0, n_new = 0;
  for (int32_t i = 0; i < m; ++i) m_new += plan->keep_row_mask[i] != 0;
  for (int32_t j = 0; j < n; ++j) n_new += plan->keep_col_mask[j] != 0;
  const gp_csr* src = plan->has_new_A ? &plan->new_A : &cur->lp.A;
  gp_device_lp next;
  memset(&next, 0, sizeof(next));
  or_compact(src, plan->keep_row_mask, plan->keep_col_mask, &next.A);
  or_transpose(&next.A, &next.AT);
  next.c = compact_vec(plan->new_c, plan->keep_col_mask, n, n_new);
  next.l = compact_vec(plan->new_l, plan->keep_col_mask, n, n_new);
  next.u = compact_vec(plan->new_u, plan->keep_col_mask, n, n_new);
  next.AL = compact_vec(plan->new_AL, plan->keep_row_mask, m, m_new);
  next.AU = compact_vec(plan->new_AU, plan->keep_row_mask, m, m_new);
  next.obj_constant = cur->lp.obj_constant + plan->obj_constant_delta;
  /* recovery: tape first (current maps), then fixed columns, then maps */
  flush_plan_tape(r, plan);
  if (!captured && n_new != n) append_fixed(r, plan, &cur->lp, plan->keep_col_mask, NULL);
  {
    int32_t k = 0;
    for (int32_t i = 0; i < m; ++i)
      if (plan->keep_row_mask[i]) r->row_red2org[k++] = r->row_red2org[i];
    k = 0;
    for (int32_t j = 0; j < n; ++j)
      if (plan->keep_col_mask[j]) r->col_red2org[k++] = r->col_red2org[j];
    r->m1 = m_new;
    r->n1 = n_new;
  }
  if (cur->owns) or_lp_free(&cur->lp);
  cur->lp = next;
  cur->owns = 1;
}

static void deferred_reset(or_deferred* d) {
  if (d->active) {
    or_plan_destroy(&d->plan);
    or_stats_destroy(&d->stats);
    free(d->entry_cols);
    free(d->reference_bounds);
  }
  memset(d, 0, sizeof(*d));
}

static void trace_push(or_result* r, const gp_round_trace* t) {
  if (r->trace_count + 1 > r->trace_cap) {
    r->trace_cap = 2 * r->trace_cap + 16;
    r->trace = (gp_round_trace*)or_xrealloc(r->trace, sizeof(gp_round_trace) * (size_t)r->trace_cap);
  }
  r->trace[r->trace_count++] = *t;
}

typedef struct or_sched {
  or_working cur;
  or_deferred d;
  or_result* r;
  const gp_params* p;
  int terminal;
  uint64_t deletion_serial;
  double start;
  int32_t fired_mask;
} or_sched;

static int time_exceeded(or_sched* s) {
  if (!(s->p->max_time > 0.0 && isfinite(s->p->max_time))) return 0;
  return now_seconds() - s->start >= s->p->max_time;
}

static void prepare_deferred(or_sched* s) {
  if (s->d.active) return;
  or_plan_create(&s->cur.lp, &s->d.plan);
  if (s->p->record_postsolve_tape) or_plan_enable_tape(&s->d.plan);
  or_stats_create(s->cur.lp.A.rows, s->cur.lp.A.cols, &s->d.stats);
  const int32_t n = s->cur.lp.A.cols;
  s->d.entry_cols = (uint8_t*)or_xcalloc((size_t)(n ? n : 1), 1);
  s->d.reference_bounds = (double*)or_xcalloc((size_t)(2 * n + 1), sizeof(double));
  s->d.plan.rule_entry_keep_col = s->d.entry_cols;
  s->d.active = s->d.stats_dirty = s->d.nnz_dirty = 1;
}

static void refresh_deferred_stats(or_sched* s) {
  if (s->d.stats_dirty) {
    or_refresh_live_stats(&s->d.stats, &s->cur.lp, &s->d.plan);
    s->d.stats_dirty = 0;
  }
}

static int64_t live_nnz(or_sched* s) {
  if (!s->d.active) return s->cur.lp.A.nnz;
  refresh_deferred_stats(s);
  s->d.nnz_dirty = 0;
  return s->d.stats.live_nnz;
}

/* deferred_rule (gpu_presolve.cu:2289-2338) */
static int deferred_rule(or_sched* s, int enabled, int32_t rule) {
  if (!enabled || s->terminal || time_exceeded(s)) return 0;
  prepare_deferred(s);
  refresh_deferred_stats(s);
  gp_plan* plan = &s->d.plan;
  const int32_t n = s->cur.lp.A.cols;
  if (n) memcpy(s->d.entry_cols, plan->keep_col_mask, (size_t)n);
  plan->has_change = plan->has_row_action = plan->has_col_action = 0;
  plan->has_redundant_row_deletion = plan->has_row_deletion = 0;
  plan->obj_constant_delta = 0.0;
  or_rule_apply(rule, plan, &s->cur.lp, &s->d.stats, s->p);
  if (getenv("GP_ORACLE_TRACE_RULES") && (plan->has_infeasible || plan->has_unbounded))
    fprintf(stderr, "[oracle] rule %d: infeasible=%d unbounded=%d\n", rule, plan->has_infeasible,
            plan->has_unbounded);
  s->r->has_infeasible = plan->has_infeasible;
  s->r->has_unbounded = plan->has_unbounded;
  s->terminal = plan->has_infeasible || plan->has_unbounded;
  const int changed = plan->has_change && !s->terminal;
  if (changed) {
    if (plan->has_row_deletion || plan->has_col_action) ++s->deletion_serial;
    s->cur.lp.obj_constant += plan->obj_constant_delta;
    capture_cleanup_recovery(s->r, plan, &s->cur.lp, s->d.entry_cols);
    s->d.pending = 1;
    if (plan->has_row_action || plan->has_col_action) s->d.stats_dirty = s->d.nnz_dirty = 1;
    s->fired_mask |= 1 << rule;
  }
  plan->obj_constant_delta = 0.0;
  return changed;
}

static void materialize(or_sched* s) {
  if (s->d.pending) {
    s->d.plan.has_change = 1;
    apply_plan(&s->cur, &s->d.plan, s->r, 1);
  }
  deferred_reset(&s->d);
}

/* run_materialized_rule (gpu_presolve.cu:2144-2174) */
static int materialized_rule(or_sched* s, int enabled, int32_t rule) {
  if (!enabled || s->terminal || time_exceeded(s)) return 0;
  gp_plan plan;
  gp_stats stats;
  memset(&stats, 0, sizeof(stats));
  or_plan_create(&s->cur.lp, &plan);
  if (s->p->record_postsolve_tape) or_plan_enable_tape(&plan);
  or_rule_apply(rule, &plan, &s->cur.lp, &stats, s->p);
  s->r->has_infeasible = plan.has_infeasible;
  s->r->has_unbounded = plan.has_unbounded;
  s->terminal = plan.has_infeasible || plan.has_unbounded;
  const int changed = !s->terminal && plan.has_change;
  if (changed) {
    apply_plan(&s->cur, &plan, s->r, 0);
    s->fired_mask |= 1 << rule;
  }
  or_plan_destroy(&plan);
  return changed;
}

/* rule enabled by the parameters and its group chosen for this round */
static int gate_on(int enabled, int32_t rule, uint32_t run, gp_round_trace* tr) {
  if (!enabled) return 0;
  if ((run >> or_rule_group(rule)) & 1u) return 1;
  tr->skipped_mask |= 1 << rule;
  return 0;
}

/* The Startup projection batch (gpu_presolve.cu:2361-2410): column
 * domination, two-row elimination and zero-cost products share one
 * deletion-only plan, allocated when the first of them applies; each
 * analysis after that sees the plan's masks.  The plan is applied once. */
static void startup_batch(or_sched* s) {
  static const int32_t order[3] = {GP_STARTUP_COLUMN_DOMINATION, GP_STARTUP_TWO_ROW_ELIMINATION,
                                   GP_STARTUP_ZERO_COST_PRODUCTS};
  const int enabled[3] = {s->p->enable_column_domination, s->p->enable_two_row_variable_elimination,
                          s->p->enable_zero_cost_product_variables};
  gp_plan plan;
  int allocated = 0, flags = 0;
  for (int k = 0; k < 3; ++k) {
    if (!enabled[k] || s->terminal || time_exceeded(s)) continue;
    int32_t count = 0;
    if (!or_startup_rule(order[k], &s->cur.lp, s->p, allocated ? &plan : NULL, allocated, &count)) continue;
    if (!allocated) {
      or_plan_create(&s->cur.lp, &plan);
      if (s->p->record_postsolve_tape) or_plan_enable_tape(&plan);
      allocated = 1;
      or_startup_rule(order[k], &s->cur.lp, s->p, &plan, 0, &count);
    }
    flags |= order[k] == GP_STARTUP_COLUMN_DOMINATION ? GP_PRIMAL_ONLY_COLUMN_DOMINATION
                                                       : GP_PRIMAL_ONLY_PROJECTED_AUXILIARY;
  }
  if (!allocated) return;
  if (plan.has_change) {
    apply_plan(&s->cur, &plan, s->r, 0);
    s->r->primal_only |= flags;
  }
  or_plan_destroy(&plan);
}

static int good_progress(int64_t before, int64_t after, int64_t minimum) {
  return before >= minimum && after <= before - minimum;
}

int or_presolve(const gp_device_lp* lp, const gp_params* p, or_result* out) {
  memset(out, 0, sizeof(*out));
  const int32_t m0 = lp->A.rows, n0 = lp->A.cols;
  out->m0 = out->m1 = m0;
  out->n0 = out->n1 = n0;
  out->row_red2org = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(m0 ? m0 : 1));
  out->col_red2org = (int32_t*)or_xmalloc(sizeof(int32_t) * (size_t)(n0 ? n0 : 1));
  for (int32_t i = 0; i < m0; ++i) out->row_red2org[i] = i;
  for (int32_t j = 0; j < n0; ++j) out->col_red2org[j] = j;
  or_tape_init(&out->tape);

  or_sched s;
  memset(&s, 0, sizeof(s));
  s.cur.lp = *lp; /* borrowed */
  s.cur.owns = 0;
  s.r = out;
  s.p = p;
  s.start = now_seconds();

  /* Startup (gpu_presolve.cu:2361-2454): the structural rules, then one
   * iteration counted. */
  if (p->enable_startup && p->max_iters > 0 && !time_exceeded(&s)) {
    if (p->enable_structure_specific_rules) startup_batch(&s);
    out->iterations++;
  }

  int phase = GP_PHASE_CLEANUP;
  const int64_t min_dec = 100;
  int bound_only_rounds = 0;
  const int64_t initial_live = p->max_iters > 0 && !s.terminal && !time_exceeded(&s) ? live_nnz(&s) : 0;
  const int configured = p->bound_strengthening_max_bound_only_rounds > 1
                             ? p->bound_strengthening_max_bound_only_rounds
                             : 1;
  int limit = configured;
  if (p->bound_strengthening_bound_only_nnz_round_budget > 0 && initial_live > 0) {
    int64_t q = p->bound_strengthening_bound_only_nnz_round_budget / initial_live;
    if (q < 1) q = 1;
    if (q < limit) limit = (int)q;
  }
  {
    const int64_t m = s.cur.lp.A.rows, n = s.cur.lp.A.cols;
    if (m >= 512 && m <= 2048 && n >= 128 * m && n <= 256 * m && initial_live >= 2048 * m &&
        initial_live <= 8192 * m) {
      const int two = configured < 2 ? configured : 2;
      if (two > limit) limit = two;
    }
  }

  /* Learned pass scheduler (no reference counterpart): the network scores the
   * rule groups before each round and groups below their threshold are
   * skipped; when the gated loop stops, the plain schedule runs on. */
  const int learned = p->learned_schedule != 0;
  if (learned && !p->scheduler_weights) return GP_ERR_INVALID_ARGUMENT;
  int gating = learned;
  gp_round_trace* hist = NULL;
  int32_t hist_count = 0, hist_cap = 0;
  const int32_t m_in = lp->A.rows, n_in = lp->A.cols;
  while (p->max_iters > 0 && !s.terminal && !time_exceeded(&s)) {
    if (phase != GP_PHASE_CLEANUP) {
      if (out->iterations >= p->max_iters) {
        out->stopped_at_iteration_limit = 1;
        break;
      }
      out->iterations++;
    }
    const double t0 = now_seconds();
    const int64_t before = live_nnz(&s);
    gp_round_trace tr;
    memset(&tr, 0, sizeof(tr));
    tr.phase = phase;
    tr.nnz_before = before;
    tr.bound_only_rounds = bound_only_rounds;
    tr.bound_only_limit = limit;
    tr.final_sweep = learned && !gating;
    uint32_t run = ~0u;
    if (gating) {
      float f[GP_MLP_IN], logits[GP_MLP_OUT];
      or_schedule_features(hist, hist_count, phase, before, initial_live, m_in, n_in, bound_only_rounds, limit, f);
      or_mlp_forward(p->scheduler_weights, f, 1, logits);
      run = 0;
      for (int g = 0; g < GP_MLP_OUT; ++g)
        if (logits[g] >= p->scheduler_weights[GP_MLP_PARAMS + g]) run |= 1u << g;
    }
#define ON(enabled, rule) gate_on(enabled, rule, run, &tr)
    s.fired_mask = 0;
    int64_t after = before;
    int stop = 0;
    if (phase == GP_PHASE_CLEANUP) {
      const uint64_t serial = s.deletion_serial;
      int changed = deferred_rule(&s, ON(p->enable_infeasible_and_fixed_variables, GP_RULE_FIXED_VARIABLES),
                                  GP_RULE_FIXED_VARIABLES);
      changed = deferred_rule(&s, ON(p->enable_empty_columns, GP_RULE_EMPTY_COLUMNS), GP_RULE_EMPTY_COLUMNS) || changed;
      changed = deferred_rule(&s, ON(p->enable_empty_rows, GP_RULE_EMPTY_ROWS), GP_RULE_EMPTY_ROWS) || changed;
      changed = deferred_rule(&s, ON(p->enable_singleton_rows, GP_RULE_SINGLETON_ROWS), GP_RULE_SINGLETON_ROWS) ||
                changed;
      changed = deferred_rule(&s, ON(p->enable_singleton_columns, GP_RULE_SINGLETON_COLUMNS),
                              GP_RULE_SINGLETON_COLUMNS) ||
                changed;
      if (s.deletion_serial != serial) bound_only_rounds = 0;
      after = live_nnz(&s);
      phase = changed ? GP_PHASE_CLEANUP : GP_PHASE_PROPAGATION;
    } else if (phase == GP_PHASE_PROPAGATION) {
      const uint64_t serial = s.deletion_serial;
      int bounds_changed = 0;
      if (ON(p->enable_bound_strengthening, GP_RULE_BOUND_STRENGTHENING) && !s.terminal && !time_exceeded(&s)) {
        prepare_deferred(&s);
        const int32_t n = s.cur.lp.A.cols;
        if (n) {
          memcpy(s.d.reference_bounds, s.d.plan.new_l, sizeof(double) * (size_t)n);
          memcpy(s.d.reference_bounds + n, s.d.plan.new_u, sizeof(double) * (size_t)n);
          s.d.plan.propagation_reference_l = s.d.reference_bounds;
          s.d.plan.propagation_reference_u = s.d.reference_bounds + n;
        }
        bounds_changed = deferred_rule(&s, 1, GP_RULE_BOUND_STRENGTHENING);
      }
      deferred_rule(&s, ON(p->enable_infeasible_and_redundant_rows, GP_RULE_REDUNDANT_ROWS), GP_RULE_REDUNDANT_ROWS);
      deferred_rule(&s, ON(p->enable_dual_fix, GP_RULE_DUAL_FIX), GP_RULE_DUAL_FIX);
      const int deleted = s.deletion_serial != serial;
      bound_only_rounds = deleted ? 0 : bound_only_rounds + 1;
      materialize(&s);
      after = live_nnz(&s);
      const int nnz_progress = good_progress(before, after, min_dec);
      const int bound_progress = bounds_changed && bound_only_rounds < limit;
      phase = nnz_progress || bound_progress ? GP_PHASE_CLEANUP : GP_PHASE_AGGREGATION;
      if (phase == GP_PHASE_AGGREGATION) bound_only_rounds = 0;
    } else if (phase == GP_PHASE_AGGREGATION) {
      materialized_rule(&s, ON(p->enable_doubleton_equations, GP_RULE_DOUBLETON_EQUATIONS),
                        GP_RULE_DOUBLETON_EQUATIONS);
      after = live_nnz(&s);
      phase = good_progress(before, after, min_dec) ? GP_PHASE_CLEANUP : GP_PHASE_PARALLEL;
    } else {
      int changed = deferred_rule(&s, ON(p->enable_parallel_rows, GP_RULE_PARALLEL_ROWS), GP_RULE_PARALLEL_ROWS);
      changed = deferred_rule(&s, ON(p->enable_parallel_columns, GP_RULE_PARALLEL_COLUMNS), GP_RULE_PARALLEL_COLUMNS) ||
                changed;
      materialize(&s);
      after = live_nnz(&s);
      const int cont = changed && good_progress(before, after, min_dec);
      if (!cont) stop = 1;
      else phase = GP_PHASE_CLEANUP;
    }
#undef ON
    tr.next_phase = stop ? GP_PHASE_END : phase;
    tr.nnz_after = after;
    tr.fired_mask = s.fired_mask;
    tr.rows_live = s.cur.lp.A.rows;
    tr.cols_live = s.cur.lp.A.cols;
    tr.seconds = now_seconds() - t0;
    if (learned) {
      if (hist_count + 1 > hist_cap) {
        hist_cap = 2 * hist_cap + 16;
        hist = (gp_round_trace*)or_xrealloc(hist, sizeof(gp_round_trace) * (size_t)hist_cap);
      }
      hist[hist_count++] = tr;
    }
    if (p->collect_trace) trace_push(out, &tr);
    if (stop) {
      if (!gating) break;
      gating = 0; /* the plain schedule takes over from here */
      phase = GP_PHASE_CLEANUP;
    }
  }
  free(hist);
  materialize(&s);

  out->m1 = s.cur.lp.A.rows;
  out->n1 = s.cur.lp.A.cols;
  out->obj_constant_delta = s.cur.lp.obj_constant - lp->obj_constant;
  out->has_change = s.cur.owns;
  if (s.cur.owns) {
    out->reduced = s.cur.lp;
  } else {
    or_lp_copy(lp, &out->reduced);
  }
  return 0;
}

void or_result_free(or_result* r) {
  or_lp_free(&r->reduced);
  free(r->row_red2org);
  free(r->col_red2org);
  free(r->fixed_idx);
  free(r->fixed_val);
  or_tape_free(&r->tape);
  free(r->trace);
  memset(r, 0, sizeof(*r));
}
