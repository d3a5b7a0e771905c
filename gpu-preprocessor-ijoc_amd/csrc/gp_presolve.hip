// gp_presolve.hip -- presolve scheduler, deferred plans, materialization,
// device recovery record and the C ABI of the engine.
//
// The phase machine restates run_gpu_presolve_impl (gpu_presolve.cu:2081-2645):
// Cleanup (fixed, empty cols, empty rows, singleton rows, singleton cols,
// repeated while anything changes) -> Propagation (bound strengthening,
// redundant rows, dual fix; materialize) -> Aggregation (doubleton
// equations) -> Parallel (parallel rows, parallel columns; materialize).
// Rules write into one deferred plan over the physical LP; the LP is
// compacted (rows/cols dropped, A and A^T rebuilt) only at materialize points.
#include <chrono>
#include <memory>
#include <cstdio>

#include "gp_fold.hpp"
#include "gp_startup.hpp"
#include "gp_tape.hpp"

namespace gp {
const char* last_error_cstr();

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---- recovery record ---------------------------------------------------------------------
struct Record {
  std::unique_ptr<FoldState> fold;  // the folding stage, when it folded the LP
  int32_t m0 = 0, n0 = 0, m1 = 0, n1 = 0;
  DVec<int32_t> rows, cols;  // physical -> original maps
  DVec<int32_t> fixed_idx;
  DVec<double> fixed_val;
  int64_t fixed_count = 0;
  DeviceTape tape;
  std::vector<gp_round_trace> trace;
};

__global__ void k_iota(int32_t* p, int32_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = static_cast<int32_t>(i);
}

__global__ void k_fixed_flag(int32_t n, const uint8_t* keep, const uint8_t* prior, const double* l, const double* u,
                             int32_t* flag) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  flag[j] = (!prior || prior[j]) && !keep[j] && isfinite(l[j]) && isfinite(u[j]) && fabs(l[j] - u[j]) <= 1e-9;
}

__global__ void k_fixed_capture(int32_t n, const int32_t* flag_excl, const uint8_t* keep, const uint8_t* prior,
                                const double* l, const double* u, const int32_t* cols, int32_t* idx, double* val,
                                int64_t base) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const bool f = (!prior || prior[j]) && !keep[j] && isfinite(l[j]) && isfinite(u[j]) && fabs(l[j] - u[j]) <= 1e-9;
  if (!f) return;
  const int64_t k = base + flag_excl[j];
  idx[k] = cols[j];
  val[k] = 0.5 * (l[j] + u[j]);
}

// append_fixed (device_recovery.cu:233-276)
void append_fixed(Record& r, const gp_plan& plan, const gp_device_lp& lp, const uint8_t* keep, const uint8_t* prior) {
  const int32_t n = lp.A.cols;
  if (n == 0) return;
  const double* l = plan.new_l ? plan.new_l : lp.l;
  const double* u = plan.new_u ? plan.new_u : lp.u;
  DVec<int32_t> flag(static_cast<size_t>(n));
  hipLaunchKernelGGL(k_fixed_flag, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, keep, prior, l, u, flag.data());
  GP_LAUNCH_CHECK("fixed flag");
  const int64_t total = exclusive_scan_i32_total(flag.data(), n);
  if (total == 0) return;
  r.fixed_idx.grow_keep(static_cast<size_t>(r.fixed_count + total), static_cast<size_t>(r.fixed_count));
  r.fixed_val.grow_keep(static_cast<size_t>(r.fixed_count + total), static_cast<size_t>(r.fixed_count));
  hipLaunchKernelGGL(k_fixed_capture, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, flag.data(), keep, prior, l, u,
                     r.cols.data(), r.fixed_idx.data(), r.fixed_val.data(), r.fixed_count);
  GP_LAUNCH_CHECK("fixed capture");
  r.fixed_count += total;
}

void flush_plan_tape(Record& r, gp_plan& plan) {
  DeviceTape* t = plan_tape(plan);
  if (!t || t->count == 0) return;
  tape_append_global(r.tape, *t, r.rows.data(), r.m1, r.cols.data(), r.n1);
  t->clear();
}

// ---- working LP -------------------------------------------------------------------------------
struct Working {
  gp_device_lp lp{};
  bool owns = false;
  EntryBounds store;  // per-entry column bounds of lp.A for the row folds
  void free_owned() {
    if (!owns) return;
    csr_pool_free(lp.A);
    csr_pool_free(lp.AT);
    pool_free(lp.c);
    pool_free(lp.l);
    pool_free(lp.u);
    pool_free(lp.AL);
    pool_free(lp.AU);
    owns = false;
  }
};

}  // namespace

// ---- plans and statistics -----------------------------------------------------------------------
void plan_create(const gp_device_lp& lp, gp_plan& plan) {
  std::memset(&plan, 0, sizeof(plan));
  const int32_t m = lp.A.rows, n = lp.A.cols;
  plan.keep_row_mask = pool_alloc<uint8_t>(m);
  plan.keep_col_mask = pool_alloc<uint8_t>(n);
  plan.new_c = pool_alloc<double>(n);
  plan.new_l = pool_alloc<double>(n);
  plan.new_u = pool_alloc<double>(n);
  plan.new_AL = pool_alloc<double>(m);
  plan.new_AU = pool_alloc<double>(m);
  fill_u8(plan.keep_row_mask, 1, m);
  fill_u8(plan.keep_col_mask, 1, n);
  d2d(plan.new_c, lp.c, n);
  d2d(plan.new_l, lp.l, n);
  d2d(plan.new_u, lp.u, n);
  d2d(plan.new_AL, lp.AL, m);
  d2d(plan.new_AU, lp.AU, m);
}

void plan_destroy(gp_plan& plan) {
  pool_free(plan.keep_row_mask);
  pool_free(plan.keep_col_mask);
  pool_free(plan.new_c);
  pool_free(plan.new_l);
  pool_free(plan.new_u);
  pool_free(plan.new_AL);
  pool_free(plan.new_AU);
  if (plan.has_new_A) csr_pool_free(plan.new_A);
  delete plan_tape(plan);
  rule_cache_free(plan.rule_cache);
  std::memset(&plan, 0, sizeof(plan));
}

void stats_create(int32_t m, int32_t n, gp_stats& s) {
  std::memset(&s, 0, sizeof(s));
  s.row_nnz = pool_alloc<int32_t>(m);
  s.empty_row_mask = pool_alloc<uint8_t>(m);
  s.singleton_row_mask = pool_alloc<uint8_t>(m);
  s.singleton_row_col = pool_alloc<int32_t>(m);
  s.singleton_row_val = pool_alloc<double>(m);
  s.col_nnz = pool_alloc<int32_t>(n);
  s.empty_col_mask = pool_alloc<uint8_t>(n);
  s.singleton_column_mask = pool_alloc<uint8_t>(n);
  s.singleton_column_row = pool_alloc<int32_t>(n);
  s.singleton_column_val = pool_alloc<double>(n);
}

void stats_destroy(gp_stats& s) {
  pool_free(s.row_nnz);
  pool_free(s.empty_row_mask);
  pool_free(s.singleton_row_mask);
  pool_free(s.singleton_row_col);
  pool_free(s.singleton_row_val);
  pool_free(s.col_nnz);
  pool_free(s.empty_col_mask);
  pool_free(s.singleton_column_mask);
  pool_free(s.singleton_column_row);
  pool_free(s.singleton_column_val);
  std::memset(&s, 0, sizeof(s));
}

namespace {

bool plan_keeps_all(const gp_plan& plan, int32_t m, int32_t n) {
  return count_u8(plan.keep_row_mask, m) == m && count_u8(plan.keep_col_mask, n) == n;
}

// apply_plan_to_working_lp (gpu_presolve.cu:1953-2072) + update_device_recovery
void apply_plan(Working& cur, gp_plan& plan, Record& r, bool captured) {
  const int32_t m = cur.lp.A.rows, n = cur.lp.A.cols;
  if (cur.owns && !plan.has_new_A && plan_keeps_all(plan, m, n)) {
    flush_plan_tape(r, plan);
    d2d(cur.lp.c, plan.new_c, n);
    d2d(cur.lp.l, plan.new_l, n);
    d2d(cur.lp.u, plan.new_u, n);
    d2d(cur.lp.AL, plan.new_AL, m);
    d2d(cur.lp.AU, plan.new_AU, m);
    cur.lp.obj_constant += plan.obj_constant_delta;
    return;
  }
  DVec<int32_t> row_o2n(m ? m : 1), col_o2n(n ? n : 1);
  const int32_t m_new = build_old_to_new(plan.keep_row_mask, m, row_o2n.data());
  const int32_t n_new = build_old_to_new(plan.keep_col_mask, n, col_o2n.data());
  const gp_csr& src = plan.has_new_A ? plan.new_A : cur.lp.A;
  gp_device_lp next{};
  // a new matrix: the entry-bounds store starts over at its first full fold
  cur.store.drop();
  compact_csr(src, plan.keep_row_mask, plan.keep_col_mask, row_o2n.data(), m_new, n_new, next.A);
  // Without a rewritten matrix the transpose of the compacted A is the compacted
  // AT: dropping dead entries keeps each column's increasing row order.
  if (plan.has_new_A)
    transpose_csr(next.A, next.AT);
  else
    compact_csr(cur.lp.AT, plan.keep_col_mask, plan.keep_row_mask, col_o2n.data(), n_new, m_new,
                next.AT);
  {
    // c, l, u, AL, AU read through the old -> new maps, written compacted
    TimedScope ts(GP_KT_VECTORS, 24.0 * n + 16.0 * m + 24.0 * n_new + 16.0 * m_new + 4.0 * (m + n));
    next.c = pool_alloc<double>(n_new);
    next.l = pool_alloc<double>(n_new);
    next.u = pool_alloc<double>(n_new);
    next.AL = pool_alloc<double>(m_new);
    next.AU = pool_alloc<double>(m_new);
    gather_compact(next.c, plan.new_c, col_o2n.data(), n);
    gather_compact(next.l, plan.new_l, col_o2n.data(), n);
    gather_compact(next.u, plan.new_u, col_o2n.data(), n);
    gather_compact(next.AL, plan.new_AL, row_o2n.data(), m);
    gather_compact(next.AU, plan.new_AU, row_o2n.data(), m);
  }
  next.obj_constant = cur.lp.obj_constant + plan.obj_constant_delta;
  flush_plan_tape(r, plan);
  if (!captured && n_new != n) append_fixed(r, plan, cur.lp, plan.keep_col_mask, nullptr);
  {
    DVec<int32_t> rows(m_new ? m_new : 1), cols(n_new ? n_new : 1);
    gather_compact_i32(rows.data(), r.rows.data(), row_o2n.data(), m);
    gather_compact_i32(cols.data(), r.cols.data(), col_o2n.data(), n);
    r.rows = std::move(rows);
    r.cols = std::move(cols);
    r.m1 = m_new;
    r.n1 = n_new;
  }
  cur.free_owned();
  cur.lp = next;
  cur.owns = true;
}

struct Deferred {
  gp_plan plan{};
  gp_stats stats{};
  DVec<uint8_t> entry_cols;
  DVec<int64_t> live_total;
  // masks the statistics were last computed for, and incremental-update scratch
  DVec<uint8_t> snap_rows, snap_cols, touched_rows, touched_cols;
  DVec<int32_t> inc_scratch;
  bool active = false, pending = false, stats_dirty = false, nnz_dirty = false, have_snapshot = false;
  bool fresh = false;          // plan masks still all ones
  bool entry_current = false;  // entry_cols equals the plan's column mask
  void reset() {
    if (active) {
      plan_destroy(plan);
      stats_destroy(stats);
      entry_cols.reset();
      live_total.reset();
      snap_rows.reset();
      snap_cols.reset();
      touched_rows.reset();
      touched_cols.reset();
      inc_scratch.reset();
    }
    active = pending = stats_dirty = nnz_dirty = have_snapshot = fresh = entry_current = false;
  }
};

struct Scheduler {
  Working cur;
  Deferred d;
  Record* r = nullptr;
  const gp_params* p = nullptr;
  gp_summary* out = nullptr;
  bool terminal = false;
  uint64_t deletion_serial = 0;
  double start = 0.0;
  int32_t fired = 0;

  bool time_exceeded() const {
    if (!(p->max_time > 0.0 && std::isfinite(p->max_time))) return false;
    return now_s() - start >= p->max_time;
  }

  void prepare_deferred() {
    if (d.active) return;
    plan_create(cur.lp, d.plan);
    rule_caches(d.plan).store = &cur.store;
    if (p->record_postsolve_tape) d.plan.tape = new DeviceTape();
    stats_create(cur.lp.A.rows, cur.lp.A.cols, d.stats);
    const int32_t n = cur.lp.A.cols;
    d.entry_cols.resize(n ? n : 1);
    d.live_total.resize(1);
    d.plan.rule_entry_keep_col = d.entry_cols.data();
    d.active = d.stats_dirty = d.nnz_dirty = d.fresh = true;
  }

  // Live statistics of the plan's masks: a full sweep for a new plan, then
  // updates that visit only the rows and columns deleted since the last one.
  void refresh_stats() {
    if (!d.stats_dirty) return;
    const int32_t m = cur.lp.A.rows, n = cur.lp.A.cols;
    bool done = false, synced = false;
    if (d.have_snapshot && p->incremental_live_stats) {
      // an accepted update leaves the snapshot equal to the masks
      done = synced = update_live_stats(d.stats, cur.lp, d.plan, d.snap_rows.data(), d.snap_cols.data(),
                                        d.touched_rows.data(), d.touched_cols.data(), d.live_total.data(),
                                        d.inc_scratch.data(), true);
    }
    if (!done && d.fresh) {
      live_stats_all_live(d.stats, cur.lp, d.live_total.data());
      done = true;
    }
    if (!done || !d.have_snapshot) {
      if (!done) refresh_live_stats(d.stats, cur.lp, d.plan, d.live_total.data());
      if (!d.have_snapshot) {
        d.snap_rows.resize(m ? m : 1);
        d.snap_cols.resize(n ? n : 1);
        d.touched_rows.resize(m ? m : 1);
        d.touched_cols.resize(n ? n : 1);
        fill_u8(d.touched_rows.data(), 0, m ? m : 1);  // updates keep them zero from here on
        fill_u8(d.touched_cols.data(), 0, n ? n : 1);
        d.inc_scratch.resize(4);
        d.have_snapshot = true;
      }
    }
    if (!synced) {
      d2d(d.snap_rows.data(), d.plan.keep_row_mask, static_cast<size_t>(m));
      d2d(d.snap_cols.data(), d.plan.keep_col_mask, static_cast<size_t>(n));
    }
    d.stats_dirty = false;
    d.fresh = false;
  }

  int64_t live_nnz() {
    if (!d.active) return cur.lp.A.nnz;
    refresh_stats();
    if (d.nnz_dirty) {
      d.stats.live_nnz = read1(d.live_total.data());
      d.nnz_dirty = false;
    }
    return d.stats.live_nnz;
  }

  // deferred_rule (gpu_presolve.cu:2289-2338)
  bool deferred_rule(bool enabled, int32_t rule) {
    if (!enabled || terminal || time_exceeded()) return false;
    prepare_deferred();
    refresh_stats();
    gp_plan& plan = d.plan;
    const int32_t n = cur.lp.A.cols;
    // the column mask at the rule's entry; only a rule with a column action
    // deletes columns, so the copy is redone after such a rule only
    if (!d.entry_current) {
      d2d(d.entry_cols.data(), plan.keep_col_mask, n);
      d.entry_current = true;
    }
    plan.has_change = plan.has_row_action = plan.has_col_action = 0;
    plan.has_redundant_row_deletion = plan.has_row_deletion = 0;
    plan.obj_constant_delta = 0.0;
    apply_rule(rule, plan, cur.lp, d.stats, *p);
    out->has_infeasible = plan.has_infeasible;
    out->has_unbounded = plan.has_unbounded;
    terminal = plan.has_infeasible || plan.has_unbounded;
    if (plan.has_col_action || terminal) d.entry_current = false;
    const bool changed = plan.has_change && !terminal;
    if (changed) {
      if (plan.has_row_deletion || plan.has_col_action) ++deletion_serial;
      cur.lp.obj_constant += plan.obj_constant_delta;
      flush_plan_tape(*r, plan);
      if (plan.has_col_action) append_fixed(*r, plan, cur.lp, plan.keep_col_mask, d.entry_cols.data());
      d.pending = true;
      if (plan.has_row_action || plan.has_col_action) d.stats_dirty = d.nnz_dirty = true;
      fired |= 1 << rule;
    }
    plan.obj_constant_delta = 0.0;
    return changed;
  }

  void materialize() {
    if (d.pending) {
      d.plan.has_change = 1;
      apply_plan(cur, d.plan, *r, true);
    }
    d.reset();
  }

  // run_materialized_rule (gpu_presolve.cu:2144-2174)
  bool materialized_rule(bool enabled, int32_t rule) {
    if (!enabled || terminal || time_exceeded()) return false;
    gp_plan plan;
    gp_stats stats;
    std::memset(&stats, 0, sizeof(stats));
    plan_create(cur.lp, plan);
    if (p->record_postsolve_tape) plan.tape = new DeviceTape();
    try {
      apply_rule(rule, plan, cur.lp, stats, *p);
      out->has_infeasible = plan.has_infeasible;
      out->has_unbounded = plan.has_unbounded;
      terminal = plan.has_infeasible || plan.has_unbounded;
      const bool changed = !terminal && plan.has_change;
      if (changed) {
        apply_plan(cur, plan, *r, false);
        fired |= 1 << rule;
      }
      plan_destroy(plan);
      return changed;
    } catch (...) {
      plan_destroy(plan);
      throw;
    }
  }
};

// The Startup projection batch (gpu_presolve.cu:2361-2410): column
// domination, two-row elimination and zero-cost products share one
// deletion-only plan, created when the first of them applies; each analysis
// after that reads the plan's masks.  The plan is applied once.
void startup_batch(Scheduler& s) {
  static const int32_t order[3] = {GP_STARTUP_COLUMN_DOMINATION, GP_STARTUP_TWO_ROW_ELIMINATION,
                                   GP_STARTUP_ZERO_COST_PRODUCTS};
  const gp_params& p = *s.p;
  const int enabled[3] = {p.enable_column_domination, p.enable_two_row_variable_elimination,
                          p.enable_zero_cost_product_variables};
  gp_plan plan;
  bool allocated = false;
  int32_t flags = 0;
  auto get_plan = [&]() -> gp_plan* {
    if (!allocated) {
      plan_create(s.cur.lp, plan);
      if (p.record_postsolve_tape) plan.tape = new DeviceTape();
      allocated = true;
    }
    return &plan;
  };
  try {
    for (int k = 0; k < 3; ++k) {
      if (!enabled[k] || s.terminal || s.time_exceeded()) continue;
      flags |= startup_rule(order[k], s.cur.lp, p, allocated ? &plan : nullptr, get_plan).primal_only;
    }
    if (allocated && plan.has_change) {
      apply_plan(s.cur, plan, *s.r, false);
      s.out->primal_only_reductions |= flags;
    }
    if (allocated) plan_destroy(plan);
  } catch (...) {
    if (allocated) plan_destroy(plan);
    throw;
  }
}

bool good_progress(int64_t before, int64_t after, int64_t minimum) {
  return before >= minimum && after <= before - minimum;
}

void run_presolve(const gp_device_lp& lp, const gp_params& p, gp_summary& out) {
  Record* rec = new Record();
  out.record = rec;
  rec->m0 = rec->m1 = lp.A.rows;
  rec->n0 = rec->n1 = lp.A.cols;
  rec->rows.resize(lp.A.rows ? lp.A.rows : 1);
  rec->cols.resize(lp.A.cols ? lp.A.cols : 1);
  if (lp.A.rows) hipLaunchKernelGGL(k_iota, dim3(grid_for(lp.A.rows)), dim3(kBlock), 0, 0, rec->rows.data(), lp.A.rows);
  if (lp.A.cols) hipLaunchKernelGGL(k_iota, dim3(grid_for(lp.A.cols)), dim3(kBlock), 0, 0, rec->cols.data(), lp.A.cols);
  GP_LAUNCH_CHECK("identity maps");
  GP_CHECK(hipDeviceSynchronize());
  if (BatchContext* bc = batch_context()) {
    bc->row_map = &rec->rows;
    bc->col_map = &rec->cols;
  }

  Scheduler s;
  s.cur.lp = lp;
  s.cur.owns = false;
  s.r = rec;
  s.p = &p;
  s.out = &out;
  s.start = now_s();
  try {
    // Startup (gpu_presolve.cu:2361-2454): the structural rules, one iteration
    if (p.enable_startup && p.max_iters > 0 && !s.time_exceeded()) {
      if (p.enable_structure_specific_rules) startup_batch(s);
      out.iterations++;
    }
    int phase = GP_PHASE_CLEANUP;
    const int64_t min_dec = 100;
    int bound_only_rounds = 0;
    const int64_t initial_live = p.max_iters > 0 && !s.terminal && !s.time_exceeded() ? s.live_nnz() : 0;
    const int configured = p.bound_strengthening_max_bound_only_rounds > 1 ? p.bound_strengthening_max_bound_only_rounds : 1;
    int limit = configured;
    if (p.bound_strengthening_bound_only_nnz_round_budget > 0 && initial_live > 0) {
      int64_t q = p.bound_strengthening_bound_only_nnz_round_budget / initial_live;
      if (q < 1) q = 1;
      if (q < limit) limit = static_cast<int>(q);
    }
    {
      const int64_t m = s.cur.lp.A.rows, n = s.cur.lp.A.cols;
      if (m >= 512 && m <= 2048 && n >= 128 * m && n <= 256 * m && initial_live >= 2048 * m &&
          initial_live <= 8192 * m) {
        const int two = configured < 2 ? configured : 2;
        if (two > limit) limit = two;
      }
    }
    // Learned pass scheduler: before each round the network scores the rule
    // groups and only groups at or above their threshold run.  When the gated
    // loop stops, the plain schedule runs on from that state (fixpoint check
    // and fallback) until its own stop.
    const bool learned = p.learned_schedule != 0;
    LearnedGate gate;
    if (learned) gate.init(p.scheduler_weights);
    bool gating = learned;
    std::vector<gp_round_trace> hist;
    const int32_t m0 = lp.A.rows, n0 = lp.A.cols;
    while (p.max_iters > 0 && !s.terminal && !s.time_exceeded()) {
      if (phase != GP_PHASE_CLEANUP) {
        if (out.iterations >= p.max_iters) {
          out.stopped_at_iteration_limit = 1;
          break;
        }
        out.iterations++;
      }
      const double t0 = now_s();
      const int64_t before = s.live_nnz();
      gp_round_trace tr{};
      tr.phase = phase;
      tr.nnz_before = before;
      tr.bound_only_rounds = bound_only_rounds;
      tr.bound_only_limit = limit;
      tr.final_sweep = learned && !gating;
      uint32_t run = ~0u;
      if (gating) {
        float f[GP_MLP_IN];
        schedule_features(hist.data(), static_cast<int32_t>(hist.size()), phase, before, initial_live, m0, n0,
                          bound_only_rounds, limit, f);
        run = gate.decide(f);
      }
      // rule enabled by the parameters and its group chosen for this round
      auto on = [&](int32_t enabled, int32_t rule) -> bool {
        if (!enabled) return false;
        if ((run >> rule_group(rule)) & 1u) return true;
        tr.skipped_mask |= 1 << rule;
        return false;
      };
      s.fired = 0;
      int64_t after = before;
      bool stop = false;
      if (phase == GP_PHASE_CLEANUP) {
        const uint64_t serial = s.deletion_serial;
        bool changed = s.deferred_rule(on(p.enable_infeasible_and_fixed_variables, GP_RULE_FIXED_VARIABLES),
                                       GP_RULE_FIXED_VARIABLES);
        changed = s.deferred_rule(on(p.enable_empty_columns, GP_RULE_EMPTY_COLUMNS), GP_RULE_EMPTY_COLUMNS) || changed;
        changed = s.deferred_rule(on(p.enable_empty_rows, GP_RULE_EMPTY_ROWS), GP_RULE_EMPTY_ROWS) || changed;
        changed = s.deferred_rule(on(p.enable_singleton_rows, GP_RULE_SINGLETON_ROWS), GP_RULE_SINGLETON_ROWS) ||
                  changed;
        changed = s.deferred_rule(on(p.enable_singleton_columns, GP_RULE_SINGLETON_COLUMNS),
                                  GP_RULE_SINGLETON_COLUMNS) ||
                  changed;
        if (s.deletion_serial != serial) bound_only_rounds = 0;
        after = s.live_nnz();
        phase = changed ? GP_PHASE_CLEANUP : GP_PHASE_PROPAGATION;
      } else if (phase == GP_PHASE_PROPAGATION) {
        const uint64_t serial = s.deletion_serial;
        bool bounds_changed = false;
        if (on(p.enable_bound_strengthening, GP_RULE_BOUND_STRENGTHENING) && !s.terminal && !s.time_exceeded()) {
          s.prepare_deferred();
          // The reference bounds are the plan's bounds as this call starts
          // (the propagation phase, gpu_presolve.cu:2456-2470, copies them).  The
          // rule reads them before it writes any bound, so here it reads the
          // plan's own arrays.
          if (s.cur.lp.A.cols) {
            s.d.plan.propagation_reference_l = s.d.plan.new_l;
            s.d.plan.propagation_reference_u = s.d.plan.new_u;
          }
          bounds_changed = s.deferred_rule(true, GP_RULE_BOUND_STRENGTHENING);
        }
        s.deferred_rule(on(p.enable_infeasible_and_redundant_rows, GP_RULE_REDUNDANT_ROWS), GP_RULE_REDUNDANT_ROWS);
        s.deferred_rule(on(p.enable_dual_fix, GP_RULE_DUAL_FIX), GP_RULE_DUAL_FIX);
        const bool deleted = s.deletion_serial != serial;
        bound_only_rounds = deleted ? 0 : bound_only_rounds + 1;
        s.materialize();
        after = s.live_nnz();
        const bool nnz_progress = good_progress(before, after, min_dec);
        const bool bound_progress = bounds_changed && bound_only_rounds < limit;
        phase = nnz_progress || bound_progress ? GP_PHASE_CLEANUP : GP_PHASE_AGGREGATION;
        if (phase == GP_PHASE_AGGREGATION) bound_only_rounds = 0;
      } else if (phase == GP_PHASE_AGGREGATION) {
        s.materialized_rule(on(p.enable_doubleton_equations, GP_RULE_DOUBLETON_EQUATIONS),
                            GP_RULE_DOUBLETON_EQUATIONS);
        after = s.live_nnz();
        phase = good_progress(before, after, min_dec) ? GP_PHASE_CLEANUP : GP_PHASE_PARALLEL;
      } else {
        bool changed = s.deferred_rule(on(p.enable_parallel_rows, GP_RULE_PARALLEL_ROWS), GP_RULE_PARALLEL_ROWS);
        changed = s.deferred_rule(on(p.enable_parallel_columns, GP_RULE_PARALLEL_COLUMNS), GP_RULE_PARALLEL_COLUMNS) ||
                  changed;
        s.materialize();
        after = s.live_nnz();
        const bool cont = changed && good_progress(before, after, min_dec);
        if (!cont) stop = true;
        else phase = GP_PHASE_CLEANUP;
      }
      tr.next_phase = stop ? GP_PHASE_END : phase;
      tr.nnz_after = after;
      tr.fired_mask = s.fired;
      tr.rows_live = s.cur.lp.A.rows;
      tr.cols_live = s.cur.lp.A.cols;
      tr.seconds = now_s() - t0;
      if (learned) hist.push_back(tr);
      if (p.collect_trace) rec->trace.push_back(tr);
      if (stop) {
        if (!gating) break;
        gating = false;  // the plain schedule takes over from here
        phase = GP_PHASE_CLEANUP;
      }
    }
    s.materialize();
    GP_CHECK(hipDeviceSynchronize());
  } catch (...) {
    s.d.reset();
    s.cur.free_owned();
    throw;
  }
  out.elapsed_seconds = now_s() - s.start;
  out.reduced_rows = s.cur.lp.A.rows;
  out.reduced_cols = s.cur.lp.A.cols;
  out.reduced_nnz = s.cur.lp.A.nnz;
  out.obj_constant_delta = s.cur.lp.obj_constant - lp.obj_constant;
  out.has_change = s.cur.owns;
  out.reduced_lp = s.cur.lp;
  out.owns_reduced_lp = s.cur.owns;
  s.cur.owns = false;
  if (!rec->trace.empty()) {
    out.trace_count = static_cast<int32_t>(rec->trace.size());
    out.trace = rec->trace.data();
  }
}

void validate_lp(const gp_device_lp* lp) {
  if (!lp) throw std::invalid_argument("null LP");
  if (lp->A.rows < 0 || lp->A.cols < 0 || lp->A.nnz < 0)
    throw std::invalid_argument("negative LP dimensions");
  if (lp->AT.rows != lp->A.cols || lp->AT.cols != lp->A.rows || lp->AT.nnz != lp->A.nnz)
    throw std::invalid_argument("AT does not match A");
  if ((lp->A.rows && (!lp->AL || !lp->AU)) || (lp->A.cols && (!lp->c || !lp->l || !lp->u)))
    throw std::invalid_argument("missing LP vector");
}

}  // namespace

RecordView record_view(const void* record) {
  const Record* r = static_cast<const Record*>(record);
  return RecordView{r->m0,
                    r->n0,
                    r->m1,
                    r->n1,
                    r->rows.data(),
                    r->cols.data(),
                    r->fixed_idx.data(),
                    r->fixed_val.data(),
                    r->fixed_count,
                    &r->tape,
                    r->fold && r->fold->applied ? r->fold.get() : nullptr};
}

}  // namespace gp

// ============================================================================ C ABI
using namespace gp;

extern "C" {

const char* gp_last_error(void) { return last_error_cstr(); }
int gp_abi_version(void) { return GP_ABI_VERSION; }

void gp_default_params(gp_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->struct_size = static_cast<int32_t>(sizeof(gp_params));
  p->max_iters = 512;
  p->max_time = HUGE_VAL;
  p->verbose = 0;
  p->record_postsolve_tape = 1;
  p->record_postsolve_tape_cpu = 0;
  p->enable_presolve = 1;
  p->enable_folding = 1;
  p->folding_tolerance = 1.0e-8;
  p->feasibility_tol = 1.0e-6;
  p->bound_tol = 1.0e-6;
  p->zero_tol = 1.0e-10;
  p->bound_strengthening_upper_relaxation_factor = 1.0;
  p->bound_strengthening_max_rounds = 64;
  p->bound_strengthening_max_bound_only_rounds = 64;
  p->bound_strengthening_preserve_free_zero_cost = 1;
  p->doubleton_equations_max_fill_in_proxy = 96;
  p->doubleton_equations_scan = 1;
  p->doubleton_equations_min_selected_per_batch = 256;
  p->doubleton_equations_min_selected_ratio = 0.005;
  p->doubleton_equations_max_batch_rounds = 0;
  p->doubleton_equations_max_time = 0.0;
  p->enable_infeasible_and_fixed_variables = 1;
  p->enable_empty_rows = 1;
  p->enable_singleton_rows = 1;
  p->enable_infeasible_and_redundant_rows = 1;
  p->enable_bound_strengthening = 1;
  p->enable_parallel_rows = 1;
  p->enable_empty_columns = 1;
  p->enable_singleton_columns = 1;
  p->enable_doubleton_equations = 1;
  p->enable_dual_fix = 1;
  p->dual_fix_max_nonzero_cost_col_degree = 128;
  p->enable_parallel_columns = 1;
  p->parallel_columns_max_abs_ratio = 2.0;
  p->parallel_columns_scale_guard_max_group_size = 4;
  p->parallel_columns_scale_guard_min_groups = 128;
  p->enable_structure_specific_rules = 1;
  p->enable_two_row_variable_elimination = 1;
  p->two_row_variable_elimination_max_row_nnz = 8;
  p->two_row_variable_elimination_safety_tol = 1.0e-8;
  p->two_row_variable_elimination_min_candidates = 128;
  p->two_row_variable_elimination_min_candidate_ratio = 0.01;
  p->enable_zero_cost_product_variables = 1;
  p->zero_cost_product_variables_min_candidates = 128;
  p->zero_cost_product_variables_min_candidate_ratio = 0.01;
  p->enable_piecewise_linear_reformulation_variable_pair = 1;
  p->piecewise_linear_reformulation_variable_pair_epigraph_probe_pairs = 256;
  p->piecewise_linear_reformulation_variable_pair_epigraph_min_pairs = 4096;
  p->piecewise_linear_reformulation_variable_pair_epigraph_min_potential_gain = 0.20;
  p->piecewise_linear_reformulation_variable_pair_epigraph_min_actual_gain = 0.20;
  p->piecewise_linear_reformulation_variable_pair_epigraph_max_tail_cols = 8;
  p->piecewise_linear_reformulation_variable_pair_epigraph_max_quotient_terms_per_row = 5;
  p->piecewise_linear_reformulation_variable_pair_shifted_antipodal_probe_pairs = 256;
  p->piecewise_linear_reformulation_variable_pair_shifted_antipodal_min_pairs = 4096;
  p->piecewise_linear_reformulation_variable_pair_shifted_antipodal_min_potential_gain = 0.20;
  p->piecewise_linear_reformulation_variable_pair_shifted_antipodal_min_actual_gain = 0.20;
  p->piecewise_linear_reformulation_variable_pair_shifted_antipodal_max_quotient_terms_per_row = 6;
  p->enable_column_domination = 1;
  p->column_domination_probe_rows = 64;
  p->column_domination_probe_cols = 256;
  p->column_domination_probe_nnz = 256;
  p->column_domination_min_fixed_cols = 4096;
  p->column_domination_min_candidate_ratio = 0.20;
  p->enable_piecewise_linear_reformulation_absolute_value = 1;
  p->piecewise_linear_reformulation_absolute_value_residual_bound_mode = 0;
  p->piecewise_linear_reformulation_absolute_value_residual_bound_as_free_min = 1.0e8;
  p->enable_redundant_variable_bounds = 0;
  p->enable_startup = 1;
  p->bound_strengthening_bound_only_nnz_round_budget = 8000000;
  p->incremental_live_stats = 1;
}

int gp_preprocess(const gp_device_lp* input, const gp_params* params, gp_summary* out) {
  return guarded([&] {
    if (!out || !params) throw std::invalid_argument("null argument");
    if (params->struct_size != static_cast<int32_t>(sizeof(gp_params)))
      throw std::invalid_argument("gp_params.struct_size mismatch (call gp_default_params)");
    validate_lp(input);
    std::memset(out, 0, sizeof(*out));
    const double t0 = now_s();
    GP_CHECK(hipDeviceSynchronize());
    KernelClock& kc = kclock();
    kc.enabled = params->kernel_timing != 0;
    for (int k = 0; k < GP_KT_COUNT; ++k) {
      kc.ms[k] = 0.0;
      kc.launches[k] = 0;
      kc.bytes[k] = 0.0;
    }
    out->original_rows = input->A.rows;
    out->original_cols = input->A.cols;
    out->presolve_enabled = params->enable_presolve;
    out->folding_applied = 0;
    // Requested stages this engine does not run: the piecewise-linear
    // reformulation branches of Startup (gpu_presolve.cu:2412-2449).  The
    // reduction is the one with those branches off.
    if (params->enable_presolve && params->enable_startup && params->max_iters > 0 &&
        params->enable_structure_specific_rules &&
        (params->enable_piecewise_linear_reformulation_variable_pair ||
         params->enable_piecewise_linear_reformulation_absolute_value))
      out->unsupported_stages |= GP_STAGE_STRUCTURE_SPECIFIC_RULES;
    const double setup_end = now_s();
    out->workspace_setup_seconds = setup_end - t0;
    try {
      // folding first (run_gpu_presolve_with_folding, folding.cu:173-209):
      // presolve then runs on the folded LP
      std::unique_ptr<FoldState> fold;
      if (params->enable_folding) fold = fold_lp(*input, params->folding_tolerance);
      if (fold) fold->fill(out->folding);
      else {
        out->folding.original_rows = out->folding.folded_rows = input->A.rows;
        out->folding.original_cols = out->folding.folded_cols = input->A.cols;
      }
      out->folding_applied = fold && fold->applied ? 1 : 0;
      const gp_device_lp* work = out->folding_applied ? &fold->lp : input;
      if (params->enable_presolve) {
        run_presolve(*work, *params, *out);
      } else {
        // identity record, reduced LP borrows the input (folding.cu:145-170)
        Record* rec = new Record();
        out->record = rec;
        rec->m0 = rec->m1 = work->A.rows;
        rec->n0 = rec->n1 = work->A.cols;
        rec->rows.resize(work->A.rows ? work->A.rows : 1);
        rec->cols.resize(work->A.cols ? work->A.cols : 1);
        if (work->A.rows)
          hipLaunchKernelGGL(k_iota, dim3(grid_for(work->A.rows)), dim3(kBlock), 0, 0, rec->rows.data(), work->A.rows);
        if (work->A.cols)
          hipLaunchKernelGGL(k_iota, dim3(grid_for(work->A.cols)), dim3(kBlock), 0, 0, rec->cols.data(), work->A.cols);
        out->reduced_lp = *work;
        out->reduced_rows = work->A.rows;
        out->reduced_cols = work->A.cols;
        out->reduced_nnz = work->A.nnz;
      }
      // the record keeps the folded LP and the map alive (the reduced LP may
      // borrow the folded arrays)
      if (out->folding_applied) static_cast<Record*>(out->record)->fold = std::move(fold);
      kc.flush();
    } catch (...) {
      kc.segments.clear();
      kc.open.clear();
      kc.enabled = false;
      gp_free_summary(out);
      throw;
    }
    for (int k = 0; k < GP_KT_COUNT; ++k) {
      out->kernel_ms[k] = kc.ms[k];
      out->kernel_launches[k] = kc.launches[k];
      out->kernel_bytes[k] = kc.bytes[k];
    }
    kc.enabled = false;
    GP_CHECK(hipDeviceSynchronize());
    out->preprocessing_seconds = now_s() - t0;
  });
}

void gp_free_summary(gp_summary* s) {
  if (!s) return;
  if (s->owns_reduced_lp) {
    csr_pool_free(s->reduced_lp.A);
    csr_pool_free(s->reduced_lp.AT);
    pool_free(s->reduced_lp.c);
    pool_free(s->reduced_lp.l);
    pool_free(s->reduced_lp.u);
    pool_free(s->reduced_lp.AL);
    pool_free(s->reduced_lp.AU);
  }
  delete static_cast<Record*>(s->record);
  std::memset(s, 0, sizeof(*s));
}

int gp_record_get_sizes(const gp_summary* s, gp_record_sizes* sizes) {
  return guarded([&] {
    if (!s || !sizes || !s->record) throw std::invalid_argument("no record");
    const Record* r = static_cast<const Record*>(s->record);
    sizes->m0 = r->m0;
    sizes->n0 = r->n0;
    sizes->m1 = r->m1;
    sizes->n1 = r->n1;
    sizes->fixed_count = static_cast<int32_t>(r->fixed_count);
    sizes->record_count = r->tape.count;
    sizes->index_count = r->tape.icount;
    sizes->value_count = r->tape.vcount;
  });
}

int gp_record_copy(const gp_summary* s, int32_t* row_red2org, int32_t* col_red2org, int32_t* fixed_idx,
                   double* fixed_val, int32_t* types, gp_offset* index_starts, gp_offset* value_starts,
                   uint8_t* dual_modes, int32_t* indices, double* vals) {
  return guarded([&] {
    if (!s || !s->record) throw std::invalid_argument("no record");
    const Record* r = static_cast<const Record*>(s->record);
    if (row_red2org) d2h(row_red2org, r->rows.data(), static_cast<size_t>(r->m1));
    if (col_red2org) d2h(col_red2org, r->cols.data(), static_cast<size_t>(r->n1));
    if (fixed_idx) d2h(fixed_idx, r->fixed_idx.data(), static_cast<size_t>(r->fixed_count));
    if (fixed_val) d2h(fixed_val, r->fixed_val.data(), static_cast<size_t>(r->fixed_count));
    tape_copy_to_host(r->tape, types, index_starts, value_starts, dual_modes, indices, vals);
  });
}

// ---- rule-level operator API -------------------------------------------------------------
int gp_plan_create(const gp_device_lp* lp, gp_plan* plan) {
  return guarded([&] {
    if (!lp || !plan) throw std::invalid_argument("null argument");
    plan_create(*lp, *plan);
    GP_CHECK(hipDeviceSynchronize());
  });
}

void gp_plan_destroy(gp_plan* plan) {
  if (!plan) return;
  try {
    plan_destroy(*plan);
  } catch (...) {
  }
}

int gp_plan_enable_tape(gp_plan* plan) {
  return guarded([&] {
    if (!plan) throw std::invalid_argument("null plan");
    if (!plan->tape) plan->tape = new DeviceTape();
  });
}

int gp_stats_create(int32_t m, int32_t n, gp_stats* stats) {
  return guarded([&] {
    if (!stats || m < 0 || n < 0) throw std::invalid_argument("bad stats arguments");
    stats_create(m, n, *stats);
  });
}

void gp_stats_destroy(gp_stats* stats) {
  if (!stats) return;
  try {
    stats_destroy(*stats);
  } catch (...) {
  }
}

int gp_refresh_live_stats(gp_stats* stats, const gp_device_lp* lp, const gp_plan* plan) {
  return guarded([&] {
    if (!stats || !lp || !plan) throw std::invalid_argument("null argument");
    DVec<int64_t> total(1);
    refresh_live_stats(*stats, *lp, *plan, total.data());
    stats->live_nnz = read1(total.data());
  });
}

int gp_update_live_stats(gp_stats* stats, const gp_device_lp* lp, const gp_plan* plan, const uint8_t* old_keep_row,
                         const uint8_t* old_keep_col, int32_t* accepted) {
  return guarded([&] {
    if (!stats || !lp || !plan || !old_keep_row || !old_keep_col || !accepted)
      throw std::invalid_argument("null argument");
    const int32_t m = lp->A.rows, n = lp->A.cols;
    DVec<uint8_t> touched_rows(m ? m : 1), touched_cols(n ? n : 1);
    DVec<int64_t> total(1);
    DVec<int32_t> scratch(4);
    const int64_t before = stats->live_nnz;
    h2d(total.data(), &before, 1);
    // the caller's masks are read only (no sync)
    const bool ok = update_live_stats(*stats, *lp, *plan, const_cast<uint8_t*>(old_keep_row),
                                      const_cast<uint8_t*>(old_keep_col), touched_rows.data(), touched_cols.data(),
                                      total.data(), scratch.data());
    *accepted = ok ? 1 : 0;
    if (ok) stats->live_nnz = read1(total.data());
  });
}

int gp_rule_apply(int32_t rule, gp_plan* plan, const gp_device_lp* lp, const gp_stats* stats, const gp_params* params) {
  return guarded([&] {
    if (!plan || !lp || !stats || !params) throw std::invalid_argument("null argument");
    // A caller of this entry point may rewrite the plan's vectors and masks,
    // or pass another LP, between calls.  The per-plan caches (stored entry
    // bounds, singleton-column row activities, column maxima) assume neither
    // happens, so each public call starts without them, as the reference's
    // stateless rule calls do; only the scheduler reuses them.
    rule_cache_free(plan->rule_cache);
    plan->rule_cache = nullptr;
    apply_rule(rule, *plan, *lp, *stats, *params);
    GP_CHECK(hipDeviceSynchronize());
  });
}

int gp_fixed_variable_tape(gp_plan* plan, const uint8_t* fixed, const uint8_t* excluded, const double* fixed_val,
                           const uint8_t* keep_row, const double* c, const gp_csr* AT, int32_t n,
                           int32_t compute_objective, double* objective) {
  return guarded([&] {
    if (objective) *objective = 0.0;
    DeviceTape* tape = plan ? plan_tape(*plan) : nullptr;
    if (n <= 0 || (!tape && !compute_objective)) return;
    if (!fixed || !fixed_val || !keep_row || !c || !AT) throw std::invalid_argument("null argument");
    if (AT->rows != n) throw std::invalid_argument("AT rows must equal n");
    const double obj = fixed_variable_records(tape, fixed, fixed_val, keep_row, c, *AT, n, true,
                                              compute_objective ? FixedObjective::Extended : FixedObjective::Skip,
                                              excluded);
    GP_CHECK(hipDeviceSynchronize());
    if (objective) *objective = obj;
  });
}

int gp_plan_tape_sizes(const gp_plan* plan, int64_t* records, int64_t* indices, int64_t* values) {
  return guarded([&] {
    if (!plan) throw std::invalid_argument("null plan");
    const DeviceTape* t = plan_tape(*plan);
    if (records) *records = t ? t->count : 0;
    if (indices) *indices = t ? t->icount : 0;
    if (values) *values = t ? t->vcount : 0;
  });
}

int gp_plan_tape_copy(const gp_plan* plan, int32_t* types, gp_offset* index_starts, gp_offset* value_starts,
                      uint8_t* dual_modes, int32_t* indices, double* vals) {
  return guarded([&] {
    if (!plan || !plan->tape) throw std::invalid_argument("plan has no tape");
    tape_copy_to_host(*plan_tape(*plan), types, index_starts, value_starts, dual_modes, indices, vals);
  });
}

int gp_row_activity_summary(const gp_csr* A, const double* l, const double* u, double zero_tol, const uint8_t* keep_row,
                            const uint8_t* keep_col, double* min_fin, double* max_fin, int32_t* min_ninf,
                            int32_t* max_pinf) {
  return guarded([&] {
    if (!A) throw std::invalid_argument("null matrix");
    row_activity_summary(*A, l, u, zero_tol, keep_row, keep_col, min_fin, max_fin, min_ninf, max_pinf);
    GP_CHECK(hipDeviceSynchronize());
  });
}

int gp_parallel_row_hashes(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, double zero_tol,
                           uint64_t* hash) {
  return guarded([&] {
    if (!A || !keep_row || !hash) throw std::invalid_argument("null argument");
    parallel_row_hashes(*A, keep_row, keep_col, zero_tol, hash);
    GP_CHECK(hipDeviceSynchronize());
  });
}

int gp_radix_sort_pairs(uint64_t* keys, int32_t* vals, int64_t n, int32_t end_bit) {
  return guarded([&] {
    if (n < 0 || end_bit < 0 || end_bit > 64) throw std::invalid_argument("bad sort arguments");
    if (n > 0 && (!keys || !vals)) throw std::invalid_argument("null argument");
    sort_pairs_u64_i32(keys, vals, n, end_bit);
    GP_CHECK(hipDeviceSynchronize());
  });
}

// ---- matrix primitives ----------------------------------------------------------------------
int gp_transpose_csr(const gp_csr* A, gp_csr* AT) {
  return guarded([&] {
    if (!A || !AT) throw std::invalid_argument("null matrix");
    transpose_csr(*A, *AT);
    GP_CHECK(hipDeviceSynchronize());
  });
}

int gp_compact_csr(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, gp_csr* out) {
  return guarded([&] {
    if (!A || !out) throw std::invalid_argument("null matrix");
    DVec<int32_t> ro(A->rows ? A->rows : 1), co(A->cols ? A->cols : 1);
    int32_t m_new = A->rows, n_new = A->cols;
    const int32_t* rop = nullptr;
    if (keep_row) {
      m_new = build_old_to_new(keep_row, A->rows, ro.data());
      rop = ro.data();
    }
    if (keep_col) n_new = build_old_to_new(keep_col, A->cols, co.data());
    compact_csr(*A, keep_row, keep_col, rop, m_new, n_new, *out);
    GP_CHECK(hipDeviceSynchronize());
  });
}

void gp_csr_free(gp_csr* A) {
  if (!A) return;
  csr_pool_free(*A);
}

// ---- device memory helpers ------------------------------------------------------------------
int gp_device_count(int32_t* count) {
  return guarded([&] {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      c = 0;
    }
    if (count) *count = c;
  });
}

int gp_malloc(void** ptr, size_t bytes) {
  return guarded([&] {
    if (!ptr) throw std::invalid_argument("null pointer");
    *ptr = nullptr;
    GP_CHECK(hipMalloc(ptr, bytes ? bytes : 1));
  });
}

int gp_free(void* ptr) {
  return guarded([&] {
    if (ptr) GP_CHECK(hipFree(ptr));
  });
}

int gp_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    if (bytes) GP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}

int gp_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    if (bytes) GP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}

int gp_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    if (bytes) GP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  });
}

int gp_synchronize(void) {
  return guarded([&] { GP_CHECK(hipDeviceSynchronize()); });
}

int gp_lp_upload(int32_t m, int32_t n, gp_offset nnz, const gp_offset* row_ptr, const int32_t* col_idx,
                 const double* vals, const double* c, const double* AL, const double* AU, const double* l,
                 const double* u, double obj_constant, gp_device_lp* out) {
  return guarded([&] {
    if (!out || m < 0 || n < 0 || nnz < 0) throw std::invalid_argument("bad LP arguments");
    std::memset(out, 0, sizeof(*out));
    out->A.rows = m;
    out->A.cols = n;
    out->A.nnz = nnz;
    out->A.row_ptr = pool_alloc<gp_offset>(static_cast<size_t>(m) + 1);
    out->A.col_idx = pool_alloc<int32_t>(nnz);
    out->A.vals = pool_alloc<double>(nnz);
    h2d(out->A.row_ptr, row_ptr, static_cast<size_t>(m) + 1);
    h2d(out->A.col_idx, col_idx, static_cast<size_t>(nnz));
    h2d(out->A.vals, vals, static_cast<size_t>(nnz));
    out->c = pool_alloc<double>(n);
    out->l = pool_alloc<double>(n);
    out->u = pool_alloc<double>(n);
    out->AL = pool_alloc<double>(m);
    out->AU = pool_alloc<double>(m);
    h2d(out->c, c, n);
    h2d(out->l, l, n);
    h2d(out->u, u, n);
    h2d(out->AL, AL, m);
    h2d(out->AU, AU, m);
    out->obj_constant = obj_constant;
    transpose_csr(out->A, out->AT);
    GP_CHECK(hipDeviceSynchronize());
  });
}

void gp_lp_free(gp_device_lp* lp) {
  if (!lp) return;
  csr_pool_free(lp->A);
  csr_pool_free(lp->AT);
  pool_free(lp->c);
  pool_free(lp->l);
  pool_free(lp->u);
  pool_free(lp->AL);
  pool_free(lp->AU);
  std::memset(lp, 0, sizeof(*lp));
}

}  // extern "C"
