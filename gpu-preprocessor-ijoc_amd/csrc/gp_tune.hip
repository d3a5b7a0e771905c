// gp_tune.hip -- launch-schedule autotuner.
//
// The grouped sweeps (bound candidates, column max, live statistics) run G
// lanes per row, i.e. 64 / G rows per wavefront, the compaction tiles hold
// 1,024 * P entries, and the row folds stage R rows per workgroup.  Every
// choice computes the same results (the grouped sweeps are order-free, a row
// fold's order is its CSR order, a compaction's output is its input order),
// so the schedule can be picked by measurement.  gp_autotune presolves the given
// LP once per candidate value, reads each kernel family's device time from its
// HIP events, keeps the fastest value per knob, and finally keeps the tuned
// schedule only if a whole presolve is faster with it than with the
// mean-row-length rule.
#include <algorithm>
#include <limits>

#include "gp_common.hpp"

using namespace gp;

namespace {

bool lanes_ok(int32_t v) { return v == 0 || v == 2 || v == 4 || v == 8 || v == 16 || v == 64; }
bool rows_ok(int32_t v) { return v == 0 || v == 128 || v == 256 || v == 512; }
bool per_lane_ok(int32_t v) { return v == 0 || v == 4 || v == 8 || v == 16; }

gp_tuning to_c(const Tuning& t) {
  return gp_tuning{t.lanes_candidates, t.lanes_col_max, t.compact_per_lane, t.lanes_live_stats, t.fold_rows};
}

Tuning from_c(const gp_tuning& t) {
  Tuning r;
  r.lanes_candidates = t.lanes_candidates;
  r.lanes_col_max = t.lanes_col_max;
  r.compact_per_lane = t.compact_per_lane;
  r.lanes_live_stats = t.lanes_live_stats;
  r.fold_rows = t.fold_rows;
  return r;
}

// one timed presolve: per-family device ms and the presolve core wall time
struct Sample {
  double family[GP_KT_COUNT] = {};
  double wall = 0.0;
};

Sample run_once(const gp_device_lp* lp, const gp_params& p) {
  gp_summary s;
  const int rc = gp_preprocess(lp, &p, &s);
  if (rc != GP_OK) throw Error(rc, std::string("autotune presolve: ") + gp_last_error());
  Sample out;
  for (int k = 0; k < GP_KT_COUNT; ++k) out.family[k] = s.kernel_ms[k];
  out.wall = s.elapsed_seconds * 1e3;
  gp_free_summary(&s);
  return out;
}

Sample best_of(const gp_device_lp* lp, const gp_params& p, int reps) {
  Sample best;
  for (int k = 0; k < GP_KT_COUNT; ++k) best.family[k] = std::numeric_limits<double>::infinity();
  best.wall = std::numeric_limits<double>::infinity();
  for (int r = 0; r < reps; ++r) {
    const Sample s = run_once(lp, p);
    for (int k = 0; k < GP_KT_COUNT; ++k) best.family[k] = std::min(best.family[k], s.family[k]);
    best.wall = std::min(best.wall, s.wall);
  }
  return best;
}

}  // namespace

extern "C" {

int gp_get_tuning(gp_tuning* t) {
  return guarded([&] {
    if (!t) throw std::invalid_argument("null tuning");
    *t = to_c(tuning());
  });
}

int gp_set_tuning(const gp_tuning* t) {
  return guarded([&] {
    if (!t) throw std::invalid_argument("null tuning");
    if (!lanes_ok(t->lanes_candidates) || !lanes_ok(t->lanes_col_max) || !lanes_ok(t->lanes_live_stats))
      throw std::invalid_argument("lanes per row must be 0, 2, 4, 8, 16 or 64");
    if (!per_lane_ok(t->compact_per_lane)) throw std::invalid_argument("compaction entries per lane must be 0, 4, 8 or 16");
    if (!rows_ok(t->fold_rows)) throw std::invalid_argument("fold rows must be 0, 128, 256 or 512");
    tuning() = from_c(*t);
  });
}

int gp_autotune(const gp_device_lp* lp, const gp_params* params, int32_t reps, gp_tuning* best, double* tuned_ms,
                double* default_ms) {
  const Tuning saved = tuning();
  const int rc = guarded([&] {
    if (!lp || !params || !best) throw std::invalid_argument("null argument");
    if (reps < 1) reps = 1;
    gp_params p = *params;
    p.kernel_timing = 1;
    p.collect_trace = 0;
    static const int32_t kLanes[] = {2, 4, 8, 16, 64};
    static const int32_t kRows[] = {128, 256, 512};
    static const int32_t kPerLane[] = {4, 8, 16};
    // knob -> families it schedules
    double best_ms[5];
    int32_t best_val[5] = {0, 0, 0, 0, 0};
    std::fill(best_ms, best_ms + 5, std::numeric_limits<double>::infinity());
    auto consider = [&](int knob, int32_t value, double ms) {
      if (ms > 0.0 && ms < best_ms[knob]) {
        best_ms[knob] = ms;
        best_val[knob] = value;
      }
    };
    for (int i = 0; i < 5; ++i) {
      Tuning t;
      t.lanes_candidates = t.lanes_live_stats = kLanes[i];
      // column max: grouped lanes in the first passes, entry-parallel tiles (0) in the last
      const int32_t col_max = i + 1 < 5 ? kLanes[i] : 0;
      t.lanes_col_max = col_max;
      t.fold_rows = i < 3 ? kRows[i] : 0;
      t.compact_per_lane = i < 3 ? kPerLane[i] : 0;
      tuning() = t;
      const Sample s = best_of(lp, p, reps);
      consider(0, kLanes[i], s.family[GP_KT_BOUND_CANDIDATES]);
      consider(1, col_max, s.family[GP_KT_COL_MAX]);
      consider(3, kLanes[i], s.family[GP_KT_LIVE_STATS]);
      if (i < 3) {
        consider(2, kPerLane[i], s.family[GP_KT_COMPACT]);
        consider(4, kRows[i], s.family[GP_KT_ACTIVITY] + s.family[GP_KT_SC_ACTIVITY] + s.family[GP_KT_REDUNDANT]);
      }
    }
    Tuning chosen;
    chosen.lanes_candidates = best_val[0];
    chosen.lanes_col_max = best_val[1];
    chosen.compact_per_lane = best_val[2];
    chosen.lanes_live_stats = best_val[3];
    chosen.fold_rows = best_val[4];
    gp_params q = *params;
    q.kernel_timing = 0;
    q.collect_trace = 0;
    tuning() = Tuning{};
    const double d = best_of(lp, q, reps).wall;
    tuning() = chosen;
    const double c = best_of(lp, q, reps).wall;
    if (c > d) chosen = Tuning{};
    *best = to_c(chosen);
    if (tuned_ms) *tuned_ms = std::min(c, d);
    if (default_ms) *default_ms = d;
  });
  tuning() = saved;
  return rc;
}

}  // extern "C"
