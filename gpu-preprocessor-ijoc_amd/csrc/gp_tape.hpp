// gp_tape.hpp -- device-resident postsolve tape (PostsolveTapeGpu layout:
// types, index_starts, value_starts, dual_modes, indices, vals) built by
// count -> scan -> pack kernels, never through host vectors.
#pragma once

#include "gp_common.hpp"

namespace gp {

struct DeviceTape {
  DVec<int32_t> types;
  DVec<gp_offset> istarts;  // count + 1
  DVec<gp_offset> vstarts;  // count + 1
  DVec<uint8_t> modes;
  DVec<int32_t> idx;
  DVec<double> vals;
  int64_t count = 0, icount = 0, vcount = 0;

  DeviceTape();
  void clear();
  // Make room for R more records, I more indices and V more values.
  void reserve_more(int64_t R, int64_t I, int64_t V);
  void commit(int64_t R, int64_t I, int64_t V);
};

// Per-item record/index/value counts, exclusive-scanned in place.
struct AppendCounts {
  DVec<int64_t> r, i, v;
  int64_t R = 0, I = 0, V = 0;
  explicit AppendCounts(int64_t n);
  void scan(int64_t n);
};

// Device view used by pack kernels: base offsets of the new records.
struct TapeView {
  int32_t* types;
  gp_offset* istarts;
  gp_offset* vstarts;
  uint8_t* modes;
  int32_t* idx;
  double* vals;
  int64_t rbase, ibase, vbase;
};
TapeView tape_view(DeviceTape& t);

// Append `src` to `dst`, mapping local indices through the physical->original
// maps (globalisation, by record type).
void tape_append_global(DeviceTape& dst, const DeviceTape& src, const int32_t* rows, int32_t m,
                        const int32_t* cols, int32_t n);

// Host copy of a tape (export).
void tape_copy_to_host(const DeviceTape& t, int32_t* types, gp_offset* istarts, gp_offset* vstarts,
                       uint8_t* modes, int32_t* idx, double* vals);

inline DeviceTape* plan_tape(const gp_plan& plan) { return static_cast<DeviceTape*>(plan.tape); }
// Per-plan caches of the rules, hung off gp_plan::rule_cache.  Each rule
// source file owns the type of its slot.
struct RuleCaches {
  void* sc = nullptr;        // singleton-column row activities (gp_rules2.hip)
  void* col_max = nullptr;   // column max |a| (gp_rules.hip)
  EntryBounds* store = nullptr;      // the scheduler's entry-bounds store (not owned)
  EntryBounds* own_store = nullptr;  // this plan's own store when the scheduler gives none
};
RuleCaches& rule_caches(gp_plan& plan);  // created on first use
// The entry-bounds store the plan's row folds use (the scheduler's, else the
// plan's own, created on first use).
EntryBounds* plan_store(gp_plan& plan);
void rule_cache_free(void* cache);
void sc_cache_free(void* cache);         // gp_rules2.hip
void col_max_cache_free(void* cache);    // gp_rules.hip

// Rules (gp_rules.hip / gp_rules2.hip).
void rule_fixed_variables(gp_plan& plan, const gp_device_lp& lp, const gp_params& p);
void rule_empty_columns(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void rule_empty_rows(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void rule_singleton_rows(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void rule_redundant_rows(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void rule_bound_strengthening(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void rule_singleton_columns(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void rule_dual_fix(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void rule_doubleton_equations(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void rule_parallel_rows(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void rule_parallel_columns(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);
void parallel_row_hashes(const gp_csr& A, const uint8_t* keep_row, const uint8_t* keep_col, double tol,
                         uint64_t* hash);
void apply_rule(int32_t rule, gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p);

// FixedVariable records + objective for the columns in `fixed` and not in
// `excluded` (rows limited to keep_row); returns the objective delta sum
// c_j * v_j: Pairwise = the fixed-order det_sum, Extended = the reference's
// sequential x87 long-double sum in column order (bound strengthening,
// fixed_variable_tape.cu:305-325), Skip = 0.
enum class FixedObjective { Pairwise, Extended, Skip };
double fixed_variable_records(DeviceTape* tape, const uint8_t* fixed, const double* fval, const uint8_t* keep_row,
                              const double* c, const gp_csr& AT, int32_t n, bool with_rows,
                              FixedObjective objective = FixedObjective::Pairwise, const uint8_t* excluded = nullptr);

// Read-only view of the opaque recovery record (defined in gp_presolve.hip).
struct RecordView {
  int32_t m0, n0, m1, n1;
  const int32_t* rows;
  const int32_t* cols;
  const int32_t* fixed_idx;
  const double* fixed_val;
  int64_t fixed_count;
  const DeviceTape* tape;
  const void* fold;  // FoldState of the folding stage (applied), or null
};
RecordView record_view(const void* record);

}  // namespace gp
