// gp_rules2.hip -- HIP kernels of the column-side and structural rules:
//   singleton columns    rule_singleton_columns.cu:518-1190, 1562-1827
//   dual fix             rule_dual_fix.cu:186-296, 325-645
//   doubleton equations  rule_doubleton_equations.cu:150-272, 824-1340, 1512-1812 (batch form)
//   parallel rows        rule_parallel_rows.cu:141-466, 648-888
//   parallel columns     rule_parallel_columns.cu:189-612, 636-1127
// Execution-order policy as in gp_rules.hip: row shifts are blocked Neumaier
// sums in CSR order (RowAcc), objective deltas use det_sum, and every work
// unit the reference runs per hash group / per selected pair / per row owner
// runs sequentially inside one thread, so results do not depend on scheduling.
#include <climits>

#include "gp_tape.hpp"

namespace gp {

namespace {
inline bool tape_on(const gp_plan& plan, const gp_params& p) { return plan.tape && p.record_postsolve_tape; }

__device__ inline int64_t gtid() { return static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; }

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  const int64_t i = gtid();
  if (i < n) p[i] = v;
}

// ============================================================== singleton columns
enum : uint8_t { SC_NONE = 0, SC_SET_ROW_SIDE = 1, SC_TIGHTEN_LHS = 2, SC_TIGHTEN_RHS = 3, SC_ELIM_FREE = 4,
                 SC_ELIM_ABOVE = 5, SC_ELIM_BELOW = 6 };

__global__ void k_sc_count(int32_t n, const uint8_t* keep_col, const uint8_t* smask, int32_t* flags) {
  // existence test only: any kept singleton column
  for (int64_t j = gtid(); j < n; j += static_cast<int64_t>(gridDim.x) * blockDim.x)
    if (keep_col[j] && smask[j]) {
      wave_flag(&flags[0], 1);
      break;
    }
}

// Activity of the support row over live entries, no zero filter
// (_kernel_singleton_row_activity_cache), then the bounds of the rest of the
// row without column j (_singleton_column_activity_bounds_from_cache).  Every
// singleton column of a row computes the same row activity.
// Activity of a row over its live entries, in CSR order (the singleton-column
// row activity cache, rule_singleton_columns.cu:555-754): finite parts of the
// minimum / maximum activity, counts of -inf / +inf terms on each side, and
// whether any term is NaN.
struct ScAct {
  double fmin, fmax;
  int32_t min_neg, min_pos, max_neg, max_pos;
  int32_t live, bad;
};

struct ScActivityFold {
  const uint8_t* keep_row;
  ScAct* out;
  struct State {
    RowAcc mn, mx;
    int32_t c0, c1, c2, c3, live, bad;
  };
  __device__ bool active(int64_t i) const { return keep_row[i] != 0; }
  __device__ bool use(double) const { return true; }
  __device__ void init(State& s) const {
    s.mn = RowAcc{};
    s.mx = RowAcc{};
    s.c0 = s.c1 = s.c2 = s.c3 = s.live = s.bad = 0;
  }
  __device__ void add(State& s, int64_t k, double tmin, double tmax, bool skip) const {
    s.mn.pos(k);
    s.mx.pos(k);
    if (skip) return;
    ++s.live;
    if (isnan(tmin) || isnan(tmax)) {
      s.bad = 1;
      return;
    }
    if (isinf(tmin)) {
      if (tmin < 0.0) ++s.c0;
      else ++s.c1;
    } else {
      s.mn.add(tmin);
    }
    if (isinf(tmax)) {
      if (tmax < 0.0) ++s.c2;
      else ++s.c3;
    } else {
      s.mx.add(tmax);
    }
  }
  __device__ void finish(int64_t i, const State& s, bool active) const {
    ScAct r;
    r.fmin = s.mn.value();
    r.fmax = s.mx.value();
    r.min_neg = s.c0;
    r.min_pos = s.c1;
    r.max_neg = s.c2;
    r.max_pos = s.c3;
    r.live = active ? s.live : 0;
    r.bad = s.bad;
    out[i] = r;
  }
};

// The same row activity for one row on one lane (the sparse-update path).
__device__ ScAct sc_row_activity(const gp_csr& A, int64_t i, Bits keep_col, const double2* lu) {
  ScActivityFold f{nullptr, nullptr};
  ScActivityFold::State st;
  f.init(st);
  const gp_offset p0 = A.row_ptr[i], p1 = A.row_ptr[i + 1];
  for (gp_offset q = p0; q < p1; ++q) {
    const int32_t j = A.col_idx[q];
    const double a = A.vals[q];
    const bool skip = !keep_col[j];
    double lo = 0.0, hi = 0.0;
    if (!skip) {
      const double2 b = lu[j];
      lo = a >= 0.0 ? a * b.x : a * b.y;
      hi = a >= 0.0 ? a * b.y : a * b.x;
    }
    f.add(st, q - p0, lo, hi, skip);
  }
  ScAct r;
  r.fmin = st.mn.value();
  r.fmax = st.mx.value();
  r.min_neg = st.c0;
  r.min_pos = st.c1;
  r.max_neg = st.c2;
  r.max_pos = st.c3;
  r.live = st.live;
  r.bad = st.bad;
  return r;
}

// Row activities kept with the plan between singleton-column calls.  A row's
// entry is recomputed only when one of its columns changed bounds or
// liveness, or the row's own liveness changed, since the last call; the
// values are those a full sweep would produce.
struct ScCache {
  int32_t m = -1, n = -1;
  bool valid = false;
  DVec<ScAct> act;
  DVec<double2> lu;
  DVec<uint8_t> keep_row, keep_col;
};

// marks the rows of columns whose bounds or mask changed; the cached column
// mask is brought up to date on the way
__global__ void k_sc_dirty_cols(gp_csr AT, const double2* lu, const double2* lu_old, const uint8_t* keep_col,
                                uint8_t* keep_col_old, uint8_t* dirty) {
  const int64_t j = gtid();
  if (j >= AT.rows) return;
  const double2 a = lu[j], b = lu_old[j];
  const uint8_t now = keep_col[j];
  const bool mask_same = now == keep_col_old[j];
  if (!mask_same) keep_col_old[j] = now;
  const bool same = __double_as_longlong(a.x) == __double_as_longlong(b.x) &&
                    __double_as_longlong(a.y) == __double_as_longlong(b.y) && mask_same;
  if (same) return;
  for (gp_offset p = AT.row_ptr[j]; p < AT.row_ptr[j + 1]; ++p) dirty[AT.col_idx[p]] = 1;
}

// dirty rows appended to the list with one atomic per wavefront (any order:
// listed rows are refreshed independently)
// (also brings the cached row mask up to date)
__global__ void k_sc_dirty_rows(int32_t m, const uint8_t* keep_row, uint8_t* keep_row_old, uint8_t* dirty,
                                int32_t* list, int32_t* count) {
  const int64_t i = gtid();
  bool d = false;
  if (i < m) {
    const uint8_t now = keep_row[i];
    d = dirty[i] != 0;
    if (now != keep_row_old[i]) {
      keep_row_old[i] = now;
      d = true;
    }
  }
  const unsigned long long b = __ballot(d);
  if (!b) return;
  const int lane = static_cast<int>(__lane_id());
  const int leader = __ffsll(static_cast<long long>(b)) - 1;
  int32_t base = 0;
  if (lane == leader) base = atomicAdd(count, static_cast<int32_t>(__popcll(b)));
  base = __shfl(base, leader, 64);
  if (d) list[base + __popcll(b & ((1ull << lane) - 1ull))] = static_cast<int32_t>(i);
}

__global__ void k_sc_refresh_rows(gp_csr A, const int32_t* list, int32_t count, const uint8_t* keep_row,
                                  Bits keep_col, const double2* lu, ScAct* act) {
  const int64_t k = gtid();
  if (k >= count) return;
  const int32_t i = list[k];
  ScAct r = sc_row_activity(A, i, keep_col, lu);
  if (!keep_row[i]) r.live = 0;
  act[i] = r;
}

// Activity of the support row without column j's own term
// (_singleton_column_activity_bounds_from_cache, :756-816).
__device__ bool sc_rest_bounds(const ScAct& act, double a, double lj, double uj, double* rmin, double* rmax) {
  if (act.bad) return false;
  const double tmin = a >= 0.0 ? a * lj : a * uj;
  const double tmax = a >= 0.0 ? a * uj : a * lj;
  if (isnan(tmin) || isnan(tmax)) return false;
  int32_t c0 = act.min_neg, c1 = act.min_pos, c2 = act.max_neg, c3 = act.max_pos;
  double vmin = act.fmin, vmax = act.fmax;
  if (isinf(tmin)) {
    if (tmin < 0.0) --c0;
    else --c1;
  } else {
    vmin -= tmin;
  }
  if (isinf(tmax)) {
    if (tmax < 0.0) --c2;
    else --c3;
  } else {
    vmax -= tmax;
  }
  if (c0 < 0 || c1 < 0 || c2 < 0 || c3 < 0 || (c0 > 0 && c1 > 0) || (c2 > 0 && c3 > 0)) return false;
  *rmin = c0 > 0 ? -INFINITY : (c1 > 0 ? INFINITY : vmin);
  *rmax = c2 > 0 ? -INFINITY : (c3 > 0 ? INFINITY : vmax);
  return !isnan(*rmin) && !isnan(*rmax);
}

__global__ void k_sc_propose(int32_t n, int32_t m, const ScAct* row_act, const uint8_t* keep_col,
                             const uint8_t* smask, const int32_t* srow, const double* sval, const double* c,
                             const double* l, const double* u, const double* AL, const double* AU, double zt,
                             double tol, uint8_t* action, double* side, int32_t* owner, int32_t* flags,
                             BatchMarks bm) {
  const int64_t j = gtid();
  if (j >= n || !keep_col[j] || !smask[j]) return;
  const int32_t row = srow[j];
  if (row < 0 || row >= m) return;
  const double a = sval[j];
  if (fabs(a) <= zt) return;
  const double cj = c[j], lj = l[j], uj = u[j];
  // A deleted row has live = 0 in the cache (the sweep and the row refresh
  // both write it so), so this also returns for every row with keep_row 0.
  const ScAct act = row_act[row];
  if (act.live <= 1) return;
  double rmin = 0.0, rmax = 0.0;
  const bool rest_ok = sc_rest_bounds(act, a, lj, uj, &rmin, &rmax);
  // The row sides matter only to the unboundedness test, which needs a cost
  // against an infinite bound, and to a column whose rest of row is bounded:
  // every other column returns here without gathering them.
  if (!rest_ok && !(cj > zt && !isfinite(lj)) && !(cj < -zt && !isfinite(uj))) return;
  const double lhs = AL[row], rhs = AU[row];
  if ((cj > zt && a > zt && !isfinite(lhs) && !isfinite(lj)) ||
      (cj > zt && a < -zt && !isfinite(rhs) && !isfinite(lj)) ||
      (cj < -zt && a < -zt && !isfinite(lhs) && !isfinite(uj)) ||
      (cj < -zt && a > zt && !isfinite(rhs) && !isfinite(uj))) {
    wave_flag(&flags[1], 1);
    bm.col(j, 2);
    return;
  }
  if (!rest_ok) return;
  uint8_t ac = SC_NONE;
  double sd = 0.0;
  if (isfinite(lhs) && isfinite(rhs) && fabs(lhs - rhs) <= tol) {
    const double x1 = (rhs - rmin) / a, x2 = (rhs - rmax) / a;
    const double ilb = fmin(x1, x2), iub = fmax(x1, x2);
    const bool above = !isfinite(uj) || iub <= uj + tol;
    const bool below = !isfinite(lj) || ilb >= lj - tol;
    if (!(above || below)) return;
    ac = above && below ? SC_ELIM_FREE : (above ? SC_ELIM_ABOVE : SC_ELIM_BELOW);
    sd = rhs;
  } else {
    bool above = true, below = true;
    if (isfinite(uj)) {
      double iu = INFINITY;
      if (a > 0.0 && isfinite(rhs)) iu = (rhs - rmin) / a;
      else if (a < 0.0 && isfinite(lhs)) iu = (lhs - rmax) / a;
      above = iu <= uj + tol;
    }
    if (isfinite(lj)) {
      double il = -INFINITY;
      if (a > 0.0 && isfinite(lhs)) il = (lhs - rmax) / a;
      else if (a < 0.0 && isfinite(rhs)) il = (rhs - rmin) / a;
      below = il >= lj - tol;
    }
    if (above && below) {
      if ((cj > zt && a > 0.0) || (cj < -zt && a < 0.0)) sd = lhs;
      else if ((cj > zt && a < 0.0) || (cj < -zt && a > 0.0)) sd = rhs;
      else sd = isfinite(lhs) ? lhs : rhs;
      ac = isfinite(sd) ? SC_SET_ROW_SIDE : SC_NONE;
    } else if (((cj < -zt && a > 0.0 && above) || (cj > zt && a < 0.0 && below)) && isfinite(rhs)) {
      sd = rhs;
      ac = SC_TIGHTEN_LHS;
    } else if (((cj > zt && a > 0.0 && below) || (cj < -zt && a < 0.0 && above)) && isfinite(lhs)) {
      sd = lhs;
      ac = SC_TIGHTEN_RHS;
    }
    if (ac == SC_NONE) return;
  }
  action[j] = ac;
  side[j] = sd;
  atomicMin(&owner[row], ac == SC_ELIM_FREE ? static_cast<int32_t>(j - INT_MAX) : static_cast<int32_t>(j));
  wave_flag(&flags[4], 1);
}

__global__ void k_sc_clear(int32_t m, int32_t n, uint8_t* row_delete, uint8_t* lower_change, uint8_t* upper_change,
                           uint8_t* col_delete, double* contrib) {
  const int64_t i = gtid();
  if (i < m) {
    row_delete[i] = 0;
    lower_change[i] = 0;
    upper_change[i] = 0;
  }
  if (i < n) {
    col_delete[i] = 0;
    contrib[i] = 0.0;
  }
}

__global__ void k_sc_stage(int32_t m, int32_t n, int32_t* owner, const uint8_t* action, const double* side,
                           const double* sval, const double* c, const double* l, const double* u, const double* AL,
                           const double* AU, double tol, uint8_t* row_delete, uint8_t* lower_change,
                           uint8_t* upper_change, uint8_t* col_delete, double* AL_new, double* AU_new,
                           double* contrib, int32_t* flags) {
  const int64_t row = gtid();
  if (row >= m) return;
  const int32_t o = owner[row];
  const int32_t j = o < 0 ? o + INT_MAX : o;
  if (j < 0 || j >= n) return;
  owner[row] = j;
  const double lhs = AL[row], rhs = AU[row];
  double nl = lhs, nu = rhs;
  const uint8_t ac = action[j];
  if (ac == SC_ELIM_FREE || ac == SC_ELIM_ABOVE || ac == SC_ELIM_BELOW) {
    const double a = sval[j];
    contrib[j] = c[j] * rhs / a;
    if (ac == SC_ELIM_FREE) {
      row_delete[row] = 1;
    } else {
      const bool above = ac == SC_ELIM_ABOVE;
      const double bound = above ? l[j] : u[j];
      const bool keep_lower = above ? a < 0.0 : a > 0.0;
      const double shifted = nu - a * bound;
      if (keep_lower) {
        nl = shifted;
        nu = INFINITY;
      } else {
        nl = -INFINITY;
        nu = shifted;
      }
    }
    col_delete[j] = 1;
    AL_new[row] = nl;
    AU_new[row] = nu;
    wave_flag(&flags[2], 1);
    return;
  }
  if (ac == SC_SET_ROW_SIDE) {
    const double sd = side[j];
    if (!isfinite(lhs) || fabs(lhs - sd) > tol) lower_change[row] = 1;
    if (!isfinite(rhs) || fabs(rhs - sd) > tol) upper_change[row] = 1;
    nl = sd;
    nu = sd;
  } else if (ac == SC_TIGHTEN_LHS) {
    lower_change[row] = 1;
    nl = rhs;
  } else if (ac == SC_TIGHTEN_RHS) {
    upper_change[row] = 1;
    nu = lhs;
  }
  AL_new[row] = nl;
  AU_new[row] = nu;
  wave_flag(&flags[3], 1);
}

// Columns met by an eliminated row (owner column deleted); only their cost
// changes.
__global__ void k_sc_touch(gp_csr A, int32_t n, const int32_t* owner, const uint8_t* col_delete, uint8_t* touched) {
  const int64_t row = gtid();
  if (row >= A.rows) return;
  const int32_t o = owner[row];
  if (o < 0 || o >= n || !col_delete[o]) return;
  for (gp_offset q = A.row_ptr[row]; q < A.row_ptr[row + 1]; ++q) touched[A.col_idx[q]] = 1;
}

// _kernel_apply_singleton_objective_updates_deterministic: plain sequential sum
// in AT order starting from c (untouched columns keep c exactly).  G lanes per
// column load G consecutive entries and form their terms at once; every lane
// then adds the valid terms in entry order, so the additions are those of the
// one-lane loop (invalid entries are skipped, not added as zero).
template <int G>
__global__ void __launch_bounds__(kBlock) k_sc_cost(gp_csr AT, const uint8_t* keep_col, const uint8_t* col_delete,
                                                     const uint8_t* touched, const int32_t* owner, const double* sval,
                                                     double* c) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  const int32_t n = AT.rows;
  for (int64_t j = gtid() / G; j < n; j += groups) {
    if (!touched[j] || !keep_col[j] || col_delete[j]) continue;  // uniform over the group
    double updated = c[j];
    const gp_offset p1 = AT.row_ptr[j + 1];
    for (gp_offset b = AT.row_ptr[j]; b < p1; b += G) {
      const gp_offset q = b + lane;
      double t = 0.0;
      int ok = 0;
      if (q < p1) {
        const int32_t o = owner[AT.col_idx[q]];
        if (o >= 0 && o < n && col_delete[o]) {
          t = -(c[o] * AT.vals[q] / sval[o]);
          ok = 1;
        }
      }
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const double tk = __shfl(t, k, G);
        if (__shfl(ok, k, G)) updated += tk;
      }
    }
    if (lane == 0) c[j] = updated;  // reads only deleted owner columns, which never change here
  }
}
template <int G>
struct ScCostK {
  static constexpr auto kernel = k_sc_cost<G>;
};

__global__ void k_sc_sub_counts(gp_csr A, int32_t n, const int32_t* owner, const uint8_t* col_delete,
                                const uint8_t* keep_col, double zt, int64_t* r, int64_t* ni, int64_t* nv) {
  const int64_t row = gtid();
  if (row >= A.rows) return;
  const int32_t j = owner[row];
  int64_t k = 0;
  bool sel = j >= 0 && j < n && col_delete[j];
  if (sel) {
    double pivot = 0.0;
    for (gp_offset q = A.row_ptr[row]; q < A.row_ptr[row + 1]; ++q) {
      const int32_t col = A.col_idx[q];
      if (col == j) pivot = A.vals[q];
      else if (keep_col[col]) ++k;
    }
    sel = fabs(pivot) > zt;
  }
  r[row] = sel ? 1 : 0;
  ni[row] = sel ? 3 + k : 0;
  nv[row] = sel ? 6 + k : 0;
}

// SingletonColumnSubstitution: idx [col, row, k, support cols], vals [pivot,
// side, l, u, c, row_deleted, support coefficients].
__global__ void k_sc_sub_pack(TapeView out, gp_csr A, const int64_t* r, const int64_t* ni, const int64_t* nv,
                              const int32_t* owner, const uint8_t* keep_col, const uint8_t* row_delete,
                              const double* side, const double* l, const double* u, const double* c) {
  const int64_t row = gtid();
  if (row >= A.rows || r[row + 1] == r[row]) return;
  const int32_t j = owner[row];
  const int64_t rec = out.rbase + r[row];
  const int64_t i0 = out.ibase + ni[row], v0 = out.vbase + nv[row];
  int64_t k = 0;
  double pivot = 0.0;
  for (gp_offset q = A.row_ptr[row]; q < A.row_ptr[row + 1]; ++q) {
    const int32_t col = A.col_idx[q];
    if (col == j) {
      pivot = A.vals[q];
    } else if (keep_col[col]) {
      out.idx[i0 + 3 + k] = col;
      out.vals[v0 + 6 + k] = A.vals[q];
      ++k;
    }
  }
  out.types[rec] = GP_RED_SINGLETON_COLUMN_SUBSTITUTION;
  out.modes[rec] = GP_DUAL_MINIMAL;
  out.idx[i0] = j;
  out.idx[i0 + 1] = static_cast<int32_t>(row);
  out.idx[i0 + 2] = static_cast<int32_t>(k);
  out.vals[v0] = pivot;
  out.vals[v0 + 1] = side[j];
  out.vals[v0 + 2] = l[j];
  out.vals[v0 + 3] = u[j];
  out.vals[v0 + 4] = c[j];
  out.vals[v0 + 5] = row_delete[row] ? 1.0 : 0.0;
  out.istarts[rec + 1] = static_cast<gp_offset>(i0 + 3 + k);
  out.vstarts[rec + 1] = static_cast<gp_offset>(v0 + 6 + k);
}

__global__ void k_rds_counts(int32_t m, const uint8_t* lo, const uint8_t* up, int64_t* r, int64_t* ni, int64_t* nv) {
  const int64_t i = gtid();
  if (i >= m) return;
  const int64_t s = (lo[i] || up[i]) ? 1 : 0;
  r[i] = s;
  ni[i] = s;
  nv[i] = s;
}

__global__ void k_rds_pack(TapeView out, int32_t m, const uint8_t* lo, const uint8_t* up, const int64_t* r) {
  const int64_t i = gtid();
  if (i >= m || !(lo[i] || up[i])) return;
  const int64_t rec = out.rbase + r[i];
  out.types[rec] = GP_RED_ROW_DUAL_SHIFT;
  out.modes[rec] = GP_DUAL_MINIMAL;
  out.idx[out.ibase + r[i]] = static_cast<int32_t>(i);
  out.vals[out.vbase + r[i]] = 0.0;
  out.istarts[rec + 1] = static_cast<gp_offset>(out.ibase + r[i] + 1);
  out.vstarts[rec + 1] = static_cast<gp_offset>(out.vbase + r[i] + 1);
}

__global__ void k_sc_commit(int32_t m, int32_t n, const uint8_t* row_delete, const int32_t* owner,
                            const double* AL_new, const double* AU_new, const uint8_t* col_delete, uint8_t* keep_row,
                            uint8_t* keep_col, double* AL, double* AU) {
  const int64_t q = gtid();
  if (q < m) {
    if (row_delete[q]) keep_row[q] = 0;
    if (owner[q] >= 0 && owner[q] < n) {
      AL[q] = AL_new[q];
      AU[q] = AU_new[q];
    }
  }
  if (q < n && col_delete[q]) keep_col[q] = 0;
}

// ============================================================== dual fix
// Row sides for the dual fix: 4 for a live row, plus 1 when AL is finite and
// 2 when AU is finite.
__global__ void k_df_row_sides(int32_t m, const uint8_t* keep_row, const double* AL, const double* AU,
                               uint8_t* side) {
  const int64_t i = gtid();
  if (i >= m) return;
  side[i] = keep_row[i] ? static_cast<uint8_t>(4 | (isfinite(AL[i]) ? 1 : 0) | (isfinite(AU[i]) ? 2 : 0)) : 0;
}

// Per column, OR over its live entries: bit 0 ("down") when a finite row side
// limits decreasing the column -- AL for a > zt, AU for a < -zt -- and bit 1
// ("up") for the opposite side.  Order-free, so the column's entries are
// reduced entry-parallel (seg_reduce_tiles over AT).
struct DfSidesOp {
  const uint8_t* side;
  double zt;
  using Slot = uint32_t;
  __device__ Slot value(int32_t, int32_t r, double a) const {
    const uint32_t f = side[r];
    if (!(f & 4u)) return 0u;
    if (a > zt) return f & 3u;
    if (a < -zt) return ((f & 2u) >> 1) | ((f & 1u) << 1);
    return 0u;
  }
  __device__ static Slot merge(Slot a, Slot b) { return a | b; }
  __device__ static void combine(Slot* s, Slot v) { atomicOr(s, v); }
};

__global__ void k_df_candidates(gp_csr AT, const uint8_t* keep_col, const uint32_t* sides, const double* c,
                                const double* l, const double* u, double zt, double bt, int32_t maxdeg,
                                const int32_t* live_col_nnz, uint8_t* fixed, uint8_t* inf, uint8_t* ffix,
                                double* val, double* contrib, int32_t* flags, BatchMarks bm) {
  const int64_t j = gtid();
  if (j >= AT.rows) return;
  fixed[j] = 0;
  inf[j] = 0;
  ffix[j] = 0;
  val[j] = 0.0;
  contrib[j] = 0.0;
  if (!keep_col[j]) return;
  const double lj = l[j], uj = u[j];
  if (lj > uj + bt) {
    wave_flag(&flags[0], 1);
    bm.col(j, 1);
    return;
  }
  const gp_offset a0 = AT.row_ptr[j], a1 = AT.row_ptr[j + 1];
  const int64_t degree = live_col_nnz ? static_cast<int64_t>(live_col_nnz[j]) : static_cast<int64_t>(a1 - a0);
  if (fabs(c[j]) > zt && degree > maxdeg) return;
  const bool down = (sides[j] & 1u) != 0, up = (sides[j] & 2u) != 0;
  const double cj = c[j];
  bool fix = false, infinite = false;
  double v = 0.0;
  if (cj > zt) {
    if (!down) {
      if (!isfinite(lj)) {
        wave_flag(&flags[0], 2);
        bm.col(j, 2);
        return;
      }
      v = lj;
      fix = true;
    }
  } else if (cj < -zt) {
    if (!up) {
      if (!isfinite(uj)) {
        wave_flag(&flags[0], 2);
        bm.col(j, 2);
        return;
      }
      v = uj;
      fix = true;
    }
  } else if (fabs(cj) <= zt) {
    if (!down) {
      fix = true;
      if (isfinite(lj)) {
        v = lj;
      } else {
        v = -INFINITY;
        infinite = true;
      }
    } else if (!up) {
      fix = true;
      if (isfinite(uj)) {
        v = uj;
      } else {
        v = INFINITY;
        infinite = true;
      }
    }
  }
  if (fix) {
    fixed[j] = 1;
    inf[j] = infinite ? 1 : 0;
    ffix[j] = infinite ? 0 : 1;
    val[j] = v;
    contrib[j] = infinite ? 0.0 : cj * v;
    wave_flag(&flags[1], 1);
    if (infinite) wave_flag(&flags[2], 1);
  }
}

// The row pass only matters for rows that meet a fixed or infinite-bound
// column, usually few: list_rows_touching lists them and only the listed rows
// take the sequential fold.  Every other row keeps shift 0, no removal.
__global__ void k_df_touch(int32_t n, const uint8_t* ffix, const uint8_t* inf, uint8_t* touched) {
  const int64_t j = gtid();
  if (j < n) touched[j] = (ffix[j] || inf[j]) ? 1 : 0;
}

__global__ void k_df_rows_list(gp_csr A, const int32_t* list, const int32_t* count, const uint8_t* ffix,
                               const uint8_t* inf, const double* val, double* shift, uint8_t* row_remove,
                               uint8_t* keep_after, int32_t* flags) {
  const int32_t total = *count;
  for (int64_t k = gtid(); k < total; k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t i = list[k];
    RowAcc acc;
    uint8_t rem = 0;
    const gp_offset p0 = A.row_ptr[i];
    for (gp_offset q = p0; q < A.row_ptr[i + 1]; ++q) {
      acc.pos(q - p0);
      const int32_t j = A.col_idx[q];
      if (ffix[j]) acc.add(A.vals[q] * val[j]);
      if (inf[j]) rem = 1;
    }
    if (rem) wave_flag(&flags[3], 1);
    shift[i] = acc.value();
    row_remove[i] = rem;
    keep_after[i] = rem ? 0 : 1;
  }
}

__device__ inline int64_t entry_len(const gp_csr& A, int32_t r, const uint8_t* entry) {
  if (!entry) return A.row_ptr[r + 1] - A.row_ptr[r];
  int64_t k = 0;
  for (gp_offset e = A.row_ptr[r]; e < A.row_ptr[r + 1]; ++e) k += entry[A.col_idx[e]] != 0;
  return k;
}

__global__ void k_df_inf_counts(gp_csr A, gp_csr AT, const uint8_t* fixed, const uint8_t* inf,
                                const uint8_t* keep_row, const uint8_t* entry, int64_t* r, int64_t* ni,
                                int64_t* nv) {
  const int64_t j = gtid();
  if (j >= AT.rows) return;
  const bool sel = fixed[j] && inf[j];
  int64_t size = 0;
  if (sel)
    for (gp_offset q = AT.row_ptr[j]; q < AT.row_ptr[j + 1]; ++q) {
      const int32_t row = AT.col_idx[q];
      if (keep_row[row]) size += 1 + entry_len(A, row, entry);
    }
  r[j] = sel ? 1 : 0;
  ni[j] = sel ? 3 + size : 0;
  nv[j] = sel ? 1 + size : 0;
}

// InfiniteBoundDualFix: idx [+-1, col, active rows, (len, cols...)...],
// vals [finite bound, (row side, coefficients...)...].
__global__ void k_df_inf_pack(TapeView out, gp_csr A, gp_csr AT, const uint8_t* fixed, const uint8_t* inf,
                              const uint8_t* keep_row, const uint8_t* entry, const double* val, const double* l,
                              const double* u, const double* AL, const double* AU, const int64_t* r,
                              const int64_t* ni, const int64_t* nv) {
  const int64_t j = gtid();
  if (j >= AT.rows || !(fixed[j] && inf[j])) return;
  const int64_t rec = out.rbase + r[j];
  int64_t ip = out.ibase + ni[j], vp = out.vbase + nv[j];
  const bool positive = val[j] > 0.0;
  out.types[rec] = GP_RED_INFINITE_BOUND_DUAL_FIX;
  out.modes[rec] = GP_DUAL_EXACT;
  out.idx[ip++] = positive ? 1 : -1;
  out.idx[ip++] = static_cast<int32_t>(j);
  const int64_t count_pos = ip++;
  out.vals[vp++] = positive ? l[j] : u[j];
  int32_t active = 0;
  for (gp_offset q = AT.row_ptr[j]; q < AT.row_ptr[j + 1]; ++q) {
    const int32_t row = AT.col_idx[q];
    if (!keep_row[row]) continue;
    ++active;
    const int64_t header = ip++;
    int32_t len = 0;
    out.vals[vp++] = isfinite(AL[row]) ? AL[row] : AU[row];
    for (gp_offset e = A.row_ptr[row]; e < A.row_ptr[row + 1]; ++e) {
      if (entry && !entry[A.col_idx[e]]) continue;
      ++len;
      out.idx[ip++] = A.col_idx[e];
      out.vals[vp++] = A.vals[e];
    }
    out.idx[header] = len;
  }
  out.idx[count_pos] = active;
  out.istarts[rec + 1] = static_cast<gp_offset>(ip);
  out.vstarts[rec + 1] = static_cast<gp_offset>(vp);
}

__global__ void k_df_apply(int32_t m, int32_t n, const uint8_t* fixed, const double* val, double* l, double* u,
                           uint8_t* keep_col, const double* shift, const uint8_t* row_remove, double* AL, double* AU,
                           uint8_t* keep_row) {
  const int64_t k = gtid();
  if (k < n && fixed[k]) {
    l[k] = val[k];
    u[k] = val[k];
    keep_col[k] = 0;
  }
  if (k < m) {
    AL[k] -= shift[k];
    AU[k] -= shift[k];
    if (row_remove[k]) keep_row[k] = 0;
  }
}

// ============================================================== hashing
constexpr uint64_t kFnvOffset = 0xcbf29ce484222325ULL;
constexpr uint64_t kFnvPrime = 0x100000001b3ULL;
__device__ inline uint64_t fnv_mix(uint64_t h, uint64_t x) { return (h ^ x) * kFnvPrime; }
__device__ inline uint64_t dbits(double v) { return static_cast<uint64_t>(__double_as_longlong(v)); }
__device__ inline uint64_t as_u64(int64_t v) { return static_cast<uint64_t>(v); }

// ============================================================== parallel rows
// ---- staged hashing ------------------------------------------------------------------------------
// The row / column hashes are sequential FNV chains over each line's live
// entries, so they cannot be split across lanes.  Two passes instead:
// (1) G lanes per line count the live entries and find the first one
// (order-free); (2) k_line_stream stages the lines' (index, value) pairs
// through LDS with coalesced loads and each lane runs its own line's chain
// from LDS, in order.  The results equal the one-lane-per-line hash.
struct LivePr {  // parallel rows: entry live when its column is live
  const uint8_t* keep;
  __device__ bool operator()(int32_t idx, double) const { return !keep || keep[idx]; }
};
struct LivePc {  // parallel columns: live row and |a| above the tolerance
  const uint8_t* keep;
  double tol;
  __device__ bool operator()(int32_t idx, double v) const { return keep[idx] && fabs(v) > tol; }
};

template <int G, class Live>
__global__ void __launch_bounds__(kBlock) k_first_live(gp_csr M, Live live, int32_t* len, gp_offset* first) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t i = gtid() / G; i < M.rows; i += groups) {
    int32_t c = 0;
    gp_offset f = INT_MAX;
    const gp_offset p1 = M.row_ptr[i + 1];
    for (gp_offset p = M.row_ptr[i] + lane; p < p1; p += G)
      if (live(M.col_idx[p], M.vals[p])) {
        ++c;
        f = p < f ? p : f;
      }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      c += __shfl_xor(c, o, G);
      const gp_offset g = __shfl_xor(f, o, G);
      f = g < f ? g : f;
    }
    if (lane == 0) {
      len[i] = c;
      first[i] = f;
    }
  }
}

template <class Live>
void first_live(const gp_csr& M, Live live, int32_t* len, gp_offset* first) {
  if (M.rows <= 0) return;
  const int lanes = lanes_for_mean(static_cast<double>(M.nnz) / M.rows);
  const unsigned grid = grid_capped(static_cast<int64_t>(M.rows) * lanes, kBlock, 16384);
  switch (lanes) {
    case 2: hipLaunchKernelGGL((k_first_live<2, Live>), dim3(grid), dim3(kBlock), 0, 0, M, live, len, first); break;
    case 4: hipLaunchKernelGGL((k_first_live<4, Live>), dim3(grid), dim3(kBlock), 0, 0, M, live, len, first); break;
    case 8: hipLaunchKernelGGL((k_first_live<8, Live>), dim3(grid), dim3(kBlock), 0, 0, M, live, len, first); break;
    case 16: hipLaunchKernelGGL((k_first_live<16, Live>), dim3(grid), dim3(kBlock), 0, 0, M, live, len, first); break;
    default: hipLaunchKernelGGL((k_first_live<64, Live>), dim3(grid), dim3(kBlock), 0, 0, M, live, len, first); break;
  }
  GP_LAUNCH_CHECK("first live entries");
}

constexpr int kStreamRows = 256;
constexpr int kStreamChunk = 2048;
// Policy P: State; void init(State&, line); void add(State&, index, value); void finish(line, const State&)
template <class P>
__global__ void __launch_bounds__(kStreamRows) k_line_stream(gp_csr M, P pol) {
  __shared__ int32_t s_idx[kStreamChunk];
  __shared__ double s_val[kStreamChunk];
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kStreamRows;
  const int64_t r1 = r0 + kStreamRows < M.rows ? r0 + kStreamRows : M.rows;
  const int64_t line = r0 + threadIdx.x;
  const bool mine = line < r1 && pol.active(line);
  typename P::State st;
  if (mine) pol.init(st, line);
  if (__syncthreads_or(mine)) {
    const gp_offset e0 = M.row_ptr[r0], e1 = M.row_ptr[r1];
    const gp_offset p0 = mine ? M.row_ptr[line] : 0, p1 = mine ? M.row_ptr[line + 1] : 0;
    for (gp_offset c0 = e0; c0 < e1; c0 += kStreamChunk) {
      const gp_offset c1 = e1 - c0 > kStreamChunk ? c0 + kStreamChunk : e1;
      for (gp_offset q = c0 + threadIdx.x; q < c1; q += kStreamRows) {
        s_idx[q - c0] = ld_stream(M.col_idx + q);
        s_val[q - c0] = ld_stream(M.vals + q);
      }
      __syncthreads();
      if (mine) {
        const gp_offset a0 = p0 > c0 ? p0 : c0, a1 = p1 < c1 ? p1 : c1;
        for (gp_offset q = a0; q < a1; ++q) pol.add(st, s_idx[q - c0], s_val[q - c0]);
      }
      __syncthreads();
    }
  }
  if (line < r1) pol.finish(line, st, mine);
}

template <class P>
void launch_line_stream(const gp_csr& M, const P& pol) {
  if (M.rows <= 0) return;
  const int64_t blocks = (static_cast<int64_t>(M.rows) + kStreamRows - 1) / kStreamRows;
  hipLaunchKernelGGL(k_line_stream<P>, dim3(static_cast<unsigned>(blocks)), dim3(kStreamRows), 0, 0, M, pol);
  GP_LAUNCH_CHECK("line stream");
}

// hash of a row over its live columns (rule_parallel_rows.cu:132-192)
struct PrHash {
  const uint8_t* keep_row;
  LivePr live;
  const int32_t* len;
  const gp_offset* first;
  const double* vals;
  double tol;
  int use_xor;
  uint64_t* hash;
  int32_t* ids;
  struct State {
    uint64_t h, agg;
    double pivot;
    bool ok;
  };
  __device__ bool active(int64_t i) const { return keep_row[i] != 0; }
  __device__ void init(State& st, int64_t i) const {
    st.ok = len[i] > 0;
    st.pivot = st.ok ? vals[first[i]] : 0.0;
    if (st.ok && fabs(st.pivot) <= tol) st.ok = false;
    st.h = fnv_mix(kFnvOffset, as_u64(len[i]));
    st.agg = 0;
  }
  __device__ void add(State& st, int32_t col, double v) const {
    if (!st.ok || !live(col, v)) return;
    const uint64_t nb = dbits(v / st.pivot);
    if (use_xor) {
      st.agg ^= fnv_mix(fnv_mix(kFnvOffset, as_u64(col)), nb);
    } else {
      st.h = fnv_mix(st.h, as_u64(col));
      st.h = fnv_mix(st.h, nb);
    }
  }
  __device__ void finish(int64_t i, const State& st, bool mine) const {
    ids[i] = static_cast<int32_t>(i);
    if (!mine) hash[i] = as_u64(i);
    else if (!st.ok) hash[i] = fnv_mix(kFnvOffset, as_u64(i));
    else hash[i] = use_xor ? fnv_mix(st.h, st.agg) : st.h;
  }
};

// hash of a column over its live rows (rule_parallel_columns.cu:189-240)
struct PcHash {
  const uint8_t* keep_col;
  LivePc live;
  const int32_t* len;
  const gp_offset* first;
  const double* vals;
  uint64_t* hash;
  int32_t* ids;
  struct State {
    uint64_t h;
    double pivot;
    bool ok;
  };
  __device__ bool active(int64_t j) const { return keep_col[j] != 0; }
  __device__ void init(State& st, int64_t j) const {
    st.ok = len[j] > 0;
    st.pivot = st.ok ? vals[first[j]] : 0.0;
    st.h = fnv_mix(kFnvOffset, as_u64(len[j]));
  }
  __device__ void add(State& st, int32_t row, double v) const {
    if (!st.ok || !live(row, v)) return;
    st.h = fnv_mix(st.h, as_u64(row));
    st.h = fnv_mix(st.h, dbits(v / st.pivot));
  }
  __device__ void finish(int64_t j, const State& st, bool mine) const {
    ids[j] = static_cast<int32_t>(j);
    if (!mine) hash[j] = as_u64(j);
    else if (!st.ok) hash[j] = fnv_mix(kFnvOffset, as_u64(j));
    else hash[j] = st.h;
  }
};


__device__ bool pr_ratio(const gp_csr& A, int32_t i, int32_t k, const uint8_t* keep_col, double tol, double* out) {
  gp_offset li = A.row_ptr[i], le = A.row_ptr[i + 1], ri = A.row_ptr[k], re = A.row_ptr[k + 1];
  while (keep_col && li < le && !keep_col[A.col_idx[li]]) ++li;
  while (keep_col && ri < re && !keep_col[A.col_idx[ri]]) ++ri;
  if (li == le && ri == re) {
    *out = 1.0;
    return true;
  }
  double ratio = 0.0;
  bool set = false;
  while (li < le && ri < re) {
    if (A.col_idx[li] != A.col_idx[ri]) return false;
    const double ai = A.vals[li], ak = A.vals[ri];
    if (!set) {
      if (fabs(ak) <= tol) return false;
      ratio = ai / ak;
      set = true;
    }
    if (fabs(ai - ratio * ak) > tol) return false;
    ++li;
    while (keep_col && li < le && !keep_col[A.col_idx[li]]) ++li;
    ++ri;
    while (keep_col && ri < re && !keep_col[A.col_idx[ri]]) ++ri;
  }
  *out = ratio;
  return set && li == le && ri == re;
}

__device__ inline void scaled_interval(double lo, double hi, double ratio, double* sl, double* su) {
  const double rl = lo * ratio, ru = hi * ratio;
  if (ratio >= 0.0) {
    *sl = rl;
    *su = ru;
  } else {
    *sl = ru;
    *su = rl;
  }
}

struct PrRecs {
  int32_t* step;
  int32_t* kept;
  double* vals;  // 5 per deleted row: ratio, kept AL, kept AU, deleted AL, deleted AU
};

__global__ void k_pr_groups(int32_t m, const uint64_t* hash, const int32_t* ids, gp_csr A, const uint8_t* keep_row,
                            const uint8_t* keep_col, double* AL, double* AU, double ctol, double itol,
                            uint8_t* row_delete, PrRecs rec, int32_t* flags, BatchMarks bm) {
  const int64_t gs = gtid();
  if (gs >= m) return;
  if (gs > 0 && hash[gs - 1] == hash[gs]) return;
  int64_t ge = gs;
  while (ge + 1 < m && hash[ge + 1] == hash[gs]) ++ge;
  if (ge == gs) return;
  int32_t seed = -1;
  bool seed_eq = false;
  for (int64_t a = gs; a <= ge; ++a) {
    const int32_t row = ids[a];
    if (!keep_row[row]) continue;
    const bool eq = isfinite(AL[row]) && isfinite(AU[row]) && fabs(AL[row] - AU[row]) <= itol;
    if (seed < 0 || (eq && !seed_eq)) {
      seed = row;
      seed_eq = eq;
    }
  }
  if (seed < 0) return;
  int32_t rep = seed;
  double rsl = AL[seed], rsu = AU[seed];
  for (int64_t a = gs; a <= ge; ++a) {
    const int32_t row = ids[a];
    if (row == seed || !keep_row[row]) continue;
    double r = 0.0;
    if (!pr_ratio(A, seed, row, keep_col, ctol, &r)) continue;
    double sl, su;
    scaled_interval(AL[row], AU[row], r, &sl, &su);
    if (rsl <= sl + itol && su <= rsu + itol) {
      rep = row;
      rsl = sl;
      rsu = su;
    }
  }
  double rl = AL[rep], ru = AU[rep];
  bool found = false;
  int32_t step = 0;
  for (int64_t a = gs; a <= ge; ++a) {
    const int32_t row = ids[a];
    if (row == rep || !keep_row[row]) continue;
    double r = 0.0;
    if (!pr_ratio(A, rep, row, keep_col, ctol, &r)) continue;
    found = true;
    double sl, su;
    scaled_interval(AL[row], AU[row], r, &sl, &su);
    if (cmax(rl, sl) > cmin(ru, su) + itol) {
      wave_flag(&flags[0], 1);
      bm.row(row, 1);
      return;
    }
    ++step;
    rec.step[row] = step;
    rec.kept[row] = rep;
    rec.vals[5 * static_cast<int64_t>(row) + 0] = r;
    rec.vals[5 * static_cast<int64_t>(row) + 1] = rl;
    rec.vals[5 * static_cast<int64_t>(row) + 2] = ru;
    rec.vals[5 * static_cast<int64_t>(row) + 3] = AL[row];
    rec.vals[5 * static_cast<int64_t>(row) + 4] = AU[row];
    rl = cmax(rl, sl);
    ru = cmin(ru, su);
    row_delete[row] = 1;
  }
  if (found) {
    AL[rep] = rl;
    AU[rep] = ru;
  }
}

__global__ void k_pr_apply(int32_t m, const uint8_t* row_delete, uint8_t* keep_row, int32_t* flags) {
  const int64_t i = gtid();
  if (i < m && row_delete[i]) {
    keep_row[i] = 0;
    wave_count(&flags[1]);
  }
}

// Record order keys: (step, deleted row).
__global__ void k_pr_keys(int32_t m, const uint8_t* row_delete, const int32_t* o2n, const int32_t* step,
                          uint64_t* keys, int32_t* rows) {
  const int64_t i = gtid();
  if (i >= m || !row_delete[i]) return;
  const int32_t t = o2n[i];
  keys[t] = (static_cast<uint64_t>(static_cast<uint32_t>(step[i])) << 32) | static_cast<uint32_t>(i);
  rows[t] = static_cast<int32_t>(i);
}

__global__ void k_pr_pack(TapeView out, int32_t count, const int32_t* rows, PrRecs rec) {
  const int64_t t = gtid();
  if (t >= count) return;
  const int32_t row = rows[t];
  const int64_t r = out.rbase + t;
  out.types[r] = GP_RED_PARALLEL_ROW;
  out.modes[r] = GP_DUAL_EXACT;
  out.idx[out.ibase + 2 * t] = rec.kept[row];
  out.idx[out.ibase + 2 * t + 1] = row;
  for (int k = 0; k < 5; ++k) out.vals[out.vbase + 5 * t + k] = rec.vals[5 * static_cast<int64_t>(row) + k];
  out.istarts[r + 1] = static_cast<gp_offset>(out.ibase + 2 * (t + 1));
  out.vstarts[r + 1] = static_cast<gp_offset>(out.vbase + 5 * (t + 1));
}

// ============================================================== parallel columns
__device__ inline bool pc_live(const gp_csr& AT, gp_offset q, const uint8_t* keep_row, double zt) {
  return keep_row[AT.col_idx[q]] && fabs(AT.vals[q]) > zt;
}
__device__ inline gp_offset pc_next(const gp_csr& AT, gp_offset q, gp_offset stop, const uint8_t* keep_row, double zt) {
  while (q < stop && !pc_live(AT, q, keep_row, zt)) ++q;
  return q;
}


__device__ bool pc_ratio(const gp_csr& AT, int32_t j, int32_t k, const uint8_t* keep_row, double zt, double ctol,
                         double* out) {
  gp_offset pj = AT.row_ptr[j], pk = AT.row_ptr[k];
  const gp_offset sj = AT.row_ptr[j + 1], sk = AT.row_ptr[k + 1];
  double ratio = 0.0;
  bool set = false;
  for (;;) {
    pj = pc_next(AT, pj, sj, keep_row, zt);
    pk = pc_next(AT, pk, sk, keep_row, zt);
    if (pj >= sj || pk >= sk) break;
    if (AT.col_idx[pj] != AT.col_idx[pk]) return false;
    const double aj = AT.vals[pj], ak = AT.vals[pk];
    if (!set) {
      if (fabs(aj) <= zt) return false;
      ratio = ak / aj;
      if (fabs(ratio) <= zt || !isfinite(ratio)) return false;
      set = true;
    }
    const double expected = ratio * aj;
    const double scale = fmax(fabs(ak), fabs(expected));
    if (!isfinite(expected) || fabs(ak - expected) > ctol * scale) return false;
    ++pj;
    ++pk;
  }
  pj = pc_next(AT, pj, sj, keep_row, zt);
  pk = pc_next(AT, pk, sk, keep_row, zt);
  if (pj < sj || pk < sk || !set) return false;
  *out = ratio;
  return true;
}

__device__ inline bool out_of_guard(double ratio, double max_abs) {
  const double ar = fabs(ratio);
  return max_abs < 1.0 || ar > max_abs || ar * max_abs < 1.0;
}

__global__ void k_pc_guard(int32_t n, const uint64_t* hash, const int32_t* ids, gp_csr AT, const uint8_t* keep_col,
                           const uint8_t* keep_row, const double* c, double zt, double ctol, double otol,
                           double max_abs, int32_t max_group, int32_t* flags) {
  const int64_t gs = gtid();
  if (gs >= n) return;
  if (gs > 0 && hash[gs - 1] == hash[gs]) return;
  int64_t ge = gs;
  while (ge + 1 < n && hash[ge + 1] == hash[gs]) ++ge;
  const int64_t size = ge - gs + 1;
  if (size < 2 || size > max_group) return;
  for (int64_t a = gs; a < ge; ++a) {
    const int32_t j = ids[a];
    if (!keep_col[j]) continue;
    for (int64_t b = a + 1; b <= ge; ++b) {
      const int32_t k = ids[b];
      if (!keep_col[k]) continue;
      double r = 0.0;
      if (!pc_ratio(AT, j, k, keep_row, zt, ctol, &r)) continue;
      if (fabs(c[k] - r * c[j]) > otol) continue;
      if (out_of_guard(r, max_abs)) {
        wave_count(&flags[4]);
        return;
      }
    }
  }
}

__device__ inline double merged_lower(double tl, double sl, double su, double r) {
  if (r > 0.0) {
    if (isfinite(tl) && isfinite(sl)) return tl + r * sl;
  } else if (isfinite(tl) && isfinite(su)) {
    return tl + r * su;
  }
  return -INFINITY;
}
__device__ inline double merged_upper(double tu, double sl, double su, double r) {
  if (r > 0.0) {
    if (isfinite(tu) && isfinite(su)) return tu + r * su;
  } else if (isfinite(tu) && isfinite(sl)) {
    return tu + r * sl;
  }
  return INFINITY;
}

struct PcOut {
  uint8_t* col_delete;
  uint8_t* fixed;
  double* fval;
  int32_t* merge_to;
  double* mrec;  // 5 per merged column: ratio, source l, u, target l, u
};

__global__ void k_pc_groups(int32_t n, const uint64_t* hash, const int32_t* ids, gp_csr AT, const uint8_t* keep_col,
                            const uint8_t* keep_row, const double* c, double* l, double* u, double zt, double ctol,
                            double otol, double active_max, PcOut o, int32_t* flags, BatchMarks bm) {
  const int64_t gs = gtid();
  if (gs >= n) return;
  if (gs > 0 && hash[gs - 1] == hash[gs]) return;
  int64_t ge = gs;
  while (ge + 1 < n && hash[ge + 1] == hash[gs]) ++ge;
  for (int64_t a = gs; a < ge; ++a) {
    const int32_t j = ids[a];
    if (!keep_col[j] || o.col_delete[j] || o.fixed[j]) continue;
    for (int64_t b = a + 1; b <= ge; ++b) {
      const int32_t k = ids[b];
      if (!keep_col[k] || o.col_delete[k] || o.fixed[k]) continue;
      if (isfinite(l[j]) && isfinite(u[j]) && isfinite(l[k]) && isfinite(u[k])) {
        const gp_offset fj = pc_next(AT, AT.row_ptr[j], AT.row_ptr[j + 1], keep_row, zt);
        const gp_offset fk = pc_next(AT, AT.row_ptr[k], AT.row_ptr[k + 1], keep_row, zt);
        if (fj < AT.row_ptr[j + 1] && fk < AT.row_ptr[k + 1]) {
          const double fr = AT.vals[fk] / AT.vals[fj];
          if (fabs(c[k] - fr * c[j]) > otol) continue;
        }
      }
      double r = 0.0;
      if (!pc_ratio(AT, j, k, keep_row, zt, ctol, &r)) continue;
      if (isfinite(active_max) && out_of_guard(r, active_max)) continue;
      const double gap = c[k] - r * c[j];
      const double tl = l[j], tu = u[j], sl = l[k], su = u[k];
      if (fabs(gap) <= otol) {
        o.merge_to[k] = j;
        double* mr = o.mrec + 5 * static_cast<int64_t>(k);
        mr[0] = r;
        mr[1] = sl;
        mr[2] = su;
        mr[3] = tl;
        mr[4] = tu;
        l[j] = merged_lower(tl, sl, su, r);
        u[j] = merged_upper(tu, sl, su, r);
        o.col_delete[k] = 1;
        wave_flag(&flags[1], 1);
        continue;
      }
      bool xk_lo = false, xk_up = false, xj_lo = false, xj_up = false;
      if (gap > otol) {
        if (r > 0.0) {
          xk_lo = !isfinite(tu);
          xj_up = !isfinite(sl);
        } else {
          xk_lo = !isfinite(tl);
          xj_lo = !isfinite(sl);
        }
      } else {
        if (r > 0.0) {
          xk_up = !isfinite(tl);
          xj_lo = !isfinite(su);
        } else {
          xk_up = !isfinite(tu);
          xj_up = !isfinite(su);
        }
      }
      if (xk_lo || xk_up) {
        const double v = xk_lo ? sl : su;
        if (!isfinite(v)) {
          wave_flag(&flags[0], 1);
          bm.col(k, 2);
          return;
        }
        o.fixed[k] = 1;
        o.fval[k] = v;
        wave_flag(&flags[2], 1);
        continue;
      }
      if (xj_lo || xj_up) {
        const double v = xj_lo ? tl : tu;
        if (!isfinite(v)) {
          wave_flag(&flags[0], 1);
          bm.col(j, 2);
          return;
        }
        o.fixed[j] = 1;
        o.fval[j] = v;
        wave_flag(&flags[2], 1);
        break;
      }
    }
  }
}

__global__ void k_pc_counts(int32_t n, const uint8_t* col_delete, int64_t* r, int64_t* ni, int64_t* nv) {
  const int64_t j = gtid();
  if (j >= n) return;
  const int64_t s = col_delete[j] ? 1 : 0;
  r[j] = s;
  ni[j] = 2 * s;
  nv[j] = 5 * s;
}

__global__ void k_pc_pack(TapeView out, int32_t n, const uint8_t* col_delete, const int32_t* merge_to,
                          const double* mrec, const int64_t* r) {
  const int64_t j = gtid();
  if (j >= n || !col_delete[j]) return;
  const int64_t t = r[j], rec = out.rbase + t;
  out.types[rec] = GP_RED_PARALLEL_COLUMN;
  out.modes[rec] = GP_DUAL_MINIMAL;
  out.idx[out.ibase + 2 * t] = static_cast<int32_t>(j);
  out.idx[out.ibase + 2 * t + 1] = merge_to[j];
  for (int k = 0; k < 5; ++k) out.vals[out.vbase + 5 * t + k] = mrec[5 * j + k];
  out.istarts[rec + 1] = static_cast<gp_offset>(out.ibase + 2 * (t + 1));
  out.vstarts[rec + 1] = static_cast<gp_offset>(out.vbase + 5 * (t + 1));
}

__global__ void k_pc_apply(int32_t n, const uint8_t* col_delete, const uint8_t* fixed, const double* fval,
                           uint8_t* keep_col, double* l, double* u) {
  const int64_t j = gtid();
  if (j >= n || !(col_delete[j] || fixed[j])) return;
  keep_col[j] = 0;
  if (fixed[j]) {
    l[j] = fval[j];
    u[j] = fval[j];
  }
}

// ============================================================== doubleton equations
constexpr double kDblMaxRatioPivot = 1.0e3;
constexpr int kDblMatchingRounds = 8;

// live entries (|a| > zt, live row) per live column, entry-parallel
// (seg_reduce_tiles over AT)
struct DbColNnzOp {
  Bits keep_row;
  const uint8_t* keep_col;
  double zt;
  using Slot = uint32_t;
  __device__ Slot value(int32_t j, int32_t r, double a) const {
    return (keep_col[j] && keep_row[r] && fabs(a) > zt) ? 1u : 0u;
  }
  __device__ static Slot merge(Slot a, Slot b) { return a + b; }
  __device__ static void combine(Slot* s, Slot v) { atomicAdd(s, v); }
};

__device__ inline bool db_integral(double num, double den, double tol) {
  if (fabs(den) <= tol) return false;
  const double r = fabs(num / den);
  if (!isfinite(r)) return false;
  return fabs(r - rint(r)) <= tol;
}

__device__ int32_t db_fill_in(const gp_csr& AT, const uint8_t* keep_row, double zt, int32_t keep, int32_t elim) {
  int32_t fill = -1;
  gp_offset jj = AT.row_ptr[keep], kk = AT.row_ptr[elim];
  const gp_offset jstop = AT.row_ptr[keep + 1], kstop = AT.row_ptr[elim + 1];
  if (jstop <= jj || kstop <= kk) return fill;
  auto live = [&](gp_offset q) { return keep_row[AT.col_idx[q]] && fabs(AT.vals[q]) > zt; };
  while (jj < jstop && kk < kstop) {
    gp_offset jl = jj, kl = kk;
    while (jl < jstop && !live(jl)) ++jl;
    while (kl < kstop && !live(kl)) ++kl;
    if (jl >= jstop || kl >= kstop) break;
    const int32_t rj = AT.col_idx[jl], rk = AT.col_idx[kl];
    if (rj == rk) {
      jj = jl + 1;
      kk = kl + 1;
    } else if (rk < rj) {
      kk = kl + 1;
      ++fill;
    } else {
      jj = jl + 1;
    }
  }
  while (kk < kstop) {
    gp_offset kl = kk;
    while (kl < kstop && !live(kl)) ++kl;
    if (kl >= kstop) break;
    ++fill;
    kk = kl + 1;
  }
  return fill;
}

struct DbCand {
  int32_t* elim;
  int32_t* keep;
  double* elim_val;
  double* keep_val;
};

__global__ void k_db_candidates(gp_csr A, gp_csr AT, const uint8_t* keep_row, const uint8_t* keep_col,
                                const double* AL, const double* AU, const int32_t* col_nnz, double zt, double bt,
                                double ratio_tol, int32_t maxfill, DbCand cand, uint8_t* active, int32_t* flags) {
  const int64_t i = gtid();
  if (i >= A.rows) return;
  active[i] = 0;
  if (!keep_row[i]) return;
  const double lhs = AL[i], rhs = AU[i];
  if (!isfinite(lhs) || !isfinite(rhs) || fabs(lhs - rhs) > bt) return;
  int32_t live = 0, c1 = -1, c2 = -1;
  double v1 = 0.0, v2 = 0.0;
  for (gp_offset q = A.row_ptr[i]; q < A.row_ptr[i + 1]; ++q) {
    const int32_t col = A.col_idx[q];
    const double v = A.vals[q];
    if (keep_col[col] && fabs(v) > zt) {
      ++live;
      if (live == 1) {
        c1 = col;
        v1 = v;
      } else if (live == 2) {
        c2 = col;
        v2 = v;
      } else {
        break;
      }
    }
  }
  if (live != 2) return;
  const bool i12 = db_integral(v1, v2, ratio_tol), i21 = db_integral(v2, v1, ratio_tol);
  int32_t ec, kc;
  double ev, kv;
  if (col_nnz[c1] == 1 && col_nnz[c2] != 1) {
    ec = c1; ev = v1; kc = c2; kv = v2;
  } else if (col_nnz[c1] != 1 && col_nnz[c2] == 1) {
    ec = c2; ev = v2; kc = c1; kv = v1;
  } else if (i12 && !i21) {
    ec = c2; ev = v2; kc = c1; kv = v1;
  } else if (i21 && !i12) {
    ec = c1; ev = v1; kc = c2; kv = v2;
  } else if (col_nnz[c1] < col_nnz[c2]) {
    ec = c1; ev = v1; kc = c2; kv = v2;
  } else {
    ec = c2; ev = v2; kc = c1; kv = v1;
  }
  if (fabs(ev) <= 0.0) return;
  const double pr = fabs(kv / ev);
  if (!(pr <= kDblMaxRatioPivot && pr >= 1.0 / kDblMaxRatioPivot)) return;
  cand.elim[i] = ec;
  cand.keep[i] = kc;
  cand.elim_val[i] = ev;
  cand.keep_val[i] = kv;
  if (col_nnz[ec] > col_nnz[kc] + maxfill + 1) return;
  if (db_fill_in(AT, keep_row, zt, kc, ec) <= maxfill) {
    active[i] = 1;
    wave_count(&flags[0]);
  }
}

__global__ void k_db_claim(int32_t m, const uint8_t* active, const uint8_t* blocked, DbCand cand, int32_t* owner) {
  const int64_t i = gtid();
  if (i >= m || !active[i]) return;
  const int32_t kc = cand.keep[i], ec = cand.elim[i];
  if (!blocked[kc] && !blocked[ec]) {
    atomicMin(&owner[kc], static_cast<int32_t>(i));
    atomicMin(&owner[ec], static_cast<int32_t>(i));
  }
}

// A row wins when it owns both columns; winners own disjoint columns, so the
// blocked marks written here never change another row's outcome.
__global__ void k_db_select(int32_t m, uint8_t* active, uint8_t* blocked, uint8_t* selected, DbCand cand,
                            const int32_t* owner) {
  const int64_t i = gtid();
  if (i >= m || !active[i]) return;
  const int32_t kc = cand.keep[i], ec = cand.elim[i];
  if (owner[kc] == i && owner[ec] == i && !blocked[kc] && !blocked[ec]) {
    selected[i] = 1;
    active[i] = 0;
    blocked[kc] = 1;
    blocked[ec] = 1;
  }
}

__global__ void k_db_rows(int32_t m, const uint8_t* selected, const int32_t* o2n, int32_t* rows) {
  const int64_t i = gtid();
  if (i < m && selected[i]) rows[o2n[i]] = static_cast<int32_t>(i);
}

struct DbPairs {
  int32_t k;
  const int32_t* rows;
  double* rec;  // 6 per pair: AU(row), elim l, u, keep l, u, elim c
  uint8_t* subst;
  uint8_t* fixedp;
  double* obj_s;
  double* obj_f;
};

__global__ void k_db_apply(DbPairs P, DbCand cand, double* l, double* u, double* c, const double* AU,
                           uint8_t* keep_row, uint8_t* keep_col, double bt, int32_t* flags, BatchMarks bm) {
  const int64_t t = gtid();
  if (t >= P.k) return;
  const int32_t row = P.rows[t], kc = cand.keep[row], ec = cand.elim[row];
  double* rc = P.rec + 6 * t;
  rc[0] = AU[row];
  rc[1] = l[ec];
  rc[2] = u[ec];
  rc[3] = l[kc];
  rc[4] = u[kc];
  rc[5] = c[ec];
  const double kv = cand.keep_val[row], ev = cand.elim_val[row];
  const double alpha = -kv / ev, beta = AU[row] / ev;
  double ml = -INFINITY, mu = INFINITY;
  if (isfinite(l[ec])) {
    const double b = (l[ec] - beta) / alpha;
    if (alpha > 0.0) ml = b;
    else mu = b;
  }
  if (isfinite(u[ec])) {
    const double b = (u[ec] - beta) / alpha;
    if (alpha > 0.0) mu = b;
    else ml = b;
  }
  if (ml > mu) {
    const double tmp = ml;
    ml = mu;
    mu = tmp;
  }
  const double kl = isfinite(ml) ? cmax(l[kc], ml) : l[kc];
  const double ku = isfinite(mu) ? cmin(u[kc], mu) : u[kc];
  P.subst[t] = 0;
  P.fixedp[t] = 0;
  P.obj_s[t] = 0.0;
  P.obj_f[t] = 0.0;
  if (kl > ku + bt) {
    wave_flag(&flags[1], 1);
    bm.row(row, 1);
    return;
  }
  if (isfinite(kl) && isfinite(ku) && ku <= kl + bt) {
    const double at = 0.5 * (kl + ku);
    l[kc] = at;
    u[kc] = at;
    keep_col[kc] = 0;
    P.fixedp[t] = 1;
    P.obj_f[t] = c[kc] * at;
  } else {
    l[kc] = kl;
    u[kc] = ku;
    c[kc] += c[ec] * alpha;
    keep_row[row] = 0;
    keep_col[ec] = 0;
    P.subst[t] = 1;
    P.obj_s[t] = c[ec] * beta;
  }
  wave_flag(&flags[2], 1);
}

__global__ void k_db_tape_counts(int32_t k, const uint8_t* subst, int64_t* r, int64_t* ni, int64_t* nv) {
  const int64_t t = gtid();
  if (t >= k) return;
  const int64_t s = subst[t] ? 1 : 0;
  r[t] = s;
  ni[t] = 4 * s;
  nv[t] = 12 * s;
}

// DoubletonEquation: idx {elim, keep, row, 0}, vals {elim a, keep a, AU(row),
// elim l, u, keep l, u, new keep l, u, elim c, 0, 0}.
__global__ void k_db_tape_pack(TapeView out, DbPairs P, DbCand cand, const double* l, const double* u,
                               const int64_t* r) {
  const int64_t t = gtid();
  if (t >= P.k || !P.subst[t]) return;
  const int32_t row = P.rows[t], kc = cand.keep[row], ec = cand.elim[row];
  const double* rc = P.rec + 6 * t;
  const int64_t s = r[t], rec = out.rbase + s;
  out.types[rec] = GP_RED_DOUBLETON_EQUATION;
  out.modes[rec] = GP_DUAL_MINIMAL;
  int32_t* ix = out.idx + out.ibase + 4 * s;
  ix[0] = ec;
  ix[1] = kc;
  ix[2] = row;
  ix[3] = 0;
  double* vx = out.vals + out.vbase + 12 * s;
  vx[0] = cand.elim_val[row];
  vx[1] = cand.keep_val[row];
  vx[2] = rc[0];
  vx[3] = rc[1];
  vx[4] = rc[2];
  vx[5] = rc[3];
  vx[6] = rc[4];
  vx[7] = l[kc];
  vx[8] = u[kc];
  vx[9] = rc[5];
  vx[10] = 0.0;
  vx[11] = 0.0;
  out.istarts[rec + 1] = static_cast<gp_offset>(out.ibase + 4 * (s + 1));
  out.vstarts[rec + 1] = static_cast<gp_offset>(out.vbase + 12 * (s + 1));
}

__global__ void k_db_pair_of(DbPairs P, DbCand cand, int32_t* pair_of) {
  const int64_t t = gtid();
  if (t >= P.k) return;
  const int32_t row = P.rows[t];
  if (P.subst[t]) pair_of[cand.elim[row]] = static_cast<int32_t>(t);
  else if (P.fixedp[t]) pair_of[cand.keep[row]] = static_cast<int32_t>(t);
}

// Row shift of kept row i: a * beta for eliminated columns (other rows of the
// pair) and a * fixed value for fixed keep columns, live entries, CSR order.
// Row shift of the substitution / fixing: only rows with an entry in a paired
// column contribute (rows are listed by list_rows_touching; the rest keep
// shift 0); each listed row is folded in CSR order.
__global__ void k_db_pair_mark(int32_t n, const int32_t* pair_of, uint8_t* mark) {
  const int64_t j = gtid();
  if (j < n) mark[j] = pair_of[j] >= 0 ? 1 : 0;
}

__global__ void k_db_shift(gp_csr A, const int32_t* list, const int32_t* count, const int32_t* pair_of, DbPairs P,
                           DbCand cand, const double* AU, const double* l, double zt, double* shift) {
  const int32_t total = *count;
  for (int64_t k = gtid(); k < total; k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t i = list[k];
    RowAcc acc;
    const gp_offset p0 = A.row_ptr[i];
    for (gp_offset q = p0; q < A.row_ptr[i + 1]; ++q) {
      acc.pos(q - p0);
      const int32_t tt = pair_of[A.col_idx[q]];
      const double a = A.vals[q];
      if (tt < 0 || !(fabs(a) > zt)) continue;
      const int32_t prow = P.rows[tt];
      if (P.subst[tt]) {
        if (prow == i) continue;
        acc.add(a * (AU[prow] / cand.elim_val[prow]));
      } else {
        acc.add(a * l[cand.keep[prow]]);
      }
    }
    shift[i] = acc.value();
  }
}

__global__ void k_db_entry_counts(DbPairs P, DbCand cand, gp_csr AT, const uint8_t* keep_row, double zt,
                                  int64_t* cnt) {
  const int64_t t = gtid();
  if (t >= P.k) return;
  int64_t c = 0;
  if (P.subst[t]) {
    const int32_t row = P.rows[t], ec = cand.elim[row];
    for (gp_offset q = AT.row_ptr[ec]; q < AT.row_ptr[ec + 1]; ++q) {
      const int32_t sr = AT.col_idx[q];
      c += (sr != row && keep_row[sr] && fabs(AT.vals[q]) > zt) ? 1 : 0;
    }
  }
  cnt[t] = c;
}

__global__ void k_db_delta(DbPairs P, DbCand cand, gp_csr AT, const uint8_t* keep_row, double zt, int32_t n,
                           const int64_t* start, uint64_t* keys, double* vals) {
  const int64_t t = gtid();
  if (t >= P.k || !P.subst[t]) return;
  const int32_t row = P.rows[t], ec = cand.elim[row], kc = cand.keep[row];
  const double alpha = -cand.keep_val[row] / cand.elim_val[row];
  int64_t w = start[t];
  for (gp_offset q = AT.row_ptr[ec]; q < AT.row_ptr[ec + 1]; ++q) {
    const int32_t sr = AT.col_idx[q];
    const double a = AT.vals[q];
    if (sr != row && keep_row[sr] && fabs(a) > zt) {
      keys[w] = static_cast<uint64_t>(static_cast<int64_t>(sr) * n + kc);
      vals[w] = alpha * a;
      ++w;
    }
  }
}

__global__ void k_db_delta_rows(int64_t total, const uint64_t* keys, int32_t n, int32_t* counts) {
  const int64_t e = gtid();
  if (e < total) atomicAdd(&counts[static_cast<int64_t>(keys[e] / static_cast<uint64_t>(n))], 1);
}

// Merge of a source row with its sorted deltas (rule_doubleton_equations.cu:1242-1335):
// equal columns are summed (source entries first), |sum| <= zero_tol dropped.
// Row merge of A with the sorted delta entries.  G lanes per row: a row without
// deltas is a stable |a| > zt filter (group ballot prefix); a row with deltas is
// merged sequentially by the group's first lane (equal columns summed A first,
// then the deltas in key order).
template <bool kWrite, int G>
__global__ void __launch_bounds__(kBlock) k_db_merge(gp_csr A, const int32_t* drp, const uint64_t* keys,
                                                      const double* dvals, int32_t n, double zt, int32_t* counts,
                                                      const gp_offset* out_ptr, int32_t* out_col, double* out_val) {
  const int lane = threadIdx.x % G;
  const int wl = static_cast<int>(__lane_id());
  const unsigned long long group_bits = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << (wl - lane);
  const unsigned long long below = (1ull << wl) - 1ull;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t i = gtid() / G; i < A.rows; i += groups) {
    gp_offset pp = A.row_ptr[i];
    const gp_offset pe = A.row_ptr[i + 1];
    int64_t qq = drp[i];
    const int64_t qe = drp[i + 1];
    if (qq == qe) {
      int64_t w = kWrite ? out_ptr[i] : 0;
      int32_t c = 0;
      for (gp_offset base = pp; base < pe; base += G) {
        const gp_offset p = base + lane;
        int32_t col = 0;
        double v = 0.0;
        bool keep = false;
        if (p < pe) {
          col = A.col_idx[p];
          v = A.vals[p];
          keep = fabs(v) > zt;
        }
        const unsigned long long bits = __ballot(keep) & group_bits;
        if (kWrite && keep) {
          const int64_t at = w + __popcll(bits & below);
          out_col[at] = col;
          out_val[at] = v;
        }
        w += __popcll(bits);
        c += __popcll(bits);
      }
      if (!kWrite && lane == 0) counts[i] = c;
      continue;
    }
    if (lane != 0) continue;
    int64_t w = kWrite ? out_ptr[i] : 0;
    int32_t c = 0;
    while (pp < pe || qq < qe) {
      const int32_t sc = pp < pe ? A.col_idx[pp] : INT_MAX;
      const int32_t dc = qq < qe ? static_cast<int32_t>(keys[qq] % static_cast<uint64_t>(n)) : INT_MAX;
      const int32_t next = sc < dc ? sc : dc;
      double merged = 0.0;
      while (pp < pe && A.col_idx[pp] == next) merged += A.vals[pp++];
      while (qq < qe && static_cast<int32_t>(keys[qq] % static_cast<uint64_t>(n)) == next) merged += dvals[qq++];
      if (fabs(merged) > zt) {
        if (kWrite) {
          out_col[w] = next;
          out_val[w] = merged;
          ++w;
        } else {
          ++c;
        }
      }
    }
    if (!kWrite) counts[i] = c;
  }
}

template <bool kWrite>
void launch_db_merge(const gp_csr& A, const int32_t* drp, const uint64_t* keys, const double* dvals, int32_t n,
                     double zt, int32_t* counts, const gp_offset* out_ptr, int32_t* out_col, double* out_val) {
  const int lanes = lanes_for_mean(A.rows ? static_cast<double>(A.nnz) / A.rows : 0.0);
  const unsigned grid = grid_capped(static_cast<int64_t>(A.rows) * lanes, kBlock, 16384);
  switch (lanes) {
    case 2: hipLaunchKernelGGL((k_db_merge<kWrite, 2>), dim3(grid), dim3(kBlock), 0, 0, A, drp, keys, dvals, n, zt, counts, out_ptr, out_col, out_val); break;
    case 4: hipLaunchKernelGGL((k_db_merge<kWrite, 4>), dim3(grid), dim3(kBlock), 0, 0, A, drp, keys, dvals, n, zt, counts, out_ptr, out_col, out_val); break;
    case 8: hipLaunchKernelGGL((k_db_merge<kWrite, 8>), dim3(grid), dim3(kBlock), 0, 0, A, drp, keys, dvals, n, zt, counts, out_ptr, out_col, out_val); break;
    case 16: hipLaunchKernelGGL((k_db_merge<kWrite, 16>), dim3(grid), dim3(kBlock), 0, 0, A, drp, keys, dvals, n, zt, counts, out_ptr, out_col, out_val); break;
    default: hipLaunchKernelGGL((k_db_merge<kWrite, 64>), dim3(grid), dim3(kBlock), 0, 0, A, drp, keys, dvals, n, zt, counts, out_ptr, out_col, out_val); break;
  }
}

__global__ void k_i32_to_offset(const int32_t* src, int64_t n, gp_offset* dst) {
  const int64_t i = gtid();
  if (i < n) dst[i] = static_cast<gp_offset>(src[i]);
}

__global__ void k_db_shift_apply(int32_t m, const double* shift, double* AL, double* AU) {
  const int64_t i = gtid();
  if (i < m) {
    AL[i] -= shift[i];
    AU[i] -= shift[i];
  }
}
}  // namespace

// ---------------------------------------------------------------- singleton columns
void rule_singleton_columns(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  if (plan.has_infeasible || plan.has_unbounded || !p.enable_singleton_columns) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  // singleton statistics, c, l, u per column; row activity cache, AL, AU per row
  TimedScope ts(GP_KT_SINGLETON_COLUMNS, 37.0 * n + 56.0 * m);
  Flags flags;
  hipLaunchKernelGGL(k_sc_count, dim3(grid_capped(n)), dim3(kBlock), 0, 0, n, plan.keep_col_mask,
                     s.singleton_column_mask, flags.d.data());
  GP_LAUNCH_CHECK("singleton columns count");
  if (flags.read()[0] == 0) return;
  DVec<double> AL_new(m), AU_new(m), side(n), contrib(n);
  DVec<int32_t> owner(m);
  DVec<uint8_t> action(n), row_delete(m), lo(m), up(m), col_delete(n);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(m)), dim3(kBlock), 0, 0, owner.data(), INT_MAX, m);
  GP_LAUNCH_CHECK("singleton columns owner");
  RuleCaches& rc = rule_caches(plan);
  ScCache* cache = static_cast<ScCache*>(rc.sc);
  if (!cache || cache->m != m || cache->n != n) {
    delete cache;
    cache = new ScCache();
    rc.sc = cache;
    cache->m = m;
    cache->n = n;
    cache->act.resize(m);
    cache->keep_row.resize(m);
    cache->keep_col.resize(n);
  }
  {
    const PackedMask live_col(plan.keep_col_mask, n);
    PackedBounds lu(plan.new_l, plan.new_u, n);
    bool full = !cache->valid, dirty_ran = false;
    if (!full) {
      dirty_ran = true;
      // rows touched since the last call; few of them -> recompute those on one lane each
      DVec<uint8_t> dirty(m);
      DVec<int32_t> list(m);
      fill_u8(dirty.data(), 0, m);
      Flags cnt;
      hipLaunchKernelGGL(k_sc_dirty_cols, dim3(grid_for(n)), dim3(kBlock), 0, 0, lp.AT, lu.data(), cache->lu.data(),
                         plan.keep_col_mask, cache->keep_col.data(), dirty.data());
      hipLaunchKernelGGL(k_sc_dirty_rows, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, plan.keep_row_mask,
                         cache->keep_row.data(), dirty.data(), list.data(), cnt.d.data());
      GP_LAUNCH_CHECK("singleton columns dirty rows");
      const int32_t count = cnt.read()[0];
      if (static_cast<int64_t>(count) * 32 > m) {
        full = true;
      } else if (count > 0) {
        hipLaunchKernelGGL(k_sc_refresh_rows, dim3(grid_for(count)), dim3(kBlock), 0, 0, lp.A, list.data(), count,
                           plan.keep_row_mask, live_col.view(), lu.data(), cache->act.data());
        GP_LAUNCH_CHECK("singleton columns row refresh");
      }
    }
    if (full) {
      // every live row is folded: the sweep refreshes the entry-bounds store
      EntryBounds* store = plan_store(plan);
      DVec<uint32_t> current;
      const StoreView view = store_view(store, lp.A, lu.data(), plan.keep_col_mask, p.zero_tol, true, current);
      {
        // A, (l, u) and a mask bit per column, keep_row, 40 bytes out per row
        TimedScope ts(GP_KT_SC_ACTIVITY, 4.0 * (m + 1) + 12.0 * lp.A.nnz + 16.0 * n + n / 8.0 + 41.0 * m);
        launch_row_fold(lp.A, lu.data(), live_col.view(), ScActivityFold{plan.keep_row_mask, cache->act.data()},
                        view);
        GP_LAUNCH_CHECK("singleton columns row activity");
      }
      store_commit(store, lp.A, lu.data(), p.zero_tol);
    }
    std::swap(cache->lu, lu.lu);
    if (!dirty_ran) {
      // the dirty-row kernels bring the cached masks up to date when they run
      d2d(cache->keep_row.data(), plan.keep_row_mask, static_cast<size_t>(m));
      d2d(cache->keep_col.data(), plan.keep_col_mask, static_cast<size_t>(n));
    }
    cache->valid = true;
  }
  const ScAct* row_act = cache->act.data();
  hipLaunchKernelGGL(k_sc_propose, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, m, row_act, plan.keep_col_mask,
                     s.singleton_column_mask, s.singleton_column_row, s.singleton_column_val, plan.new_c, plan.new_l, plan.new_u, plan.new_AL, plan.new_AU, p.zero_tol,
                     p.bound_tol, action.data(), side.data(), owner.data(), flags.d.data(), batch_marks());
  GP_LAUNCH_CHECK("singleton columns propose");
  const int32_t* pf = flags.read();
  if (pf[1]) {
    plan.has_unbounded = 1;
    return;
  }
  if (!pf[4]) return;  // no proposal
  // the stage's outputs start cleared; a call without proposals never needs them
  hipLaunchKernelGGL(k_sc_clear, dim3(grid_for(m > n ? m : n)), dim3(kBlock), 0, 0, m, n, row_delete.data(), lo.data(),
                     up.data(), col_delete.data(), contrib.data());
  hipLaunchKernelGGL(k_sc_stage, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, n, owner.data(), action.data(), side.data(),
                     s.singleton_column_val, plan.new_c, plan.new_l, plan.new_u, plan.new_AL, plan.new_AU, p.bound_tol,
                     row_delete.data(), lo.data(), up.data(), col_delete.data(), AL_new.data(), AU_new.data(),
                     contrib.data(), flags.d.data());
  GP_LAUNCH_CHECK("singleton columns stage");
  const int32_t* f = flags.read();
  const bool status1 = f[2] != 0, status2 = f[3] != 0;
  if (!status1 && !status2) return;
  if (status1 && tape_on(plan, p)) {
    DeviceTape* tape = plan_tape(plan);
    AppendCounts ac(static_cast<int64_t>(m) + 1);
    hipLaunchKernelGGL(k_sc_sub_counts, dim3(grid_for(m)), dim3(kBlock), 0, 0, lp.A, n, owner.data(),
                       col_delete.data(), plan.keep_col_mask, p.zero_tol, ac.r.data(), ac.i.data(), ac.v.data());
    GP_LAUNCH_CHECK("singleton substitution counts");
    GP_CHECK(hipMemsetAsync(ac.r.data() + m, 0, sizeof(int64_t), 0));
    GP_CHECK(hipMemsetAsync(ac.i.data() + m, 0, sizeof(int64_t), 0));
    GP_CHECK(hipMemsetAsync(ac.v.data() + m, 0, sizeof(int64_t), 0));
    ac.scan(static_cast<int64_t>(m) + 1);
    if (ac.R > 0) {
      tape->reserve_more(ac.R, ac.I, ac.V);
      hipLaunchKernelGGL(k_sc_sub_pack, dim3(grid_for(m)), dim3(kBlock), 0, 0, tape_view(*tape), lp.A, ac.r.data(),
                         ac.i.data(), ac.v.data(), owner.data(), plan.keep_col_mask, row_delete.data(), side.data(),
                         plan.new_l, plan.new_u, plan.new_c);
      GP_LAUNCH_CHECK("singleton substitution pack");
      tape->commit(ac.R, ac.I, ac.V);
    }
  }
  if (status2 && tape_on(plan, p)) {
    DeviceTape* tape = plan_tape(plan);
    AppendCounts ac(m);
    hipLaunchKernelGGL(k_rds_counts, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, lo.data(), up.data(), ac.r.data(),
                       ac.i.data(), ac.v.data());
    GP_LAUNCH_CHECK("row dual shift counts");
    ac.scan(m);
    if (ac.R > 0) {
      tape->reserve_more(ac.R, ac.I, ac.V);
      hipLaunchKernelGGL(k_rds_pack, dim3(grid_for(m)), dim3(kBlock), 0, 0, tape_view(*tape), m, lo.data(), up.data(),
                         ac.r.data());
      GP_LAUNCH_CHECK("row dual shift pack");
      tape->commit(ac.R, ac.I, ac.V);
    }
  }
  if (status1) {
    // cost update after the tape, which records the pre-update costs
    DVec<uint8_t> touched(n);
    fill_u8(touched.data(), 0, n);
    hipLaunchKernelGGL(k_sc_touch, dim3(grid_for(m)), dim3(kBlock), 0, 0, lp.A, n, owner.data(), col_delete.data(),
                       touched.data());
    GP_LAUNCH_CHECK("singleton columns touch");
    launch_grouped<ScCostK>(lanes_for_mean(n ? static_cast<double>(lp.AT.nnz) / n : 0.0), n, lp.AT,
                            static_cast<const uint8_t*>(plan.keep_col_mask),
                            static_cast<const uint8_t*>(col_delete.data()), static_cast<const uint8_t*>(touched.data()),
                            static_cast<const int32_t*>(owner.data()),
                            static_cast<const double*>(s.singleton_column_val), plan.new_c);
    GP_LAUNCH_CHECK("singleton columns cost");
  }
  const int32_t mn = m > n ? m : n;
  hipLaunchKernelGGL(k_sc_commit, dim3(grid_for(mn)), dim3(kBlock), 0, 0, m, n, row_delete.data(), owner.data(),
                     AL_new.data(), AU_new.data(), col_delete.data(), plan.keep_row_mask, plan.keep_col_mask,
                     plan.new_AL, plan.new_AU);
  GP_LAUNCH_CHECK("singleton columns commit");
  plan.obj_constant_delta += obj_delta(contrib.data(), n);
  plan.has_row_action = 1;
  plan.has_col_action = plan.has_col_action || status1;
  plan.has_change = 1;
}

void sc_cache_free(void* cache) { delete static_cast<ScCache*>(cache); }

// ---------------------------------------------------------------- dual fix
void rule_dual_fix(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t n = lp.A.cols, m = lp.A.rows;
  if (n == 0) return;
  // one pass over AT; c, l, u and keep_col per column; AL, AU and keep_row per row
  TimedScope ts(GP_KT_DUAL_FIX, 4.0 * (n + 1) + 12.0 * lp.AT.nnz + 25.0 * n + 17.0 * m);
  Flags flags;
  DVec<uint8_t> fixed(n), inf(n), ffix(n);
  DVec<double> val(n), contrib(n);
  {
    DVec<uint8_t> side(m ? m : 1);
    DVec<uint32_t> sides(n);
    if (m > 0) {
      hipLaunchKernelGGL(k_df_row_sides, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, plan.keep_row_mask, plan.new_AL,
                         plan.new_AU, side.data());
      GP_LAUNCH_CHECK("dual fix row sides");
    }
    seg_reduce_tiles(lp.AT, DfSidesOp{side.data(), p.zero_tol}, sides.data());
    hipLaunchKernelGGL(k_df_candidates, dim3(grid_for(n)), dim3(kBlock), 0, 0, lp.AT, plan.keep_col_mask,
                       static_cast<const uint32_t*>(sides.data()), plan.new_c, plan.new_l, plan.new_u, p.zero_tol,
                       p.bound_tol, p.dual_fix_max_nonzero_cost_col_degree, s.live_masks ? s.col_nnz : nullptr,
                       fixed.data(), inf.data(), ffix.data(), val.data(), contrib.data(), flags.d.data(),
                       batch_marks());
  }
  GP_LAUNCH_CHECK("dual fix candidates");
  const int32_t* f = flags.read();
  if (f[0] == 1) {
    plan.has_infeasible = 1;
    return;
  }
  if (f[0] == 2) {
    plan.has_unbounded = 1;
    return;
  }
  if (!f[1]) return;
  const bool any_inf = f[2] != 0;
  DVec<double> shift(m ? m : 1);
  DVec<uint8_t> row_remove(m ? m : 1), keep_after(m ? m : 1);
  if (m > 0) {
    GP_CHECK(hipMemsetAsync(shift.data(), 0, sizeof(double) * static_cast<size_t>(m), 0));  // +0.0
    fill_u8(row_remove.data(), 0, m);
    d2d(keep_after.data(), plan.keep_row_mask, static_cast<size_t>(m));
    DVec<uint8_t> touched(n);
    hipLaunchKernelGGL(k_df_touch, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, ffix.data(), inf.data(),
                       touched.data());
    DVec<int32_t> list(m);
    int32_t* count = flags.d.data() + 4;
    list_rows_touching(lp.A, &lp.AT, plan.keep_row_mask, touched.data(), list.data(), count);
    hipLaunchKernelGGL(k_df_rows_list, dim3(grid_capped(m, kBlock, 4096)), dim3(kBlock), 0, 0, lp.A, list.data(),
                       count, ffix.data(), inf.data(), val.data(), shift.data(), row_remove.data(), keep_after.data(),
                       flags.d.data());
    GP_LAUNCH_CHECK("dual fix rows");
  }
  const bool row_changed = flags.read()[3] != 0;
  if (tape_on(plan, p)) {
    DeviceTape* tape = plan_tape(plan);
    if (any_inf) {
      AppendCounts ac(n);
      hipLaunchKernelGGL(k_df_inf_counts, dim3(grid_for(n)), dim3(kBlock), 0, 0, lp.A, lp.AT, fixed.data(), inf.data(),
                         plan.keep_row_mask, plan.rule_entry_keep_col, ac.r.data(), ac.i.data(), ac.v.data());
      GP_LAUNCH_CHECK("infinite dual fix counts");
      ac.scan(n);
      if (ac.R > 0) {
        tape->reserve_more(ac.R, ac.I, ac.V);
        hipLaunchKernelGGL(k_df_inf_pack, dim3(grid_for(n)), dim3(kBlock), 0, 0, tape_view(*tape), lp.A, lp.AT,
                           fixed.data(), inf.data(), plan.keep_row_mask, plan.rule_entry_keep_col, val.data(),
                           plan.new_l, plan.new_u, plan.new_AL, plan.new_AU, ac.r.data(), ac.i.data(), ac.v.data());
        GP_LAUNCH_CHECK("infinite dual fix pack");
        tape->commit(ac.R, ac.I, ac.V);
      }
    }
    if (p.record_postsolve_tape_cpu)
      fixed_variable_records(tape, ffix.data(), val.data(), keep_after.data(), plan.new_c, lp.AT, n, true);
  }
  const int32_t mn = m > n ? m : n;
  hipLaunchKernelGGL(k_df_apply, dim3(grid_for(mn)), dim3(kBlock), 0, 0, m, n, fixed.data(), val.data(), plan.new_l,
                     plan.new_u, plan.keep_col_mask, shift.data(), row_remove.data(), plan.new_AL, plan.new_AU,
                     plan.keep_row_mask);
  GP_LAUNCH_CHECK("dual fix apply");
  plan.obj_constant_delta += obj_delta(contrib.data(), n);
  plan.has_row_action = plan.has_row_action || row_changed;
  plan.has_col_action = 1;
  plan.has_change = 1;
}

// compute_parallel_row_hashes (rule_parallel_rows.cu:583-627)
void parallel_row_hashes(const gp_csr& A, const uint8_t* keep_row, const uint8_t* keep_col, double tol,
                         uint64_t* hash) {
  const int32_t m = A.rows;
  if (m <= 0) return;
  const int use_xor = static_cast<int64_t>(A.nnz) > 32 * static_cast<int64_t>(m) ? 1 : 0;
  DVec<int32_t> len(m), ids(m);
  DVec<gp_offset> first(m);
  first_live(A, LivePr{keep_col}, len.data(), first.data());
  launch_line_stream(A, PrHash{keep_row, LivePr{keep_col}, len.data(), first.data(), A.vals, tol, use_xor, hash,
                               ids.data()});
}

// ---------------------------------------------------------------- parallel rows
void rule_parallel_rows(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows;
  if (m <= 1) return;
  const uint8_t* keep_col = s.live_masks ? plan.keep_col_mask : nullptr;
  const int use_xor = static_cast<int64_t>(lp.A.nnz) > 32 * static_cast<int64_t>(m) ? 1 : 0;
  DVec<uint64_t> hash(m);
  DVec<int32_t> ids(m);
  {
    // hash of every row (one pass over A, 8 bytes out per row) and the sort of the (hash, row) pairs
    TimedScope ts(GP_KT_HASH_SORT, 4.0 * (m + 1) + 12.0 * lp.A.nnz + 8.0 * m + 24.0 * m);
    DVec<int32_t> len(m);
    DVec<gp_offset> first(m);
    first_live(lp.A, LivePr{keep_col}, len.data(), first.data());
    launch_line_stream(lp.A, PrHash{plan.keep_row_mask, LivePr{keep_col}, len.data(), first.data(), lp.A.vals,
                                    p.zero_tol, use_xor, hash.data(), ids.data()});
    sort_pairs_u64_i32(hash.data(), ids.data(), m);
  }
  // sorted (hash, row) pairs, one pass over A for the ratios, AL, AU per row
  TimedScope ts(GP_KT_PARALLEL, 12.0 * m + 4.0 * (m + 1) + 12.0 * lp.A.nnz + 16.0 * m);
  Flags flags;
  DVec<uint8_t> row_delete(m);
  fill_u8(row_delete.data(), 0, m);
  DVec<int32_t> step(m), kept(m);
  DVec<double> rvals(5 * static_cast<size_t>(m));
  PrRecs rec{step.data(), kept.data(), rvals.data()};
  hipLaunchKernelGGL(k_pr_groups, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, hash.data(), ids.data(), lp.A,
                     plan.keep_row_mask, keep_col, plan.new_AL, plan.new_AU, p.zero_tol, p.feasibility_tol,
                     row_delete.data(), rec, flags.d.data(), batch_marks());
  GP_LAUNCH_CHECK("parallel rows groups");
  if (flags.read()[0]) {
    plan.has_infeasible = 1;
    return;
  }
  hipLaunchKernelGGL(k_pr_apply, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, row_delete.data(), plan.keep_row_mask,
                     flags.d.data());
  GP_LAUNCH_CHECK("parallel rows apply");
  const int32_t changed = flags.read()[1];
  if (!changed) return;
  if (tape_on(plan, p)) {
    DeviceTape* tape = plan_tape(plan);
    DVec<int32_t> o2n(m), rows(changed);
    DVec<uint64_t> keys(changed);
    build_old_to_new(row_delete.data(), m, o2n.data());
    hipLaunchKernelGGL(k_pr_keys, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, row_delete.data(), o2n.data(), step.data(),
                       keys.data(), rows.data());
    GP_LAUNCH_CHECK("parallel rows keys");
    sort_pairs_u64_i32(keys.data(), rows.data(), changed);
    tape->reserve_more(changed, 2 * static_cast<int64_t>(changed), 5 * static_cast<int64_t>(changed));
    hipLaunchKernelGGL(k_pr_pack, dim3(grid_for(changed)), dim3(kBlock), 0, 0, tape_view(*tape), changed, rows.data(),
                       rec);
    GP_LAUNCH_CHECK("parallel rows pack");
    tape->commit(changed, 2 * static_cast<int64_t>(changed), 5 * static_cast<int64_t>(changed));
  }
  plan.has_row_action = 1;
  plan.has_change = 1;
}

// ---------------------------------------------------------------- parallel columns
void rule_parallel_columns(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  (void)s;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (n <= 1) return;
  const double zt = p.zero_tol;
  const double ctol = fmax(p.zero_tol, p.bound_tol), otol = ctol;
  DVec<uint64_t> hash(n);
  DVec<int32_t> ids(n);
  {
    // hash of every column (one pass over AT, 8 bytes out per column) and the sort of the pairs
    TimedScope ts(GP_KT_HASH_SORT, 4.0 * (n + 1) + 12.0 * lp.AT.nnz + 8.0 * n + 24.0 * n);
    DVec<int32_t> len(n);
    DVec<gp_offset> first(n);
    first_live(lp.AT, LivePc{plan.keep_row_mask, zt}, len.data(), first.data());
    launch_line_stream(lp.AT, PcHash{plan.keep_col_mask, LivePc{plan.keep_row_mask, zt}, len.data(), first.data(),
                                     lp.AT.vals, hash.data(), ids.data()});
    sort_pairs_u64_i32(hash.data(), ids.data(), n);
  }
  // sorted (hash, column) pairs, one pass over AT for the ratios, c, l, u per column
  TimedScope ts(GP_KT_PARALLEL, 12.0 * n + 4.0 * (n + 1) + 12.0 * lp.AT.nnz + 24.0 * n);
  Flags flags;
  double active_max = p.parallel_columns_max_abs_ratio;
  if (std::isfinite(p.parallel_columns_max_abs_ratio) && p.parallel_columns_scale_guard_min_groups > 1) {
    hipLaunchKernelGGL(k_pc_guard, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, hash.data(), ids.data(), lp.AT,
                       plan.keep_col_mask, plan.keep_row_mask, plan.new_c, zt, ctol, otol,
                       p.parallel_columns_max_abs_ratio, p.parallel_columns_scale_guard_max_group_size,
                       flags.d.data());
    GP_LAUNCH_CHECK("parallel columns scale guard");
    if (flags.read()[4] < p.parallel_columns_scale_guard_min_groups) active_max = INFINITY;
  }
  DVec<uint8_t> col_delete(n), fixed(n);
  fill_u8(col_delete.data(), 0, n);
  fill_u8(fixed.data(), 0, n);
  DVec<double> fval(n), mrec(5 * static_cast<size_t>(n));
  DVec<int32_t> merge_to(n);
  PcOut o{col_delete.data(), fixed.data(), fval.data(), merge_to.data(), mrec.data()};
  hipLaunchKernelGGL(k_pc_groups, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, hash.data(), ids.data(), lp.AT,
                     plan.keep_col_mask, plan.keep_row_mask, plan.new_c, plan.new_l, plan.new_u, zt, ctol, otol,
                     active_max, o, flags.d.data(), batch_marks());
  GP_LAUNCH_CHECK("parallel columns groups");
  const int32_t* f = flags.read();
  if (f[0]) {
    plan.has_unbounded = 1;
    return;
  }
  const bool status1 = f[1] != 0, status2 = f[2] != 0;
  if (!status1 && !status2) return;
  if (status1 && tape_on(plan, p)) {
    DeviceTape* tape = plan_tape(plan);
    AppendCounts ac(n);
    hipLaunchKernelGGL(k_pc_counts, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, col_delete.data(), ac.r.data(),
                       ac.i.data(), ac.v.data());
    GP_LAUNCH_CHECK("parallel columns counts");
    ac.scan(n);
    if (ac.R > 0) {
      tape->reserve_more(ac.R, ac.I, ac.V);
      hipLaunchKernelGGL(k_pc_pack, dim3(grid_for(n)), dim3(kBlock), 0, 0, tape_view(*tape), n, col_delete.data(),
                         merge_to.data(), mrec.data(), ac.r.data());
      GP_LAUNCH_CHECK("parallel columns pack");
      tape->commit(ac.R, ac.I, ac.V);
    }
  }
  double obj = 0.0;
  if (status2) {
    obj = fixed_variable_records(tape_on(plan, p) ? plan_tape(plan) : nullptr, fixed.data(), fval.data(),
                                 plan.keep_row_mask, plan.new_c, lp.AT, n, true);
    if (m > 0) {
      shift_rows_by_columns(lp.A, &lp.AT, plan.keep_row_mask, fixed.data(), fval.data(), plan.new_AL,
                            plan.new_AU);
      GP_LAUNCH_CHECK("parallel columns row shift");
    }
    plan.has_row_action = 1;
  }
  hipLaunchKernelGGL(k_pc_apply, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, col_delete.data(), fixed.data(),
                     fval.data(), plan.keep_col_mask, plan.new_l, plan.new_u);
  GP_LAUNCH_CHECK("parallel columns apply");
  plan.obj_constant_delta += obj;
  plan.has_col_action = 1;
  plan.has_change = 1;
}

// ---------------------------------------------------------------- doubleton equations
// ---------------------------------------------------------------- doubleton equations, single row
// The single-row form (doubleton_equations_scan off,
// rule_doubleton_equations.cu:1948-2128): one doubleton equation per call,
// the lowest live equality row with two live entries whose pivot ratio and
// fill-in pass.  The row is picked by an atomic minimum and its data are then
// read from that row alone, so the choice never depends on thread timing.
namespace {
// fill-in proxy over every entry of the two columns of A^T, live or not
// (_doubleton_equations_fill_in_proxy_device, :216-243)
__device__ int32_t db_fill_in_all(const gp_csr& AT, int32_t keep, int32_t elim) {
  int32_t fill = -1;
  gp_offset jj = AT.row_ptr[keep], kk = AT.row_ptr[elim];
  const gp_offset jstop = AT.row_ptr[keep + 1], kstop = AT.row_ptr[elim + 1];
  if (jstop <= jj || kstop <= kk) return fill;
  while (jj < jstop && kk < kstop) {
    const int32_t rj = AT.col_idx[jj], rk = AT.col_idx[kk];
    if (rj == rk) {
      ++jj;
      ++kk;
    } else if (rk < rj) {
      ++kk;
      ++fill;
    } else {
      ++jj;
    }
  }
  return fill + static_cast<int32_t>(kstop - kk);
}

struct DbsArgs {
  gp_csr A, AT;
  const uint8_t* keep_row;
  const uint8_t* keep_col;
  const double* AL;
  const double* AU;
  const int32_t* col_nnz;
  double zt, bt, ratio_tol;
  int32_t maxfill;
};

// row i as a single-row candidate (:273-363): eliminated and kept column
__device__ bool dbs_candidate(const DbsArgs& g, int64_t i, int32_t& ec, int32_t& kc, double& ev, double& kv) {
  if (!g.keep_row[i]) return false;
  const double lhs = g.AL[i], rhs = g.AU[i];
  if (!isfinite(lhs) || !isfinite(rhs) || fabs(lhs - rhs) > g.bt) return false;
  int32_t live = 0, c1 = -1, c2 = -1;
  double v1 = 0.0, v2 = 0.0;
  for (gp_offset q = g.A.row_ptr[i]; q < g.A.row_ptr[i + 1]; ++q) {
    const int32_t col = g.A.col_idx[q];
    const double v = g.A.vals[q];
    if (g.keep_col[col] && fabs(v) > g.zt) {
      ++live;
      if (live == 1) {
        c1 = col;
        v1 = v;
      } else if (live == 2) {
        c2 = col;
        v2 = v;
      } else {
        return false;
      }
    }
  }
  if (live != 2) return false;
  const int32_t* cn = g.col_nnz;
  const bool i12 = db_integral(v1, v2, g.ratio_tol), i21 = db_integral(v2, v1, g.ratio_tol);
  if (cn[c1] == 1 && cn[c2] != 1) {
    ec = c1; ev = v1; kc = c2; kv = v2;
  } else if (cn[c1] != 1 && cn[c2] == 1) {
    ec = c2; ev = v2; kc = c1; kv = v1;
  } else if (i12 && !i21) {
    ec = c2; ev = v2; kc = c1; kv = v1;
  } else if (i21 && !i12) {
    ec = c1; ev = v1; kc = c2; kv = v2;
  } else if (cn[c1] < cn[c2]) {
    ec = c1; ev = v1; kc = c2; kv = v2;
  } else {
    ec = c2; ev = v2; kc = c1; kv = v1;
  }
  if (fabs(ev) <= 0.0) return false;
  const double pr = fabs(kv / ev);
  if (!(pr <= kDblMaxRatioPivot && pr >= 1.0 / kDblMaxRatioPivot)) return false;
  if (cn[ec] > cn[kc] + g.maxfill + 1) return false;
  return db_fill_in_all(g.AT, kc, ec) <= g.maxfill;
}

__global__ void k_dbs_target(DbsArgs g, int32_t* target) {
  const int64_t i = gtid();
  if (i >= g.A.rows) return;
  int32_t ec, kc;
  double ev, kv;
  if (dbs_candidate(g, i, ec, kc, ev, kv)) atomicMin(target, static_cast<int32_t>(i));
}

struct DbsMeta {
  int32_t status;  // 0: no candidate, 1: infeasible, 2: applied
  int32_t row, ec, kc;
  int64_t at0, at1;  // the eliminated column's entries in A^T
  double ev, kv, rhs, old_el, old_eu, old_kl, old_ku, kl, ku, elim_obj, alpha, beta, obj;
};

// the target row's substitution (_kernel_doubleton_equations_apply, :365-461),
// one lane: mapped bounds of the kept column, then the column and row updates
__global__ void k_dbs_apply(DbsArgs g, const int32_t* target, double* l, double* u, double* c, uint8_t* keep_row,
                            uint8_t* keep_col, DbsMeta* meta, BatchMarks bm) {
  DbsMeta r{};
  const int32_t row = *target;
  if (row >= g.A.rows || !dbs_candidate(g, row, r.ec, r.kc, r.ev, r.kv)) {
    *meta = r;
    return;
  }
  r.row = row;
  r.rhs = g.AU[row];
  r.old_el = l[r.ec];
  r.old_eu = u[r.ec];
  r.elim_obj = c[r.ec];
  r.old_kl = l[r.kc];
  r.old_ku = u[r.kc];
  r.alpha = -r.kv / r.ev;
  r.beta = r.rhs / r.ev;
  double ml = -INFINITY, mu = INFINITY;
  if (isfinite(r.old_el)) {
    const double b = (r.old_el - r.beta) / r.alpha;
    if (r.alpha > 0.0) ml = b;
    else mu = b;
  }
  if (isfinite(r.old_eu)) {
    const double b = (r.old_eu - r.beta) / r.alpha;
    if (r.alpha > 0.0) mu = b;
    else ml = b;
  }
  if (ml > mu) {
    const double t = ml;
    ml = mu;
    mu = t;
  }
  r.kl = isfinite(ml) ? fmax(r.old_kl, ml) : r.old_kl;
  r.ku = isfinite(mu) ? fmin(r.old_ku, mu) : r.old_ku;
  r.at0 = g.AT.row_ptr[r.ec];
  r.at1 = g.AT.row_ptr[r.ec + 1];
  if (r.kl > r.ku + g.bt) {
    r.status = 1;
    bm.row(row, 1);
    *meta = r;
    return;
  }
  l[r.kc] = r.kl;
  u[r.kc] = r.ku;
  c[r.kc] += r.elim_obj * r.alpha;
  keep_row[row] = 0;
  keep_col[r.ec] = 0;
  r.obj = r.elim_obj * r.beta;
  r.status = 2;
  *meta = r;
}

// AL, AU of every row with an entry in the eliminated column: minus that
// entry (the first for the row) times beta
__global__ void k_dbs_shift(gp_csr AT, const DbsMeta* meta, double* AL, double* AU) {
  const int64_t q = meta->at0 + gtid();
  if (q >= meta->at1) return;
  const int32_t t = AT.col_idx[q];
  if (q > meta->at0 && AT.col_idx[q - 1] == t) return;
  const double shift = AT.vals[q] * meta->beta;
  AL[t] -= shift;
  AU[t] -= shift;
}

__global__ void k_dbs_tape(TapeView v, const DbsMeta* meta) {
  const DbsMeta& r = *meta;
  v.types[v.rbase] = GP_RED_DOUBLETON_EQUATION;
  v.modes[v.rbase] = GP_DUAL_MINIMAL;
  const int32_t ix[4] = {r.ec, r.kc, r.row, 0};
  const double vx[12] = {r.ev, r.kv, r.rhs, r.old_el, r.old_eu, r.old_kl, r.old_ku, r.kl, r.ku, r.elim_obj, 0.0, 0.0};
  for (int k = 0; k < 4; ++k) v.idx[v.ibase + k] = ix[k];
  for (int k = 0; k < 12; ++k) v.vals[v.vbase + k] = vx[k];
  v.istarts[v.rbase + 1] = static_cast<gp_offset>(v.ibase + 4);
  v.vstarts[v.rbase + 1] = static_cast<gp_offset>(v.vbase + 12);
}

// one row of the rewritten A (_kernel_doubleton_equations_row_counts and
// _kernel_copy_doubleton_equations_rows, :463-521, 1814-1892): the target row
// as it is; elsewhere the kept column's entry becomes old + alpha * a_elim,
// dropped when |.| <= zero_tol and inserted in column order when the row has
// an eliminated entry and no kept entry; the eliminated column's entries stay
template <bool kWrite>
__global__ void k_dbs_rows(gp_csr A, const DbsMeta* meta, double zt, int32_t* counts, const gp_offset* out_ptr,
                           int32_t* out_col, double* out_val) {
  const int64_t r = gtid();
  if (r >= A.rows) return;
  const int32_t ec = meta->ec, kc = meta->kc;
  const double alpha = meta->alpha;
  const gp_offset first = A.row_ptr[r], last = A.row_ptr[r + 1];
  int64_t w = kWrite ? out_ptr[r] : 0;
  auto put = [&](int32_t col, double val) {
    if (kWrite) {
      out_col[w] = col;
      out_val[w] = val;
    }
    ++w;
  };
  if (r == meta->row) {
    for (gp_offset q = first; q < last; ++q) put(A.col_idx[q], A.vals[q]);
  } else {
    bool elim_present = false, keep_present = false;
    double are = 0.0, old_keep = 0.0;
    for (gp_offset q = first; q < last; ++q) {
      if (A.col_idx[q] == ec) {
        elim_present = true;
        are = A.vals[q];
      } else if (A.col_idx[q] == kc) {
        keep_present = true;
        old_keep = A.vals[q];
      }
    }
    const double new_keep = old_keep + alpha * are;
    const bool insert_keep = elim_present && !keep_present && fabs(new_keep) > zt;
    bool inserted = !insert_keep;
    for (gp_offset q = first; q < last; ++q) {
      const int32_t col = A.col_idx[q];
      if (insert_keep && !inserted && col > kc) {
        put(kc, new_keep);
        inserted = true;
      }
      if (col == kc) {
        if (fabs(new_keep) > zt) put(kc, new_keep);
      } else {
        put(col, A.vals[q]);
      }
    }
    if (insert_keep && !inserted) put(kc, new_keep);
  }
  if (!kWrite) counts[r] = static_cast<int32_t>(w);
}
}  // namespace

void rule_doubleton_single_row(gp_plan& plan, const gp_device_lp& lp, const gp_params& p) {
  const int32_t m = lp.A.rows, n = lp.A.cols;
  // A and A^T sweeps for the counts and the target; c, l, u per column; AL, AU per row
  TimedScope ts(GP_KT_DOUBLETON, 2.0 * (4.0 * (m + 1) + 12.0 * lp.A.nnz) + 24.0 * n + 16.0 * m);
  const gp_csr A = plan.has_new_A ? plan.new_A : lp.A;
  const double zt = p.zero_tol, bt = p.bound_tol;
  DVec<int32_t> col_nnz(n), target(1);
  DVec<DbsMeta> meta(1);
  const PackedMask live_rows(plan.keep_row_mask, m);
  seg_reduce_tiles(lp.AT, DbColNnzOp{live_rows.view(), plan.keep_col_mask, zt},
                   reinterpret_cast<uint32_t*>(col_nnz.data()));
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kBlock), 0, 0, target.data(), m, 1);
  const DbsArgs g{A, lp.AT, plan.keep_row_mask, plan.keep_col_mask, plan.new_AL, plan.new_AU, col_nnz.data(),
                  zt, bt, fmax(1.0e-9, bt), p.doubleton_equations_max_fill_in_proxy};
  hipLaunchKernelGGL(k_dbs_target, dim3(grid_for(m)), dim3(kBlock), 0, 0, g, target.data());
  hipLaunchKernelGGL(k_dbs_apply, dim3(1), dim3(1), 0, 0, g, static_cast<const int32_t*>(target.data()), plan.new_l,
                     plan.new_u, plan.new_c, plan.keep_row_mask, plan.keep_col_mask, meta.data(), batch_marks());
  GP_LAUNCH_CHECK("doubleton single row");
  DbsMeta h;
  d2h(&h, meta.data(), 1);
  if (h.status == 0) return;
  if (h.status == 1) {
    plan.has_infeasible = 1;
    return;
  }
  if (h.at1 > h.at0) {
    hipLaunchKernelGGL(k_dbs_shift, dim3(grid_for(h.at1 - h.at0)), dim3(kBlock), 0, 0, lp.AT,
                       static_cast<const DbsMeta*>(meta.data()), plan.new_AL, plan.new_AU);
    GP_LAUNCH_CHECK("doubleton single row shift");
  }
  if (tape_on(plan, p)) {
    DeviceTape* tape = plan_tape(plan);
    tape->reserve_more(1, 4, 12);
    hipLaunchKernelGGL(k_dbs_tape, dim3(1), dim3(1), 0, 0, tape_view(*tape), static_cast<const DbsMeta*>(meta.data()));
    GP_LAUNCH_CHECK("doubleton single row tape");
    tape->commit(1, 4, 12);
  }
  DVec<int32_t> cnt(static_cast<size_t>(m) + 1);
  hipLaunchKernelGGL(k_dbs_rows<false>, dim3(grid_for(m)), dim3(kBlock), 0, 0, A,
                     static_cast<const DbsMeta*>(meta.data()), zt, cnt.data(), nullptr, nullptr, nullptr);
  GP_LAUNCH_CHECK("doubleton single row counts");
  GP_CHECK(hipMemsetAsync(cnt.data() + m, 0, sizeof(int32_t), 0));
  const int64_t nnz = exclusive_scan_i32_total(cnt.data(), static_cast<int64_t>(m) + 1);
  if (nnz > INT32_MAX) throw std::length_error("doubleton substitution exceeds int32 offsets");
  gp_csr out;
  out.rows = m;
  out.cols = n;
  out.nnz = static_cast<gp_offset>(nnz);
  out.row_ptr = pool_alloc<gp_offset>(static_cast<size_t>(m) + 1);
  out.col_idx = pool_alloc<int32_t>(static_cast<size_t>(nnz));
  out.vals = pool_alloc<double>(static_cast<size_t>(nnz));
  hipLaunchKernelGGL(k_i32_to_offset, dim3(grid_for(m + 1)), dim3(kBlock), 0, 0, cnt.data(), static_cast<int64_t>(m) + 1,
                     out.row_ptr);
  hipLaunchKernelGGL(k_dbs_rows<true>, dim3(grid_for(m)), dim3(kBlock), 0, 0, A,
                     static_cast<const DbsMeta*>(meta.data()), zt, nullptr, out.row_ptr, out.col_idx, out.vals);
  GP_LAUNCH_CHECK("doubleton single row rewrite");
  if (plan.has_new_A) {
    GP_CHECK(hipStreamSynchronize(0));
    csr_pool_free(plan.new_A);
  }
  plan.new_A = out;
  plan.has_new_A = 1;
  plan.obj_constant_delta += h.obj;
  plan.has_row_action = 1;
  plan.has_col_action = 1;
  plan.has_change = 1;
}

void rule_doubleton_equations(gp_plan& plan, const gp_device_lp& lp, const gp_stats& s, const gp_params& p) {
  (void)s;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  if (!p.doubleton_equations_scan) {
    rule_doubleton_single_row(plan, lp, p);
    return;
  }
  // one pass over A; c, l, u per column; AL, AU per row
  TimedScope ts(GP_KT_DOUBLETON, 4.0 * (m + 1) + 12.0 * lp.A.nnz + 24.0 * n + 16.0 * m);
  const gp_csr A = plan.has_new_A ? plan.new_A : lp.A;
  const gp_csr& AT = lp.AT;
  const double zt = p.zero_tol, bt = p.bound_tol;
  const double ratio_tol = fmax(1.0e-9, bt);
  Flags flags;
  DVec<int32_t> col_nnz(n), owner(n), ce(m), ck(m);
  DVec<double> cev(m), ckv(m);
  DVec<uint8_t> active(m), selected(m), blocked(n);
  fill_u8(selected.data(), 0, m);
  fill_u8(blocked.data(), 0, n);
  DbCand cand{ce.data(), ck.data(), cev.data(), ckv.data()};
  const PackedMask live_rows(plan.keep_row_mask, m);
  static_assert(sizeof(*col_nnz.data()) == sizeof(uint32_t), "column counts are 32-bit");
  seg_reduce_tiles(AT, DbColNnzOp{live_rows.view(), plan.keep_col_mask, zt},
                   reinterpret_cast<uint32_t*>(col_nnz.data()));
  hipLaunchKernelGGL(k_db_candidates, dim3(grid_for(m)), dim3(kBlock), 0, 0, A, AT, plan.keep_row_mask,
                     plan.keep_col_mask, plan.new_AL, plan.new_AU, col_nnz.data(), zt, bt, ratio_tol,
                     p.doubleton_equations_max_fill_in_proxy, cand, active.data(), flags.d.data());
  GP_LAUNCH_CHECK("doubleton candidates");
  const int32_t acceptable = flags.read()[0];
  if (acceptable == 0) return;
  for (int round = 0; round < kDblMatchingRounds; ++round) {
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(kBlock), 0, 0, owner.data(), INT_MAX, n);
    hipLaunchKernelGGL(k_db_claim, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, active.data(), blocked.data(), cand,
                       owner.data());
    hipLaunchKernelGGL(k_db_select, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, active.data(), blocked.data(),
                       selected.data(), cand, owner.data());
    GP_LAUNCH_CHECK("doubleton matching");
  }
  DVec<int32_t> o2n(m);
  const int32_t k = build_old_to_new(selected.data(), m, o2n.data());
  const double ratio = static_cast<double>(k) / static_cast<double>(acceptable);
  if (k == 0 || (p.doubleton_equations_scan && (k < p.doubleton_equations_min_selected_per_batch ||
                                                ratio < p.doubleton_equations_min_selected_ratio)))
    return;
  DVec<int32_t> rows(k);
  DVec<double> rec(6 * static_cast<size_t>(k)), obj_s(k), obj_f(k);
  DVec<uint8_t> subst(k), fixedp(k);
  hipLaunchKernelGGL(k_db_rows, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, selected.data(), o2n.data(), rows.data());
  GP_LAUNCH_CHECK("doubleton rows");
  DbPairs P{k, rows.data(), rec.data(), subst.data(), fixedp.data(), obj_s.data(), obj_f.data()};
  hipLaunchKernelGGL(k_db_apply, dim3(grid_for(k)), dim3(kBlock), 0, 0, P, cand, plan.new_l, plan.new_u, plan.new_c,
                     plan.new_AU, plan.keep_row_mask, plan.keep_col_mask, bt, flags.d.data(), batch_marks());
  GP_LAUNCH_CHECK("doubleton apply");
  const int32_t* f = flags.read();
  if (f[1]) {
    plan.has_infeasible = 1;
    return;
  }
  if (!f[2]) return;
  const int32_t subst_count = count_u8(subst.data(), k);
  if (tape_on(plan, p) && subst_count > 0) {
    DeviceTape* tape = plan_tape(plan);
    AppendCounts ac(k);
    hipLaunchKernelGGL(k_db_tape_counts, dim3(grid_for(k)), dim3(kBlock), 0, 0, k, subst.data(), ac.r.data(),
                       ac.i.data(), ac.v.data());
    GP_LAUNCH_CHECK("doubleton tape counts");
    ac.scan(k);
    tape->reserve_more(ac.R, ac.I, ac.V);
    hipLaunchKernelGGL(k_db_tape_pack, dim3(grid_for(k)), dim3(kBlock), 0, 0, tape_view(*tape), P, cand, plan.new_l,
                       plan.new_u, ac.r.data());
    GP_LAUNCH_CHECK("doubleton tape pack");
    tape->commit(ac.R, ac.I, ac.V);
  }
  DVec<int32_t> pair_of(n);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(kBlock), 0, 0, pair_of.data(), -1, n);
  hipLaunchKernelGGL(k_db_pair_of, dim3(grid_for(k)), dim3(kBlock), 0, 0, P, cand, pair_of.data());
  GP_LAUNCH_CHECK("doubleton pair map");
  DVec<double> shift(m);
  {
    GP_CHECK(hipMemsetAsync(shift.data(), 0, sizeof(double) * static_cast<size_t>(m), 0));  // +0.0
    DVec<uint8_t> mark(n);
    hipLaunchKernelGGL(k_db_pair_mark, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, pair_of.data(), mark.data());
    DVec<int32_t> list(m), count(1);
    GP_CHECK(hipMemsetAsync(count.data(), 0, sizeof(int32_t), 0));
    list_rows_touching(A, &AT, plan.keep_row_mask, mark.data(), list.data(), count.data());
    hipLaunchKernelGGL(k_db_shift, dim3(grid_capped(m, kBlock, 4096)), dim3(kBlock), 0, 0, A, list.data(), count.data(),
                       pair_of.data(), P, cand, plan.new_AU, plan.new_l, zt, shift.data());
    GP_LAUNCH_CHECK("doubleton row shift");
  }
  if (subst_count > 0) {
    DVec<int64_t> start(static_cast<size_t>(k) + 1);
    hipLaunchKernelGGL(k_db_entry_counts, dim3(grid_for(k)), dim3(kBlock), 0, 0, P, cand, AT, plan.keep_row_mask, zt,
                       start.data());
    GP_LAUNCH_CHECK("doubleton entry counts");
    int64_t total = 0;
    exclusive_scan_i64(start.data(), k, &total);
    if (total > 0) {
      DVec<uint64_t> keys(static_cast<size_t>(total));
      DVec<double> dvals(static_cast<size_t>(total));
      hipLaunchKernelGGL(k_db_delta, dim3(grid_for(k)), dim3(kBlock), 0, 0, P, cand, AT, plan.keep_row_mask, zt, n,
                         start.data(), keys.data(), dvals.data());
      GP_LAUNCH_CHECK("doubleton deltas");
      sort_pairs_u64_f64(keys.data(), dvals.data(), total);
      DVec<int32_t> drp(static_cast<size_t>(m) + 1), cnt(static_cast<size_t>(m) + 1);
      GP_CHECK(hipMemsetAsync(drp.data(), 0, sizeof(int32_t) * (static_cast<size_t>(m) + 1), 0));
      hipLaunchKernelGGL(k_db_delta_rows, dim3(grid_for(total)), dim3(kBlock), 0, 0, total, keys.data(), n, drp.data());
      GP_LAUNCH_CHECK("doubleton delta rows");
      exclusive_scan_i32_total(drp.data(), static_cast<int64_t>(m) + 1);
      launch_db_merge<false>(A, drp.data(), keys.data(), dvals.data(), n, zt, cnt.data(), nullptr, nullptr, nullptr);
      GP_LAUNCH_CHECK("doubleton merge counts");
      GP_CHECK(hipMemsetAsync(cnt.data() + m, 0, sizeof(int32_t), 0));
      const int64_t nnz = exclusive_scan_i32_total(cnt.data(), static_cast<int64_t>(m) + 1);
      if (nnz > INT32_MAX) throw std::length_error("doubleton substitution exceeds int32 offsets");
      gp_csr out;
      out.rows = m;
      out.cols = n;
      out.nnz = static_cast<gp_offset>(nnz);
      out.row_ptr = pool_alloc<gp_offset>(static_cast<size_t>(m) + 1);
      out.col_idx = pool_alloc<int32_t>(static_cast<size_t>(nnz));
      out.vals = pool_alloc<double>(static_cast<size_t>(nnz));
      hipLaunchKernelGGL(k_i32_to_offset, dim3(grid_for(m + 1)), dim3(kBlock), 0, 0, cnt.data(),
                         static_cast<int64_t>(m) + 1, out.row_ptr);
      launch_db_merge<true>(A, drp.data(), keys.data(), dvals.data(), n, zt, nullptr, out.row_ptr, out.col_idx,
                            out.vals);
      GP_LAUNCH_CHECK("doubleton merge rows");
      if (plan.has_new_A) {
        GP_CHECK(hipStreamSynchronize(0));
        csr_pool_free(plan.new_A);
      }
      plan.new_A = out;
      plan.has_new_A = 1;
    }
  }
  hipLaunchKernelGGL(k_db_shift_apply, dim3(grid_for(m)), dim3(kBlock), 0, 0, m, shift.data(), plan.new_AL,
                     plan.new_AU);
  GP_LAUNCH_CHECK("doubleton shift apply");
  plan.has_row_action = subst_count > 0;
  plan.has_col_action = 1;
  plan.has_change = 1;
  plan.obj_constant_delta += obj_delta(obj_s.data(), k, rows.data(), true);
  plan.obj_constant_delta += obj_delta(obj_f.data(), k, rows.data(), true);
}

}  // namespace gp
