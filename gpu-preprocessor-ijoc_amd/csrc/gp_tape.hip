// gp_tape.hip -- device postsolve tape: growth, append with index
// globalisation (device_recovery.cu:62-185 record layouts), host export and
// FixedVariable record packing (fixed_variable_tape.cu:49-141).
#include "gp_tape.hpp"

namespace gp {

DeviceTape::DeviceTape() { clear(); }

void DeviceTape::clear() {
  count = icount = vcount = 0;
  istarts.grow_keep(1, 0);
  vstarts.grow_keep(1, 0);
  GP_CHECK(hipMemsetAsync(istarts.data(), 0, sizeof(gp_offset), 0));
  GP_CHECK(hipMemsetAsync(vstarts.data(), 0, sizeof(gp_offset), 0));
}

void DeviceTape::reserve_more(int64_t R, int64_t I, int64_t V) {
  if (count + R + 1 > INT32_MAX || icount + I > INT32_MAX || vcount + V > INT32_MAX)
    throw std::length_error("postsolve tape exceeds int32 offsets");
  types.grow_keep(static_cast<size_t>(count + R), static_cast<size_t>(count));
  modes.grow_keep(static_cast<size_t>(count + R), static_cast<size_t>(count));
  istarts.grow_keep(static_cast<size_t>(count + R + 1), static_cast<size_t>(count + 1));
  vstarts.grow_keep(static_cast<size_t>(count + R + 1), static_cast<size_t>(count + 1));
  idx.grow_keep(static_cast<size_t>(icount + I), static_cast<size_t>(icount));
  vals.grow_keep(static_cast<size_t>(vcount + V), static_cast<size_t>(vcount));
}

void DeviceTape::commit(int64_t R, int64_t I, int64_t V) {
  count += R;
  icount += I;
  vcount += V;
}

TapeView tape_view(DeviceTape& t) {
  return TapeView{t.types.data(), t.istarts.data(), t.vstarts.data(), t.modes.data(),
                  t.idx.data(),   t.vals.data(),    t.count,          t.icount,
                  t.vcount};
}

AppendCounts::AppendCounts(int64_t n)
    : r(static_cast<size_t>(n ? n : 1)), i(static_cast<size_t>(n ? n : 1)), v(static_cast<size_t>(n ? n : 1)) {}

void AppendCounts::scan(int64_t n) {
  DVec<int64_t> totals(3);
  exclusive_scan3_i64_device_total(r.data(), i.data(), v.data(), n, totals.data());
  int64_t h[3];
  d2h(h, totals.data(), 3);  // one round trip for the three totals
  R = h[0];
  I = h[1];
  V = h[2];
}

namespace {
__device__ inline int32_t map_col(int32_t x, const int32_t* cols, int32_t n) {
  return x >= 0 && x < n ? cols[x] : x;
}
__device__ inline int32_t map_row(int32_t x, const int32_t* rows, int32_t m) {
  return x >= 0 && x < m ? rows[x] : x;
}

__global__ void k_append_global(TapeView out, const int32_t* types, const gp_offset* is, const gp_offset* vs,
                                 const uint8_t* modes, const int32_t* idx, int64_t count, const int32_t* rows,
                                 int32_t m, const int32_t* cols, int32_t n) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int32_t type = types[k];
  out.types[out.rbase + k] = type;
  out.modes[out.rbase + k] = modes[k];
  out.istarts[out.rbase + k + 1] = static_cast<gp_offset>(out.ibase + is[k + 1]);
  out.vstarts[out.rbase + k + 1] = static_cast<gp_offset>(out.vbase + vs[k + 1]);
  const gp_offset a = is[k];
  const int64_t size = is[k + 1] - a;
  const int32_t* in = idx + a;
  int32_t* o = out.idx + out.ibase + a;
  for (int64_t p = 0; p < size; ++p) o[p] = in[p];
  switch (type) {
    case GP_RED_FIXED_VARIABLE:
      if (size > 0) {
        o[0] = map_col(o[0], cols, n);
        for (int64_t p = 1; p < size; ++p) o[p] = map_row(o[p], rows, m);
      }
      break;
    case GP_RED_INFINITE_BOUND_DUAL_FIX:
      if (size >= 3) {
        o[1] = map_col(o[1], cols, n);
        int64_t p = 3;
        while (p < size) {
          const int64_t len = o[p++];
          for (int64_t t = 0; t < len && p < size; ++t, ++p) o[p] = map_col(o[p], cols, n);
        }
      }
      break;
    case GP_RED_SINGLETON_COLUMN_SUBSTITUTION:
      if (size >= 3) {
        o[0] = map_col(o[0], cols, n);
        o[1] = map_row(o[1], rows, m);
        const int32_t support = o[2];
        for (int32_t t = 0; t < support && 3 + t < size; ++t) o[3 + t] = map_col(o[3 + t], cols, n);
      }
      break;
    case GP_RED_PARALLEL_COLUMN:
      if (size >= 2) {
        o[0] = map_col(o[0], cols, n);
        o[1] = map_col(o[1], cols, n);
      }
      break;
    case GP_RED_PARALLEL_ROW:
      if (size >= 2) {
        o[0] = map_row(o[0], rows, m);
        o[1] = map_row(o[1], rows, m);
      }
      break;
    case GP_RED_DELETED_ROW:
      if (size >= 1) {
        o[0] = map_row(o[0], rows, m);
        if (size >= 2) o[1] = map_col(o[1], cols, n);
      }
      break;
    case GP_RED_ROW_LOWER_BOUND_CHANGE:
    case GP_RED_ROW_UPPER_BOUND_CHANGE:
    case GP_RED_ROW_DUAL_SHIFT:
      if (size >= 1) o[0] = map_row(o[0], rows, m);
      break;
    case GP_RED_VARIABLE_BOUND_CHANGE:
      if (size >= 1) o[0] = map_col(o[0], cols, n);
      break;
    case GP_RED_ROW_IMPLIED_BOUND_CHANGE:
      if (size >= 2) {
        o[0] = map_col(o[0], cols, n);
        o[1] = map_row(o[1], rows, m);
      }
      break;
    case GP_RED_DOUBLETON_EQUATION:
      if (size >= 4) {
        o[0] = map_col(o[0], cols, n);
        o[1] = map_col(o[1], cols, n);
        o[2] = map_row(o[2], rows, m);
        const int32_t support = o[3];
        for (int32_t t = 0; t < support && 4 + t < size; ++t) o[4 + t] = map_row(o[4 + t], rows, m);
      }
      break;
    case GP_RED_FME_COL:
      /* [K, column, K rows, N, L, (orient, len) x L, N columns] */
      if (size >= 4) {
        const int32_t K = o[0];
        o[1] = map_col(o[1], cols, n);
        for (int32_t t = 0; t < K && 2 + t < size; ++t) o[2 + t] = map_row(o[2 + t], rows, m);
        const int64_t sp = 2 + (int64_t)K;
        if (sp + 1 < size) {
          const int32_t N = o[sp], L = o[sp + 1];
          const int64_t first = sp + 2 + 2 * (int64_t)L;
          for (int32_t t = 0; t < N && first + t < size; ++t) o[first + t] = map_col(o[first + t], cols, n);
        }
      }
      break;
    default:
      break;
  }
}
}  // namespace

void tape_append_global(DeviceTape& dst, const DeviceTape& src, const int32_t* rows, int32_t m, const int32_t* cols,
                        int32_t n) {
  if (src.count == 0) return;
  dst.reserve_more(src.count, src.icount, src.vcount);
  TapeView v = tape_view(dst);
  hipLaunchKernelGGL(k_append_global, dim3(grid_for(src.count)), dim3(kBlock), 0, 0, v, src.types.data(),
                     src.istarts.data(), src.vstarts.data(), src.modes.data(), src.idx.data(), src.count, rows, m,
                     cols, n);
  GP_LAUNCH_CHECK("tape append global");
  d2d(dst.vals.data() + dst.vcount, src.vals.data(), static_cast<size_t>(src.vcount));
  dst.commit(src.count, src.icount, src.vcount);
}

void tape_copy_to_host(const DeviceTape& t, int32_t* types, gp_offset* istarts, gp_offset* vstarts, uint8_t* modes,
                       int32_t* idx, double* vals) {
  if (types) d2h(types, t.types.data(), static_cast<size_t>(t.count));
  if (istarts) d2h(istarts, t.istarts.data(), static_cast<size_t>(t.count + 1));
  if (vstarts) d2h(vstarts, t.vstarts.data(), static_cast<size_t>(t.count + 1));
  if (modes) d2h(modes, t.modes.data(), static_cast<size_t>(t.count));
  if (idx) d2h(idx, t.idx.data(), static_cast<size_t>(t.icount));
  if (vals) d2h(vals, t.vals.data(), static_cast<size_t>(t.vcount));
}

// ---- FixedVariable records ---------------------------------------------------------------
namespace {
// a column is recorded when fixed and not excluded (fixed_variable_tape.cu:43-48)
__device__ inline bool fixed_selected(const uint8_t* fixed, const uint8_t* excluded, int64_t j) {
  return fixed[j] != 0 && (!excluded || excluded[j] == 0);
}
__global__ void k_fixed_counts(const uint8_t* fixed, const uint8_t* excluded, const uint8_t* keep_row, gp_csr AT,
                               int32_t n, int with_rows, int64_t* r, int64_t* ni, int64_t* nv) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const bool sel = fixed_selected(fixed, excluded, j);
  int64_t k = 0;
  if (sel && with_rows)
    for (gp_offset p = AT.row_ptr[j]; p < AT.row_ptr[j + 1]; ++p) k += keep_row[AT.col_idx[p]] != 0;
  r[j] = sel ? 1 : 0;
  ni[j] = sel ? 1 + k : 0;
  nv[j] = sel ? 2 + k : 0;
}
__global__ void k_fixed_pack(TapeView out, const uint8_t* fixed, const uint8_t* excluded, const double* fval,
                             const uint8_t* keep_row, const double* c, gp_csr AT, int32_t n, int with_rows,
                             const int64_t* r, const int64_t* ni, const int64_t* nv) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n || !fixed_selected(fixed, excluded, j)) return;
  const int64_t rec = out.rbase + r[j];
  int64_t ip = out.ibase + ni[j], vp = out.vbase + nv[j];
  out.types[rec] = GP_RED_FIXED_VARIABLE;
  out.modes[rec] = GP_DUAL_MINIMAL;
  out.idx[ip++] = static_cast<int32_t>(j);
  out.vals[vp++] = fval[j];
  out.vals[vp++] = c[j];
  if (with_rows)
    for (gp_offset p = AT.row_ptr[j]; p < AT.row_ptr[j + 1]; ++p) {
      const int32_t row = AT.col_idx[p];
      if (!keep_row[row]) continue;
      out.idx[ip++] = row;
      out.vals[vp++] = AT.vals[p];
    }
  out.istarts[rec + 1] = static_cast<gp_offset>(ip);
  out.vstarts[rec + 1] = static_cast<gp_offset>(vp);
}
__global__ void k_fixed_contrib(const uint8_t* fixed, const double* fval, const double* c, int32_t n, double* out) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n) out[j] = fixed[j] ? c[j] * fval[j] : 0.0;
}
__global__ void k_fixed_flags(const uint8_t* fixed, const uint8_t* excluded, int32_t n, int32_t* flag) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n) flag[j] = fixed_selected(fixed, excluded, j) ? 1 : 0;
}
__global__ void k_fixed_list(const uint8_t* fixed, const uint8_t* excluded, int32_t n, const int32_t* rank,
                             int32_t* list) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n && fixed_selected(fixed, excluded, j)) list[rank[j]] = static_cast<int32_t>(j);
}
}  // namespace

double fixed_variable_records(DeviceTape* tape, const uint8_t* fixed, const double* fval, const uint8_t* keep_row,
                              const double* c, const gp_csr& AT, int32_t n, bool with_rows, FixedObjective objective,
                              const uint8_t* excluded) {
  if (n <= 0) return 0.0;
  if (tape) {
    AppendCounts cnt(n);
    hipLaunchKernelGGL(k_fixed_counts, dim3(grid_for(n)), dim3(kBlock), 0, 0, fixed, excluded, keep_row, AT, n,
                       with_rows ? 1 : 0, cnt.r.data(), cnt.i.data(), cnt.v.data());
    GP_LAUNCH_CHECK("fixed counts");
    cnt.scan(n);
    if (cnt.R > 0) {
      tape->reserve_more(cnt.R, cnt.I, cnt.V);
      TapeView v = tape_view(*tape);
      hipLaunchKernelGGL(k_fixed_pack, dim3(grid_for(n)), dim3(kBlock), 0, 0, v, fixed, excluded, fval, keep_row, c,
                         AT, n, with_rows ? 1 : 0, cnt.r.data(), cnt.i.data(), cnt.v.data());
      GP_LAUNCH_CHECK("fixed pack");
      tape->commit(cnt.R, cnt.I, cnt.V);
    }
  }
  if (objective == FixedObjective::Skip) return 0.0;
  if (objective == FixedObjective::Extended) {
    // the selected columns in increasing order, then one sequential sum
    DVec<int32_t> rank(static_cast<size_t>(n));
    hipLaunchKernelGGL(k_fixed_flags, dim3(grid_for(n)), dim3(kBlock), 0, 0, fixed, excluded, n, rank.data());
    GP_LAUNCH_CHECK("fixed flags");
    const int64_t count = exclusive_scan_i32_total(rank.data(), n);
    if (count == 0) return 0.0;
    DVec<int32_t> list(static_cast<size_t>(count));
    hipLaunchKernelGGL(k_fixed_list, dim3(grid_for(n)), dim3(kBlock), 0, 0, fixed, excluded, n, rank.data(),
                       list.data());
    GP_LAUNCH_CHECK("fixed list");
    return obj_delta_x87(c, fval, list.data(), count);
  }
  DVec<double> contrib(static_cast<size_t>(n));
  hipLaunchKernelGGL(k_fixed_contrib, dim3(grid_for(n)), dim3(kBlock), 0, 0, fixed, fval, c, n, contrib.data());
  GP_LAUNCH_CHECK("fixed contrib");
  return obj_delta(contrib.data(), n);
}

}  // namespace gp
