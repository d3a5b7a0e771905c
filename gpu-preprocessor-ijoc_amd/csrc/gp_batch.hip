// gp_batch.hip -- batched presolve: many independent LPs presolved together as
// the block-diagonal direct sum of the instances.
//
// No presolve reduction couples two blocks of a block-diagonal LP (every rule
// acts on one row, one column, a row's entries or a column's entries; parallel
// rows / columns of different blocks never share a support), so presolving the
// direct sum presolves every instance, and each round's sweep covers all
// instances in one launch.  What the batch adds:
//  * assembly: one persistent work-queue kernel per array, each workgroup
//    taking (instance, tile) items from an atomic counter;
//  * the objective constant per instance (obj_delta sums each rule's
//    contributions per instance, in a fixed order);
//  * the split: each instance's rows / columns / nnz in the reduced direct sum
//    (reduction keeps the block order), and extraction of one reduced
//    instance as its own LP.
// The scheduler's decisions (phase changes, gates, limits) are taken for the
// batch as a whole.  A rule that finds an instance infeasible or unbounded
// marks that instance (BatchMarks); the batch then starts over without the
// marked instances, which are reported in the summary's status, so one such
// instance never ends the batch for the others.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <type_traits>

#include "gp_tape.hpp"

namespace gp {

BatchContext*& batch_context() {
  thread_local BatchContext* ctx = nullptr;
  return ctx;
}

BatchMarks batch_marks() {
  BatchMarks b;
  const BatchContext* bc = batch_context();
  if (!bc || bc->count <= 0 || !bc->row_map || !bc->col_map || bc->marks.size() == 0) return b;
  b.row_map = bc->row_map->data();
  b.col_map = bc->col_map->data();
  b.row_off = bc->row_off.data();
  b.col_off = bc->col_off.data();
  b.count = bc->count;
  b.marks = const_cast<int32_t*>(bc->marks.data());
  return b;
}

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// segment boundaries in item space: seg[k] = first item whose original line is
// at or past off[k] (lines increase with the item index)
__global__ void k_seg_bounds(const int32_t* item_line, int64_t n, const int32_t* map, const int32_t* off,
                             int32_t count, int64_t* seg) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > count) return;
  const int32_t key = off[k];
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const int32_t line = item_line ? item_line[mid] : static_cast<int32_t>(mid);
    if (map[line] >= key) hi = mid;
    else lo = mid + 1;
  }
  seg[k] = lo;
}

// one workgroup per instance: 256 strided partial sums in increasing order,
// then a fixed pairwise tree
__global__ void __launch_bounds__(256) k_seg_sum(const double* x, const int64_t* seg, double* out) {
  __shared__ double part[256];
  const int32_t k = blockIdx.x;
  const int64_t s = seg[k], e = seg[k + 1];
  double acc = 0.0;
  for (int64_t i = s + threadIdx.x; i < e; i += 256) acc += x[i];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = part[0];
}

template <class T>
struct Seg {
  const T* src;
  int64_t len;
  int64_t dst_off;
  T add;  // added to integer elements (index rebasing); unused for floating point
};

// Persistent segmented copy: workgroups take (instance, tile) items from an
// atomic counter until every tile is copied.  tiles[k] is the exclusive prefix
// of the instances' tile counts (tiles[count] = total).
template <class T>
__global__ void __launch_bounds__(256) k_seg_copy(const Seg<T>* segs, const int64_t* tiles, int32_t count,
                                                  int64_t tile, unsigned long long* queue, T* dst) {
  __shared__ long long item;
  const int64_t total = tiles[count];
  for (;;) {
    if (threadIdx.x == 0) item = static_cast<long long>(atomicAdd(queue, 1ull));
    __syncthreads();
    const int64_t it = item;
    __syncthreads();
    if (it >= total) break;
    int32_t lo = 0, hi = count - 1;  // the last instance whose first tile is <= it
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (tiles[mid] <= it) lo = mid;
      else hi = mid - 1;
    }
    const Seg<T> s = segs[lo];
    const int64_t b = (it - tiles[lo]) * tile;
    const int64_t e = b + tile < s.len ? b + tile : s.len;
    for (int64_t i = b + threadIdx.x; i < e; i += 256) {
      T v = s.src[i];
      if constexpr (std::is_integral<T>::value) v += s.add;
      dst[s.dst_off + i] = v;
    }
  }
}

__global__ void k_put_offset(gp_offset* p, int64_t at, gp_offset v) { p[at] = v; }

__global__ void k_lower_bounds(const int32_t* sorted, int32_t n, const int32_t* keys, int32_t count, int32_t* out) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > count) return;
  const int32_t key = keys[k];
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (sorted[mid] >= key) hi = mid;
    else lo = mid + 1;
  }
  out[k] = lo;
}

__global__ void k_gather_offsets(const gp_offset* row_ptr, const int32_t* at, int32_t count, gp_offset* out) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= count) out[k] = row_ptr[at[k]];
}

constexpr int64_t kCopyTile = 16384;

// copy segs (host) into dst with the persistent work-queue kernel
template <class T>
void seg_copy(const std::vector<Seg<T>>& segs, T* dst) {
  const int32_t count = static_cast<int32_t>(segs.size());
  if (count == 0) return;
  std::vector<int64_t> tiles(static_cast<size_t>(count) + 1, 0);
  for (int32_t k = 0; k < count; ++k) tiles[k + 1] = tiles[k] + (segs[k].len + kCopyTile - 1) / kCopyTile;
  if (tiles[count] == 0) return;
  DVec<Seg<T>> dsegs(static_cast<size_t>(count));
  DVec<int64_t> dtiles(static_cast<size_t>(count) + 1);
  DVec<unsigned long long> queue(1);
  h2d(dsegs.data(), segs.data(), segs.size());
  h2d(dtiles.data(), tiles.data(), tiles.size());
  GP_CHECK(hipMemsetAsync(queue.data(), 0, sizeof(unsigned long long), 0));
  int dev = 0, cus = 0;
  GP_CHECK(hipGetDevice(&dev));
  GP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t grid = std::min<int64_t>(tiles[count], static_cast<int64_t>(cus > 0 ? cus : 1) * 8);
  hipLaunchKernelGGL(k_seg_copy<T>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, 0, dsegs.data(), dtiles.data(),
                     count, kCopyTile, queue.data(), dst);
  GP_LAUNCH_CHECK("batch segmented copy");
  GP_CHECK(hipStreamSynchronize(0));  // segs / tiles / queue are released on return
}

struct BatchOwned {
  gp_device_lp lp{};
  void free() {
    csr_pool_free(lp.A);
    csr_pool_free(lp.AT);
    pool_free(lp.c);
    pool_free(lp.l);
    pool_free(lp.u);
    pool_free(lp.AL);
    pool_free(lp.AU);
    std::memset(&lp, 0, sizeof(lp));
  }
};

void validate_instance(const gp_device_lp& lp) {
  if (lp.A.rows < 0 || lp.A.cols < 0 || lp.A.nnz < 0) throw std::invalid_argument("negative LP dimensions");
  if (lp.AT.rows != lp.A.cols || lp.AT.cols != lp.A.rows || lp.AT.nnz != lp.A.nnz)
    throw std::invalid_argument("AT does not match A");
  if ((lp.A.rows && (!lp.AL || !lp.AU || !lp.A.row_ptr)) || (lp.A.cols && (!lp.c || !lp.l || !lp.u || !lp.AT.row_ptr)))
    throw std::invalid_argument("missing LP array");
  if (lp.A.nnz && (!lp.A.col_idx || !lp.A.vals || !lp.AT.col_idx || !lp.AT.vals))
    throw std::invalid_argument("missing LP array");
}

template <class T>
T* host_array(size_t count) {
  T* p = static_cast<T*>(std::calloc(count ? count : 1, sizeof(T)));
  if (!p) throw std::bad_alloc();
  return p;
}

}  // namespace

double obj_delta(const double* contrib, int64_t n, const int32_t* item_line, bool item_rows) {
  const double total = det_sum(contrib, n);
  BatchContext* bc = batch_context();
  if (!bc || n <= 0 || bc->count <= 0) return total;
  const DVec<int32_t>* map = item_rows ? bc->row_map : bc->col_map;
  if (!map) throw Error(GP_ERR_RUNTIME, "batch context without line maps");
  const int32_t count = bc->count;
  DVec<int64_t> seg(static_cast<size_t>(count) + 1);
  DVec<double> sums(static_cast<size_t>(count));
  hipLaunchKernelGGL(k_seg_bounds, dim3(grid_for(count + 1)), dim3(kBlock), 0, 0, item_line, n, map->data(),
                     item_rows ? bc->row_off.data() : bc->col_off.data(), count, seg.data());
  hipLaunchKernelGGL(k_seg_sum, dim3(static_cast<unsigned>(count)), dim3(256), 0, 0, contrib, seg.data(), sums.data());
  GP_LAUNCH_CHECK("batch objective split");
  std::vector<double> h(static_cast<size_t>(count));
  d2h(h.data(), sums.data(), h.size());
  for (int32_t k = 0; k < count; ++k) bc->obj[k] += h[k];
  return total;
}

namespace {
// one lane per segment of the ordered column list: sequential x87 sum
__global__ void k_x87_segments(const double* c, const double* v, const int32_t* cols, const int64_t* seg,
                               int32_t segs, double* out) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= segs) return;
  X80 acc = x80_zero();
  for (int64_t i = seg[k]; i < seg[k + 1]; ++i) {
    const int32_t j = cols[i];
    acc = x80_add(acc, x80_mul(x80_from_double(c[j]), x80_from_double(v[j])));
  }
  out[k] = x80_to_double(acc);
}
}  // namespace

double obj_delta_x87(const double* c, const double* v, const int32_t* cols, int64_t count) {
  if (count <= 0) return 0.0;
  DVec<int64_t> whole(2);
  const int64_t h[2] = {0, count};
  h2d(whole.data(), h, 2);
  DVec<double> total(1);
  hipLaunchKernelGGL(k_x87_segments, dim3(1), dim3(64), 0, 0, c, v, cols, static_cast<const int64_t*>(whole.data()), 1,
                     total.data());
  GP_LAUNCH_CHECK("extended objective sum");
  BatchContext* bc = batch_context();
  if (bc && bc->count > 0) {
    if (!bc->col_map) throw Error(GP_ERR_RUNTIME, "batch context without line maps");
    const int32_t segs = bc->count;
    DVec<int64_t> seg(static_cast<size_t>(segs) + 1);
    DVec<double> sums(static_cast<size_t>(segs));
    hipLaunchKernelGGL(k_seg_bounds, dim3(grid_for(segs + 1)), dim3(kBlock), 0, 0, cols, count, bc->col_map->data(),
                       bc->col_off.data(), segs, seg.data());
    hipLaunchKernelGGL(k_x87_segments, dim3(grid_for(segs)), dim3(kBlock), 0, 0, c, v, cols,
                       static_cast<const int64_t*>(seg.data()), segs, sums.data());
    GP_LAUNCH_CHECK("batch extended objective split");
    std::vector<double> hs(static_cast<size_t>(segs));
    d2h(hs.data(), sums.data(), hs.size());
    for (int32_t k = 0; k < segs; ++k) bc->obj[k] += hs[k];
  }
  return read1(total.data());
}

}  // namespace gp

using namespace gp;

extern "C" {

int gp_preprocess_batch(const gp_device_lp* instances, int32_t count, const gp_params* params,
                        gp_batch_summary* out) {
  BatchOwned owned;
  BatchContext ctx;
  int rc = guarded([&] {
    if (!out || !params || (count > 0 && !instances)) throw std::invalid_argument("null argument");
    if (count < 1) throw std::invalid_argument("batch needs at least one instance");
    std::memset(out, 0, sizeof(*out));
    const double t0 = now_s();
    for (int32_t k = 0; k < count; ++k) validate_instance(instances[k]);
    out->count = count;
    out->row_offsets = host_array<int32_t>(count + 1);
    out->col_offsets = host_array<int32_t>(count + 1);
    out->reduced_row_offsets = host_array<int32_t>(count + 1);
    out->reduced_col_offsets = host_array<int32_t>(count + 1);
    out->reduced_nnz_offsets = host_array<gp_offset>(count + 1);
    out->obj_constant = host_array<double>(count);
    out->status = host_array<int32_t>(count);
    double assemble = 0.0, presolve = 0.0;
    // An instance that a rule finds infeasible or unbounded retires alone: the
    // batch starts over without it (as an empty block) and reports it in
    // status[]; every other instance is presolved together as before.
    for (;;) {
      const double ta = now_s();
      std::vector<int64_t> ro(count + 1, 0), co(count + 1, 0), no(count + 1, 0);
      for (int32_t k = 0; k < count; ++k) {
        const bool on = out->status[k] == 0;
        ro[k + 1] = ro[k] + (on ? instances[k].A.rows : 0);
        co[k + 1] = co[k] + (on ? instances[k].A.cols : 0);
        no[k + 1] = no[k] + (on ? instances[k].A.nnz : 0);
      }
      if (ro[count] > INT32_MAX || co[count] > INT32_MAX || no[count] > INT32_MAX)
        throw std::length_error("batch exceeds int32 rows, columns or offsets");
      const int32_t m = static_cast<int32_t>(ro[count]), n = static_cast<int32_t>(co[count]);
      const gp_offset nnz = static_cast<gp_offset>(no[count]);
      gp_device_lp& lp = owned.lp;
      lp.A = gp_csr{m, n, nnz, pool_alloc<gp_offset>(static_cast<size_t>(m) + 1), pool_alloc<int32_t>(nnz),
                    pool_alloc<double>(nnz)};
      lp.AT = gp_csr{n, m, nnz, pool_alloc<gp_offset>(static_cast<size_t>(n) + 1), pool_alloc<int32_t>(nnz),
                     pool_alloc<double>(nnz)};
      lp.c = pool_alloc<double>(n);
      lp.l = pool_alloc<double>(n);
      lp.u = pool_alloc<double>(n);
      lp.AL = pool_alloc<double>(m);
      lp.AU = pool_alloc<double>(m);
      double obj0 = 0.0;
      {
        std::vector<Seg<gp_offset>> rp, tp;
        std::vector<Seg<int32_t>> ci, ti;
        std::vector<Seg<double>> av, tv, c, l, u, al, au;
        for (int32_t k = 0; k < count; ++k) {
          if (out->status[k] != 0) continue;
          const gp_device_lp& in = instances[k];
          const gp_offset nb = static_cast<gp_offset>(no[k]);
          rp.push_back({in.A.row_ptr, in.A.rows, ro[k], nb});
          tp.push_back({in.AT.row_ptr, in.AT.rows, co[k], nb});
          ci.push_back({in.A.col_idx, in.A.nnz, no[k], static_cast<int32_t>(co[k])});
          ti.push_back({in.AT.col_idx, in.AT.nnz, no[k], static_cast<int32_t>(ro[k])});
          av.push_back({in.A.vals, in.A.nnz, no[k], 0.0});
          tv.push_back({in.AT.vals, in.AT.nnz, no[k], 0.0});
          c.push_back({in.c, in.A.cols, co[k], 0.0});
          l.push_back({in.l, in.A.cols, co[k], 0.0});
          u.push_back({in.u, in.A.cols, co[k], 0.0});
          al.push_back({in.AL, in.A.rows, ro[k], 0.0});
          au.push_back({in.AU, in.A.rows, ro[k], 0.0});
          obj0 += in.obj_constant;
        }
        seg_copy(rp, lp.A.row_ptr);
        seg_copy(tp, lp.AT.row_ptr);
        hipLaunchKernelGGL(k_put_offset, dim3(1), dim3(1), 0, 0, lp.A.row_ptr, static_cast<int64_t>(m), nnz);
        hipLaunchKernelGGL(k_put_offset, dim3(1), dim3(1), 0, 0, lp.AT.row_ptr, static_cast<int64_t>(n), nnz);
        GP_LAUNCH_CHECK("batch last offsets");
        seg_copy(ci, lp.A.col_idx);
        seg_copy(ti, lp.AT.col_idx);
        seg_copy(av, lp.A.vals);
        seg_copy(tv, lp.AT.vals);
        seg_copy(c, lp.c);
        seg_copy(l, lp.l);
        seg_copy(u, lp.u);
        seg_copy(al, lp.AL);
        seg_copy(au, lp.AU);
      }
      lp.obj_constant = obj0;
      GP_CHECK(hipDeviceSynchronize());
      const double tb = now_s();
      assemble += tb - ta;

      for (int32_t k = 0; k <= count; ++k) {
        out->row_offsets[k] = static_cast<int32_t>(ro[k]);
        out->col_offsets[k] = static_cast<int32_t>(co[k]);
      }
      ctx.count = count;
      ctx.row_off.resize(count + 1);
      ctx.col_off.resize(count + 1);
      ctx.marks.resize(count);
      h2d(ctx.row_off.data(), out->row_offsets, count + 1);
      h2d(ctx.col_off.data(), out->col_offsets, count + 1);
      GP_CHECK(hipMemsetAsync(ctx.marks.data(), 0, sizeof(int32_t) * static_cast<size_t>(count), 0));
      ctx.obj.assign(count, 0.0);
      batch_context() = &ctx;
      const int prc = gp_preprocess(&lp, params, &out->whole);
      batch_context() = nullptr;
      if (prc != GP_OK) throw Error(prc, std::string("batched presolve: ") + gp_last_error());
      presolve += now_s() - tb;
      if (!(out->whole.has_infeasible || out->whole.has_unbounded)) break;
      std::vector<int32_t> marks(static_cast<size_t>(count));
      d2h(marks.data(), ctx.marks.data(), marks.size());
      bool retired = false;
      for (int32_t k = 0; k < count; ++k) {
        if (marks[k] == 0 || out->status[k] != 0) continue;
        out->status[k] = (marks[k] & 1) ? 1 : 2;
        retired = true;
      }
      if (!retired) break;  // a stage without instance marks ended the batch: reported as a whole
      gp_free_summary(&out->whole);
      owned.free();
      ++out->restarts;
    }
    const double t2 = now_s();

    // split: instance k's rows / columns in the reduced direct sum
    const RecordView rv = record_view(out->whole.record);
    DVec<int32_t> rr(count + 1), rc(count + 1);
    DVec<gp_offset> rn(count + 1);
    hipLaunchKernelGGL(k_lower_bounds, dim3(grid_for(count + 1)), dim3(kBlock), 0, 0, rv.rows, rv.m1,
                       ctx.row_off.data(), count, rr.data());
    hipLaunchKernelGGL(k_lower_bounds, dim3(grid_for(count + 1)), dim3(kBlock), 0, 0, rv.cols, rv.n1,
                       ctx.col_off.data(), count, rc.data());
    hipLaunchKernelGGL(k_gather_offsets, dim3(grid_for(count + 1)), dim3(kBlock), 0, 0,
                       out->whole.reduced_lp.A.row_ptr, rr.data(), count, rn.data());
    GP_LAUNCH_CHECK("batch split");
    d2h(out->reduced_row_offsets, rr.data(), count + 1);
    d2h(out->reduced_col_offsets, rc.data(), count + 1);
    d2h(out->reduced_nnz_offsets, rn.data(), count + 1);
    for (int32_t k = 0; k < count; ++k) out->obj_constant[k] = instances[k].obj_constant + ctx.obj[k];
    out->assembled = owned.lp;
    owned.lp = gp_device_lp{};
    out->assemble_seconds = assemble;
    out->presolve_seconds = presolve;
    out->total_seconds = now_s() - t0;
    (void)t2;
  });
  batch_context() = nullptr;
  if (rc != GP_OK) {
    owned.free();
    if (out) gp_free_batch_summary(out);
  }
  return rc;
}

int gp_postprocess_batch(const double* x, const double* y, const double* z, const gp_batch_summary* s,
                         double* x_out, double* y_out, double* z_out) {
  if (!s || !s->whole.record) {
    set_last_error("batch summary has no record");
    return GP_ERR_INVALID_ARGUMENT;
  }
  return gp_postprocess(x, y, z, &s->whole, &s->assembled, x_out, y_out, z_out);
}

int gp_batch_instance(const gp_batch_summary* s, int32_t k, gp_device_lp* out) {
  BatchOwned owned;
  int rc = guarded([&] {
    if (!s || !out || !s->whole.record) throw std::invalid_argument("null argument");
    if (k < 0 || k >= s->count) throw std::invalid_argument("instance index out of range");
    std::memset(out, 0, sizeof(*out));
    const gp_device_lp& red = s->whole.reduced_lp;
    const int32_t r0 = s->reduced_row_offsets[k], r1 = s->reduced_row_offsets[k + 1];
    const int32_t c0 = s->reduced_col_offsets[k], c1 = s->reduced_col_offsets[k + 1];
    const gp_offset z0 = s->reduced_nnz_offsets[k], z1 = s->reduced_nnz_offsets[k + 1];
    gp_offset t0 = 0, t1 = 0;
    if (red.AT.row_ptr) {
      d2h(&t0, red.AT.row_ptr + c0, 1);
      d2h(&t1, red.AT.row_ptr + c1, 1);
    }
    const int32_t m = r1 - r0, n = c1 - c0;
    const gp_offset nnz = z1 - z0;
    if (t1 - t0 != nnz) throw Error(GP_ERR_RUNTIME, "reduced batch instance is not block-diagonal");
    gp_device_lp& lp = owned.lp;
    lp.A = gp_csr{m, n, nnz, pool_alloc<gp_offset>(static_cast<size_t>(m) + 1), pool_alloc<int32_t>(nnz),
                  pool_alloc<double>(nnz)};
    lp.AT = gp_csr{n, m, nnz, pool_alloc<gp_offset>(static_cast<size_t>(n) + 1), pool_alloc<int32_t>(nnz),
                   pool_alloc<double>(nnz)};
    lp.c = pool_alloc<double>(n);
    lp.l = pool_alloc<double>(n);
    lp.u = pool_alloc<double>(n);
    lp.AL = pool_alloc<double>(m);
    lp.AU = pool_alloc<double>(m);
    seg_copy<gp_offset>({{red.A.row_ptr + r0, static_cast<int64_t>(m) + 1, 0, -z0}}, lp.A.row_ptr);
    seg_copy<gp_offset>({{red.AT.row_ptr + c0, static_cast<int64_t>(n) + 1, 0, -t0}}, lp.AT.row_ptr);
    seg_copy<int32_t>({{red.A.col_idx + z0, nnz, 0, -c0}}, lp.A.col_idx);
    seg_copy<int32_t>({{red.AT.col_idx + t0, nnz, 0, -r0}}, lp.AT.col_idx);
    seg_copy<double>({{red.A.vals + z0, nnz, 0, 0.0}}, lp.A.vals);
    seg_copy<double>({{red.AT.vals + t0, nnz, 0, 0.0}}, lp.AT.vals);
    seg_copy<double>({{red.c + c0, n, 0, 0.0}}, lp.c);
    seg_copy<double>({{red.l + c0, n, 0, 0.0}}, lp.l);
    seg_copy<double>({{red.u + c0, n, 0, 0.0}}, lp.u);
    seg_copy<double>({{red.AL + r0, m, 0, 0.0}}, lp.AL);
    seg_copy<double>({{red.AU + r0, m, 0, 0.0}}, lp.AU);
    lp.obj_constant = s->obj_constant[k];
    GP_CHECK(hipDeviceSynchronize());
    *out = lp;
    owned.lp = gp_device_lp{};
  });
  if (rc != GP_OK) owned.free();
  return rc;
}

void gp_free_batch_summary(gp_batch_summary* s) {
  if (!s) return;
  gp_free_summary(&s->whole);
  BatchOwned owned;
  owned.lp = s->assembled;
  owned.free();
  std::free(s->row_offsets);
  std::free(s->col_offsets);
  std::free(s->reduced_row_offsets);
  std::free(s->reduced_col_offsets);
  std::free(s->reduced_nnz_offsets);
  std::free(s->obj_constant);
  std::free(s->status);
  std::memset(s, 0, sizeof(*s));
}

}  // extern "C"
