// gp_primitives.hip -- device-wide primitives of the presolve engine:
// scans, prefix maps, the fixed-order global sum, CSR transpose and
// compaction, live statistics and row activity summaries.
#include "gp_common.hpp"

#include <mutex>

namespace gp {

// ---- errors / pool / clock ---------------------------------------------------------
namespace {
thread_local std::string g_last_error;
}

Tuning& tuning() {
  static Tuning t;
  return t;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error_cstr() { return g_last_error.c_str(); }

Pool& Pool::instance() {
  static Pool pool;
  return pool;
}

void* Pool::get(size_t bytes) {
  auto it = free_.find(bytes);
  if (it != free_.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    cached_ -= bytes;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // Release cached blocks and retry once before reporting exhaustion.
    (void)hipGetLastError();
    release_cached();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(GP_ERR_ALLOCATION, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
  }
  size_of_[p] = bytes;
  return p;
}

void Pool::put(void* p) {
  if (!p) return;
  auto it = size_of_.find(p);
  if (it == size_of_.end()) return;  // not a pool block
  free_[it->second].push_back(p);
  cached_ += it->second;
}

void Pool::release_cached() {
  if (cached_ == 0) return;
  (void)hipDeviceSynchronize();
  for (auto& kv : free_) {
    for (void* p : kv.second) {
      (void)hipFree(p);
      size_of_.erase(p);
    }
    kv.second.clear();
  }
  cached_ = 0;
}

Pool::~Pool() {}

KernelClock& kclock() {
  static KernelClock c;
  return c;
}

hipEvent_t KernelClock::take() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e;
  GP_CHECK(hipEventCreate(&e));
  return e;
}

void KernelClock::start_segment(int family) {
  Segment g{family, take(), take()};
  GP_CHECK(hipEventRecord(g.a, 0));
  segments.push_back(g);
}

void KernelClock::stop_segment() {
  (void)hipEventRecord(segments.back().b, 0);
}

int KernelClock::begin(int family) {
  if (!enabled) return -1;
  if (!open.empty()) stop_segment();
  open.push_back(family);
  launches[family] += 1;
  start_segment(family);
  return family;
}

void KernelClock::end(int token, double algorithmic_bytes) {
  if (token < 0 || !enabled || open.empty()) return;
  stop_segment();
  bytes[token] += algorithmic_bytes;
  open.pop_back();
  if (!open.empty()) start_segment(open.back());
}

void KernelClock::flush() {
  if (segments.empty()) return;
  GP_CHECK(hipDeviceSynchronize());
  for (auto& g : segments) {
    float t = 0.f;
    GP_CHECK(hipEventElapsedTime(&t, g.a, g.b));
    ms[g.family] += t;
    pool.push_back(g.a);
    pool.push_back(g.b);
  }
  segments.clear();
}

KernelClock::~KernelClock() {}

// ---- flag words ------------------------------------------------------------------------
namespace {
struct FlagBuffers {
  int32_t* buf[2] = {nullptr, nullptr};
  int cur = 0, next = kFlagSlots;  // next slot of buf[cur]; kFlagSlots = switch buffers
};
std::map<int, FlagBuffers> g_flag_buffers;
}  // namespace

int32_t* flag_slot() {
  int dev = 0;
  GP_CHECK(hipGetDevice(&dev));
  FlagBuffers& f = g_flag_buffers[dev];
  const size_t bytes = sizeof(int32_t) * 8 * kFlagSlots;
  if (f.next == kFlagSlots) {
    f.cur ^= 1;
    if (!f.buf[f.cur]) GP_CHECK(hipMalloc(&f.buf[f.cur], bytes));
    GP_CHECK(hipMemsetAsync(f.buf[f.cur], 0, bytes, 0));
    f.next = 0;
  }
  return f.buf[f.cur] + 8 * (f.next++);
}

// ---- scans -------------------------------------------------------------------------
namespace {
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <class T>
__device__ T block_exclusive_scan(T v, T* lds, T* total) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    T add = t >= off ? lds[t - off] : T(0);
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  const T incl = lds[t];
  *total = lds[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

template <class T>
__global__ void k_tile_reduce(const T* in, int64_t n, T* sums) {
  __shared__ T lds[kScanThreads];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  T acc = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kScanThreads + threadIdx.x;
    if (i < n) acc += in[i];
  }
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kScanThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = lds[0];
}

// Scan one tile in place; `prefix` (exclusive, per tile) may be null.
template <class T>
__global__ void k_tile_scan(T* data, int64_t n, const T* prefix, int inclusive) {
  __shared__ T tile[kScanTile];
  __shared__ T lds[kScanThreads];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kScanThreads + threadIdx.x;
    tile[k * kScanThreads + threadIdx.x] = i < n ? data[i] : T(0);
  }
  __syncthreads();
  T local[kScanItems];
  T sum = 0;
  for (int k = 0; k < kScanItems; ++k) {
    local[k] = tile[threadIdx.x * kScanItems + k];
    sum += local[k];
  }
  T total;
  T run = block_exclusive_scan<T>(sum, lds, &total);
  if (prefix) run += prefix[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    const T v = local[k];
    tile[threadIdx.x * kScanItems + k] = inclusive ? run + v : run;
    run += v;
  }
  __syncthreads();
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kScanThreads + threadIdx.x;
    if (i < n) data[i] = tile[k * kScanThreads + threadIdx.x];
  }
}

template <class T>
void scan_impl(T* data, int64_t n, bool inclusive) {
  if (n <= 0) return;
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_tile_scan<T>, dim3(1), dim3(kScanThreads), 0, 0, data, n,
                       static_cast<const T*>(nullptr), inclusive ? 1 : 0);
    GP_LAUNCH_CHECK("tile scan");
    return;
  }
  DVec<T> sums(static_cast<size_t>(tiles));
  hipLaunchKernelGGL(k_tile_reduce<T>, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0, 0,
                     static_cast<const T*>(data), n, sums.data());
  GP_LAUNCH_CHECK("tile reduce");
  scan_impl<T>(sums.data(), tiles, false);
  hipLaunchKernelGGL(k_tile_scan<T>, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0, 0,
                     data, n, static_cast<const T*>(sums.data()), inclusive ? 1 : 0);
  GP_LAUNCH_CHECK("tile scan");
}

// The total of an exclusive scan (last output + last input) is formed on the
// device, so a scan costs at most one host round trip, when the caller reads it.
template <class T>
__global__ void k_save_last(const T* data, int64_t n, T* slot) {
  *slot = data[n - 1];
}
template <class T>
__global__ void k_scan_total(const T* data, int64_t n, const T* saved, int64_t* out) {
  *out = static_cast<int64_t>(data[n - 1]) + static_cast<int64_t>(*saved);
}
__global__ void k_zero_i64(int64_t* out) { *out = 0; }

template <class T>
void exclusive_scan_dev_total(T* data, int64_t n, int64_t* dev_total) {
  if (n <= 0) {
    hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(1), 0, 0, dev_total);
    GP_LAUNCH_CHECK("scan total");
    return;
  }
  DVec<T> saved(1);
  hipLaunchKernelGGL(k_save_last<T>, dim3(1), dim3(1), 0, 0, static_cast<const T*>(data), n, saved.data());
  scan_impl<T>(data, n, false);
  hipLaunchKernelGGL(k_scan_total<T>, dim3(1), dim3(1), 0, 0, static_cast<const T*>(data), n,
                     static_cast<const T*>(saved.data()), dev_total);
  GP_LAUNCH_CHECK("scan total");
}

__global__ void k_mask_to_i32(const uint8_t* keep, int32_t n, int32_t* out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = keep[i] != 0;
}

__global__ void k_old_to_new(const uint8_t* keep, int32_t n, int32_t* map) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) map[i] = keep[i] ? map[i] : -1;
}
}  // namespace

void inclusive_scan_i32(int32_t* data, int64_t n) { scan_impl<int32_t>(data, n, true); }
void exclusive_scan_i32(int32_t* data, int64_t n) {
  if (n > 0) scan_impl<int32_t>(data, n, false);
}

void exclusive_scan_i64(int64_t* data, int64_t n, int64_t* total) {
  if (n <= 0) {
    if (total) *total = 0;
    return;
  }
  if (!total) {
    scan_impl<int64_t>(data, n, false);
    return;
  }
  DVec<int64_t> t(1);
  exclusive_scan_dev_total<int64_t>(data, n, t.data());
  *total = read1(t.data());
}

void exclusive_scan_i64_device_total(int64_t* data, int64_t n, int64_t* dev_total) {
  exclusive_scan_dev_total<int64_t>(data, n, dev_total);
}

// ---- three exclusive scans of the same length at once -----------------------------------
// One reduce pass over the three arrays' tiles, one workgroup that scans the
// three tile-sum arrays and writes the totals, one scan pass: three launches
// instead of about seven per array.
namespace {
constexpr int kScan3MaxTiles = kScanThreads * 64;

__global__ void k_tile_reduce3(const int64_t* a0, const int64_t* a1, const int64_t* a2, int64_t n, int64_t tiles,
                               int64_t* sums) {
  __shared__ int64_t lds[kScanThreads];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  const int64_t* arr[3] = {a0, a1, a2};
  for (int a = 0; a < 3; ++a) {
    int64_t acc = 0;
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t i = base + k * kScanThreads + threadIdx.x;
      if (i < n) acc += arr[a][i];
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kScanThreads / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) sums[a * tiles + blockIdx.x] = lds[0];
    __syncthreads();
  }
}

// one workgroup: exclusive scan of each array's tile sums in place, totals out
__global__ void k_scan_sums3(int64_t* sums, int64_t tiles, int64_t* totals) {
  __shared__ int64_t lds[kScanThreads];
  const int64_t per = (tiles + kScanThreads - 1) / kScanThreads;
  const int64_t b0 = threadIdx.x * per;
  for (int a = 0; a < 3; ++a) {
    int64_t* x = sums + a * tiles;
    int64_t local = 0;
    for (int64_t k = b0; k < b0 + per && k < tiles; ++k) local += x[k];
    int64_t total;
    int64_t run = block_exclusive_scan<int64_t>(local, lds, &total);
    for (int64_t k = b0; k < b0 + per && k < tiles; ++k) {
      const int64_t v = x[k];
      x[k] = run;
      run += v;
    }
    if (threadIdx.x == 0) totals[a] = total;
    __syncthreads();
  }
}

__global__ void k_tile_scan3(int64_t* a0, int64_t* a1, int64_t* a2, int64_t n, int64_t tiles, const int64_t* prefix) {
  __shared__ int64_t tile[kScanTile];
  __shared__ int64_t lds[kScanThreads];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int64_t* arr[3] = {a0, a1, a2};
  for (int a = 0; a < 3; ++a) {
    int64_t* data = arr[a];
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t i = base + k * kScanThreads + threadIdx.x;
      tile[k * kScanThreads + threadIdx.x] = i < n ? data[i] : 0;
    }
    __syncthreads();
    int64_t local[kScanItems];
    int64_t sum = 0;
    for (int k = 0; k < kScanItems; ++k) {
      local[k] = tile[threadIdx.x * kScanItems + k];
      sum += local[k];
    }
    int64_t total;
    int64_t run = block_exclusive_scan<int64_t>(sum, lds, &total) + prefix[a * tiles + blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
      tile[threadIdx.x * kScanItems + k] = run;
      run += local[k];
    }
    __syncthreads();
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t i = base + k * kScanThreads + threadIdx.x;
      if (i < n) data[i] = tile[k * kScanThreads + threadIdx.x];
    }
    __syncthreads();
  }
}
}  // namespace

void exclusive_scan3_i64_device_total(int64_t* a0, int64_t* a1, int64_t* a2, int64_t n, int64_t* dev_totals) {
  const int64_t tiles = n > 0 ? (n + kScanTile - 1) / kScanTile : 0;
  if (tiles == 0 || tiles > kScan3MaxTiles) {
    exclusive_scan_i64_device_total(a0, n, dev_totals);
    exclusive_scan_i64_device_total(a1, n, dev_totals + 1);
    exclusive_scan_i64_device_total(a2, n, dev_totals + 2);
    return;
  }
  DVec<int64_t> sums(static_cast<size_t>(3 * tiles));
  hipLaunchKernelGGL(k_tile_reduce3, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0, 0,
                     static_cast<const int64_t*>(a0), static_cast<const int64_t*>(a1),
                     static_cast<const int64_t*>(a2), n, tiles, sums.data());
  hipLaunchKernelGGL(k_scan_sums3, dim3(1), dim3(kScanThreads), 0, 0, sums.data(), tiles, dev_totals);
  hipLaunchKernelGGL(k_tile_scan3, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0, 0, a0, a1, a2, n, tiles,
                     static_cast<const int64_t*>(sums.data()));
  GP_LAUNCH_CHECK("three scans");
}

int64_t exclusive_scan_i32_total(int32_t* data, int64_t n) {
  if (n <= 0) return 0;
  DVec<int64_t> t(1);
  exclusive_scan_dev_total<int32_t>(data, n, t.data());
  return read1(t.data());
}

int32_t build_old_to_new(const uint8_t* keep, int32_t n, int32_t* old_to_new) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_mask_to_i32, dim3(grid_for(n)), dim3(kBlock), 0, 0, keep, n, old_to_new);
  GP_LAUNCH_CHECK("mask to i32");
  const int64_t total = exclusive_scan_i32_total(old_to_new, n);
  hipLaunchKernelGGL(k_old_to_new, dim3(grid_for(n)), dim3(kBlock), 0, 0, keep, n, old_to_new);
  GP_LAUNCH_CHECK("old to new");
  return static_cast<int32_t>(total);
}

// ---- fixed-order global sum ------------------------------------------------------------
namespace {
__global__ void k_det_sum_chunks(const double* x, int64_t n, double* partial) {
  __shared__ double y[512];
  const int t = threadIdx.x;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * 1024;
  const double a = base + t < n ? x[base + t] : 0.0;
  const double b = base + t + 512 < n ? x[base + t + 512] : 0.0;
  y[t] = a + b;
  __syncthreads();
  for (int s = 256; s > 0; s >>= 1) {
    if (t < s) y[t] = y[t] + y[t + s];
    __syncthreads();
  }
  if (t == 0) partial[blockIdx.x] = y[0];
}
}  // namespace

double det_sum(const double* x, int64_t n) {
  if (n <= 0) return 0.0;
  int64_t chunks = (n + 1023) / 1024;
  DVec<double> a(static_cast<size_t>(chunks));
  hipLaunchKernelGGL(k_det_sum_chunks, dim3(static_cast<unsigned>(chunks)), dim3(512), 0, 0, x, n, a.data());
  GP_LAUNCH_CHECK("det sum");
  DVec<double> b;
  while (chunks > 1) {
    const int64_t next = (chunks + 1023) / 1024;
    b.resize(static_cast<size_t>(next));
    hipLaunchKernelGGL(k_det_sum_chunks, dim3(static_cast<unsigned>(next)), dim3(512), 0, 0,
                       static_cast<const double*>(a.data()), chunks, b.data());
    GP_LAUNCH_CHECK("det sum");
    std::swap(a, b);
    chunks = next;
  }
  return read1(a.data());
}

// ---- small vector kernels ----------------------------------------------------------------
namespace {
__global__ void k_fill_u8(uint8_t* p, uint8_t v, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
template <class T>
__global__ void k_gather_compact(T* dst, const T* src, const int32_t* o2n, int32_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    const int32_t k = o2n[i];
    if (k >= 0) dst[k] = src[i];
  }
}
// grid-stride, one atomic per wavefront (a block per 256 bytes would hit the
// counter from tens of thousands of blocks)
__global__ void k_count_u8(const uint8_t* mask, int32_t n, int32_t* out) {
  int32_t c = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    c += mask[i] ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if (__lane_id() == 0 && c) atomicAdd(out, c);
}
}  // namespace

void fill_u8(uint8_t* p, uint8_t v, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n)), dim3(kBlock), 0, 0, p, v, n);
  GP_LAUNCH_CHECK("fill u8");
}

void gather_compact(double* dst, const double* src, const int32_t* o2n, int32_t n_old) {
  if (n_old <= 0) return;
  hipLaunchKernelGGL(k_gather_compact<double>, dim3(grid_for(n_old)), dim3(kBlock), 0, 0, dst, src, o2n, n_old);
  GP_LAUNCH_CHECK("gather compact");
}

void gather_compact_i32(int32_t* dst, const int32_t* src, const int32_t* o2n, int32_t n_old) {
  if (n_old <= 0) return;
  hipLaunchKernelGGL(k_gather_compact<int32_t>, dim3(grid_for(n_old)), dim3(kBlock), 0, 0, dst, src, o2n, n_old);
  GP_LAUNCH_CHECK("gather compact i32");
}

int32_t count_u8(const uint8_t* mask, int32_t n) {
  if (n <= 0) return 0;
  DVec<int32_t> c(1);
  GP_CHECK(hipMemsetAsync(c.data(), 0, sizeof(int32_t), 0));
  hipLaunchKernelGGL(k_count_u8, dim3(grid_capped(n, kBlock, 1024)), dim3(kBlock), 0, 0, mask, n, c.data());
  GP_LAUNCH_CHECK("count u8");
  return read1(c.data());
}

// ---- transpose ----------------------------------------------------------------------------
// Histogram of column counts, exclusive scan, atomic scatter, then each output
// row is sorted by input row: the unique CSR of A^T with increasing row order.
namespace {
__global__ void k_col_hist(const int32_t* col, int64_t nnz, int32_t* cnt) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < nnz) atomicAdd(cnt + col[p], 1);
}
// G lanes per row: coalesced reads of the row, one cursor atomic per entry.
// The order inside a column is restored by the segment sorts below.
template <int G>
__global__ void __launch_bounds__(kBlock) k_scatter_rows(const gp_offset* rp, const int32_t* col, const double* val,
                                                          int32_t m, int32_t* cursor, int32_t* at_col,
                                                          double* at_val) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; i < m; i += groups) {
    const gp_offset p1 = rp[i + 1];
    for (gp_offset p = rp[i] + lane; p < p1; p += G) {
      const int32_t q = atomicAdd(cursor + ld_stream(col + p), 1);
      at_col[q] = static_cast<int32_t>(i);
      at_val[q] = ld_stream(val + p);
    }
  }
}
constexpr int kShortSegment = 32;
constexpr int kLongSegment = 8192;
// Segments of at most kShortSegment entries: one lane each, insertion sort.
// Longer ones are listed for k_sort_long (up to kLongSegment entries) or for a
// device radix sort of their own (longer still: dense columns).
__global__ void k_sort_segments(const gp_offset* rp, int32_t n, int32_t* idx, double* val, int32_t* long_list,
                                int32_t* counts) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const gp_offset a = rp[j], b = rp[j + 1];
  if (b - a > kLongSegment) {
    long_list[n - 1 - atomicAdd(counts + 1, 1)] = static_cast<int32_t>(j);  // huge: listed from the end
    return;
  }
  if (b - a > kShortSegment) {
    long_list[atomicAdd(counts, 1)] = static_cast<int32_t>(j);
    return;
  }
  for (gp_offset p = a + 1; p < b; ++p) {
    const int32_t key = idx[p];
    const double v = val[p];
    gp_offset q = p - 1;
    while (q >= a && idx[q] > key) {
      idx[q + 1] = idx[q];
      val[q + 1] = val[q];
      --q;
    }
    idx[q + 1] = key;
    val[q + 1] = v;
  }
}
// One workgroup per long segment, in LDS chunks: every entry is placed by its
// rank, the number of entries with a smaller key or an equal key before it.
constexpr int kLongChunk = 2048;
__global__ void __launch_bounds__(256) k_sort_long(const gp_offset* rp, const int32_t* long_list,
                                                   const int32_t* long_count, int32_t* idx, double* val,
                                                   int32_t* tmp_idx, double* tmp_val) {
  __shared__ int32_t keys[kLongChunk];
  const int32_t total = *long_count;
  for (int32_t s = blockIdx.x; s < total; s += gridDim.x) {
    const int32_t j = long_list[s];
    const gp_offset a = rp[j], b = rp[j + 1];
    const int64_t len = b - a;
    // ranks against every chunk of the segment, accumulated per element
    for (int64_t e0 = 0; e0 < len; e0 += 256) {
      const int64_t e = e0 + threadIdx.x;
      const int32_t mine = e < len ? idx[a + e] : 0;
      int64_t rank = 0;
      for (int64_t c0 = 0; c0 < len; c0 += kLongChunk) {
        const int64_t c1 = c0 + kLongChunk < len ? c0 + kLongChunk : len;
        __syncthreads();
        for (int64_t k = c0 + threadIdx.x; k < c1; k += 256) keys[k - c0] = idx[a + k];
        __syncthreads();
        if (e < len)  // stable: equal keys keep their order
          for (int64_t k = 0; k < c1 - c0; ++k) rank += (keys[k] < mine || (keys[k] == mine && c0 + k < e)) ? 1 : 0;
      }
      if (e < len) {
        tmp_idx[a + rank] = mine;
        tmp_val[a + rank] = val[a + e];
      }
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < len; e += 256) {
      idx[a + e] = tmp_idx[a + e];
      val[a + e] = tmp_val[a + e];
    }
    __syncthreads();
  }
}
__global__ void k_widen_keys(const int32_t* idx, int64_t len, uint64_t* keys) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < len) keys[k] = static_cast<uint32_t>(idx[k]);
}
__global__ void k_narrow_keys(const uint64_t* keys, int64_t len, int32_t* idx) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < len) idx[k] = static_cast<int32_t>(keys[k]);
}
__global__ void k_i32_to_offset_shift(const int32_t* src, int32_t n, gp_offset* dst) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n) dst[j] = src[j];
}
}  // namespace

// Sort-based transpose: key (column << row_bits | row) per entry, one stable
// LSD radix sort of (key, value) pairs, then row indices and the column
// starts read off the sorted keys.  Entries of a column come out in row
// order, so the result is the unique sorted CSR of A^T.
// Key = column, payload = (row, value): the entries arrive in CSR order and
// the radix sort is stable, so within a column they stay in increasing row
// order -- the unique CSR of A^T -- with only the column bits sorted.
template <int G>
__global__ void __launch_bounds__(kBlock) k_transpose_keys(const gp_offset* rp, const int32_t* col, const double* val,
                                                            int32_t m, uint32_t* keys, RowVal* rv) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; i < m; i += groups) {
    const gp_offset p1 = rp[i + 1];
    for (gp_offset p = rp[i] + lane; p < p1; p += G) {
      keys[p] = static_cast<uint32_t>(ld_stream(col + p));
      rv[p] = RowVal{static_cast<int32_t>(i), 0, ld_stream(val + p)};
    }
  }
}
__global__ void k_transpose_out(const uint32_t* keys, const RowVal* rv, int64_t nnz, int32_t n, int32_t* at_col,
                                double* at_val, gp_offset* at_rp) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  const int64_t c = static_cast<int64_t>(keys[p]);
  const RowVal e = rv[p];
  at_col[p] = e.row;
  at_val[p] = e.val;
  const int64_t prev = p > 0 ? static_cast<int64_t>(keys[p - 1]) : -1;
  for (int64_t j = prev + 1; j <= c; ++j) at_rp[j] = static_cast<gp_offset>(p);  // columns starting here
  if (p == nnz - 1)
    for (int64_t j = c + 1; j <= n; ++j) at_rp[j] = static_cast<gp_offset>(nnz);  // trailing empty columns
}

int bits_for(int64_t v) {  // bits needed for values 0..v-1 (at least 1)
  int b = 1;
  while (b < 62 && (int64_t{1} << b) < v) ++b;
  return b;
}

void transpose_csr(const gp_csr& A, gp_csr& AT) {
  const int32_t m = A.rows, n = A.cols;
  if (A.nnz > 0 && static_cast<int64_t>(A.nnz) <= INT32_MAX) {
    // CSR in, keys + payload (4 + 16 B per entry) out and back, CSR of A^T out
    TimedScope ts(GP_KT_TRANSPOSE, 2.0 * (12.0 * A.nnz) + 4.0 * (A.rows + 1) + 4.0 * (A.cols + 1));
    const int64_t nnz = A.nnz;
    AT.rows = n;
    AT.cols = m;
    AT.nnz = A.nnz;
    AT.row_ptr = pool_alloc<gp_offset>(static_cast<size_t>(n) + 1);
    AT.col_idx = pool_alloc<int32_t>(static_cast<size_t>(nnz));
    AT.vals = pool_alloc<double>(static_cast<size_t>(nnz));
    DVec<uint32_t> keys(static_cast<size_t>(nnz));
    DVec<RowVal> rv(static_cast<size_t>(nnz));
    const unsigned grid = grid_capped(static_cast<int64_t>(m) * 8, kBlock, 16384);
    switch (lanes_for_mean(m ? static_cast<double>(nnz) / m : 0.0)) {
      case 2: hipLaunchKernelGGL(k_transpose_keys<2>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, keys.data(), rv.data()); break;
      case 4: hipLaunchKernelGGL(k_transpose_keys<4>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, keys.data(), rv.data()); break;
      case 8: hipLaunchKernelGGL(k_transpose_keys<8>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, keys.data(), rv.data()); break;
      case 16: hipLaunchKernelGGL(k_transpose_keys<16>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, keys.data(), rv.data()); break;
      default: hipLaunchKernelGGL(k_transpose_keys<64>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, keys.data(), rv.data()); break;
    }
    GP_LAUNCH_CHECK("transpose keys");
    DVec<uint32_t> keys2;
    DVec<RowVal> rv2;
    const uint32_t* sorted_keys = nullptr;
    const RowVal* sorted_rv = nullptr;
    sort_pairs_u32_rowval(keys.data(), rv.data(), nnz, bits_for(n), keys2, rv2, &sorted_keys, &sorted_rv);
    hipLaunchKernelGGL(k_transpose_out, dim3(grid_for(nnz)), dim3(kBlock), 0, 0, sorted_keys, sorted_rv, nnz, n,
                       AT.col_idx, AT.vals, AT.row_ptr);
    GP_LAUNCH_CHECK("transpose columns");
    return;
  }
  transpose_csr_scatter(A, AT);
}

void transpose_csr_scatter(const gp_csr& A, gp_csr& AT) {
  TimedScope ts(GP_KT_TRANSPOSE, 2.0 * (12.0 * A.nnz) + 4.0 * (A.rows + 1) + 4.0 * (A.cols + 1));
  const int32_t m = A.rows, n = A.cols;
  const int64_t nnz = A.nnz;
  AT.rows = n;
  AT.cols = m;
  AT.nnz = A.nnz;
  AT.row_ptr = pool_alloc<gp_offset>(static_cast<size_t>(n) + 1);
  AT.col_idx = pool_alloc<int32_t>(static_cast<size_t>(nnz));
  AT.vals = pool_alloc<double>(static_cast<size_t>(nnz));
  DVec<int32_t> cnt(static_cast<size_t>(n) + 1);
  GP_CHECK(hipMemsetAsync(cnt.data(), 0, sizeof(int32_t) * (static_cast<size_t>(n) + 1), 0));
  if (nnz > 0) {
    hipLaunchKernelGGL(k_col_hist, dim3(grid_for(nnz)), dim3(kBlock), 0, 0, A.col_idx, nnz, cnt.data());
    GP_LAUNCH_CHECK("col hist");
  }
  exclusive_scan_i32_total(cnt.data(), static_cast<int64_t>(n) + 1);
  hipLaunchKernelGGL(k_i32_to_offset_shift, dim3(grid_for(n + 1)), dim3(kBlock), 0, 0, cnt.data(), n + 1, AT.row_ptr);
  GP_LAUNCH_CHECK("row ptr");
  if (nnz > 0) {
    const unsigned grid = grid_capped(static_cast<int64_t>(m) * 8, kBlock, 16384);
    switch (lanes_for_mean(m ? static_cast<double>(nnz) / m : 0.0)) {
      case 2: hipLaunchKernelGGL(k_scatter_rows<2>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, cnt.data(), AT.col_idx, AT.vals); break;
      case 4: hipLaunchKernelGGL(k_scatter_rows<4>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, cnt.data(), AT.col_idx, AT.vals); break;
      case 8: hipLaunchKernelGGL(k_scatter_rows<8>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, cnt.data(), AT.col_idx, AT.vals); break;
      case 16: hipLaunchKernelGGL(k_scatter_rows<16>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, cnt.data(), AT.col_idx, AT.vals); break;
      default: hipLaunchKernelGGL(k_scatter_rows<64>, dim3(grid), dim3(kBlock), 0, 0, A.row_ptr, A.col_idx, A.vals, m, cnt.data(), AT.col_idx, AT.vals); break;
    }
    GP_LAUNCH_CHECK("scatter rows");
    DVec<int32_t> long_list(static_cast<size_t>(n) + 1), counts(2);
    GP_CHECK(hipMemsetAsync(counts.data(), 0, 2 * sizeof(int32_t), 0));
    hipLaunchKernelGGL(k_sort_segments, dim3(grid_for(n)), dim3(kBlock), 0, 0, AT.row_ptr, n, AT.col_idx, AT.vals,
                       long_list.data(), counts.data());
    GP_LAUNCH_CHECK("sort segments");
    DVec<int32_t> tmp_idx(static_cast<size_t>(nnz));
    DVec<double> tmp_val(static_cast<size_t>(nnz));
    hipLaunchKernelGGL(k_sort_long, dim3(grid_capped(static_cast<int64_t>(n) * 256, 256, 2048)), dim3(256), 0, 0,
                       AT.row_ptr, long_list.data(), counts.data(), AT.col_idx, AT.vals, tmp_idx.data(),
                       tmp_val.data());
    GP_LAUNCH_CHECK("sort long segments");
    int32_t huge = 0;
    d2h(&huge, counts.data() + 1, 1);
    if (huge > 0) {
      std::vector<int32_t> cols(static_cast<size_t>(huge));
      d2h(cols.data(), long_list.data() + (n - huge), static_cast<size_t>(huge));
      std::vector<gp_offset> rp(static_cast<size_t>(n) + 1);
      d2h(rp.data(), AT.row_ptr, rp.size());
      DVec<uint64_t> keys;
      for (const int32_t j : cols) {
        const int64_t a = rp[j], len = rp[j + 1] - rp[j];
        keys.resize(static_cast<size_t>(len));
        hipLaunchKernelGGL(k_widen_keys, dim3(grid_for(len)), dim3(kBlock), 0, 0, AT.col_idx + a, len, keys.data());
        sort_pairs_u64_f64(keys.data(), AT.vals + a, len, 32);
        hipLaunchKernelGGL(k_narrow_keys, dim3(grid_for(len)), dim3(kBlock), 0, 0, keys.data(), len, AT.col_idx + a);
        GP_LAUNCH_CHECK("sort dense column");
      }
    }
  }
}

void csr_pool_free(gp_csr& A) {
  pool_free(A.row_ptr);
  pool_free(A.col_idx);
  pool_free(A.vals);
  A.row_ptr = nullptr;
  A.col_idx = nullptr;
  A.vals = nullptr;
  A.rows = A.cols = 0;
  A.nnz = 0;
}

// ---- compaction ------------------------------------------------------------------------------
namespace {
// New index of column j: its rank among the kept columns, from one 8-byte
// record per 32 columns (x: the kept bits, y: kept columns before the word),
// so an entry costs one gather into an L2-resident table (2 MB at 8M
// columns) instead of an HBM gather into an n-entry old-to-new map or two
// gathers into separate bit and prefix arrays.  -1 for a dropped column.
struct ColRank {
  const uint2* words;  // null: keep every column, identity map
};
__global__ void k_rank_words(const uint32_t* bits, const int32_t* prefix, int64_t nw, uint2* out) {
  const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w < nw) out[w] = make_uint2(bits[w], static_cast<uint32_t>(prefix[w]));
}
__global__ void k_word_popc(const uint32_t* words, int64_t nw, int32_t* out) {
  const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w < nw) out[w] = __popc(words[w]);
}

// Entry-parallel stable compaction.  The output is the kept entries in input
// order, so a kept entry's position is the number of kept entries before it.
// One workgroup per tile of kCompactThreads * P consecutive entries; a
// lane's k-th entry is t0 + k * kCompactThreads + lane id, so every load and
// store stays coalesced.  tile_row[t] = the row holding the tile's first
// entry (one lane per row).
constexpr int kCompactThreads = 1024;
__global__ void k_tile_rows(const gp_offset* rp, int32_t m, int64_t tile, int32_t* tile_row) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int64_t a = rp[r], b = rp[r + 1];
  if (a >= b) return;
  for (int64_t t = (a + tile - 1) / tile; t * tile < b; ++t) tile_row[t] = static_cast<int32_t>(r);
}

// Tile status words for the decoupled look-back: a flag in the top two bits
// (1: the tile's own kept count, 2: the inclusive count of tiles 0..t) and
// the count below.  A look-back that spins past kSpinLimit polls gives up,
// marks the error word and lets the kernel finish; the host then raises an
// error.
constexpr unsigned long long kStatAgg = 1ull << 62, kStatIncl = 2ull << 62, kStatMask = (1ull << 62) - 1;
constexpr int64_t kSpinLimit = int64_t{1} << 26;
__device__ inline void publish(unsigned long long* st, unsigned long long v) {
  __hip_atomic_store(st, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Kept entries before tile t, looked back over by a whole wavefront: lane L
// reads the status word of tile q - L, so one poll covers 64 predecessors.
// The window's entries up to the nearest inclusive one are summed when none
// of them is still unpublished; without an inclusive entry the whole window
// is summed and the window moves 64 tiles back.  Called by every lane of one
// wavefront; every lane returns the same value.
__device__ inline int64_t wave_look_back(unsigned long long* status, int64_t t, int32_t* err) {
  const int lane = static_cast<int>(__lane_id());
  int64_t excl = 0, spins = 0;
  for (int64_t q = t - 1; q >= 0;) {
    const int64_t idx = q - lane;
    // positions before tile 0 read as an inclusive count of 0
    const unsigned long long v =
        idx >= 0 ? __hip_atomic_load(status + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kStatIncl;
    const unsigned long long incl = __ballot((v >> 62) == 2);
    const unsigned long long open = __ballot((v >> 62) == 0);
    // lanes up to and including the nearest inclusive entry, else all
    const unsigned long long upto = incl ? ((incl & (~incl + 1ull)) << 1) - 1ull : ~0ull;
    if (open & upto) {
      if (++spins > kSpinLimit) {
        if (lane == 0) atomicOr(err, 1);
        return 0;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    int64_t x = ((upto >> lane) & 1ull) ? static_cast<int64_t>(v & kStatMask) : 0;
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    excl += x;
    if (incl) break;
    q -= 64;
  }
  return excl;
}

// Single pass.  Workgroups take tiles in ticket order from an atomic counter,
// so every tile a workgroup waits for belongs to a workgroup that is already
// running and publishes its count before it waits on anything.  A tile keeps
// in LDS a bit per entry of a dropped row, the wavefront ballots of its keep
// flags and their exclusive prefix per 64-entry group (groups in position
// order g = k * waves + wave); an entry's output position is base +
// prefix(g) + the kept lanes of g before it, and the row pointer of a kept
// row starting at tile position a is the same expression at a.
// ctl[0]: tile ticket, ctl[1]: error word, ctl[2]: total kept entries,
// ctl[3..]: tile status words; all zero at launch.
template <int P>
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_tiles(const gp_offset* rp, const int32_t* col, const double* val, int32_t m, int64_t nnz, int64_t tiles,
                    const int32_t* tile_row, const uint8_t* keep_row, const int32_t* row_o2n, ColRank col_rank,
                    int32_t m_new, unsigned long long* ctl, gp_offset* out_rp, int32_t* out_col, double* out_val) {
  constexpr int kT = kCompactThreads, kTile = kT * P, kWaves = kT / 64;
  constexpr int kGroups = kTile / 64, kWords = kTile / 32;
  constexpr int kPerLane = kGroups / 64;  // groups of the prefix scan per lane of wave 0
  static_assert(kGroups % 64 == 0, "prefix scan layout");
  __shared__ uint32_t s_drop[kWords];
  __shared__ unsigned long long s_bal[kGroups];
  __shared__ int32_t s_pre[kGroups + 1];
  __shared__ int64_t s_tile, s_base;
  const int lane = static_cast<int>(__lane_id()), wave = threadIdx.x / 64;
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long* status = ctl + 3;
  int32_t* err = reinterpret_cast<int32_t*>(ctl + 1);
  if (threadIdx.x == 0) s_tile = static_cast<int64_t>(atomicAdd(ctl, 1ull));
  for (int w = threadIdx.x; w < kWords; w += kT) s_drop[w] = 0u;
  __syncthreads();
  const int64_t tile = s_tile;
  if (tile >= tiles) return;
  const int64_t t0 = tile * kTile;
  const int64_t t1 = t0 + kTile < nnz ? t0 + kTile : nnz;
  // every load of the lane issued before any is used (out-of-tile lanes read
  // entry t0 and are masked below)
  int32_t j[P];
  double v[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t p = t0 + static_cast<int64_t>(k) * kT + threadIdx.x;
    j[k] = ld_stream(col + (p < t1 ? p : t0));
  }
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t p = t0 + static_cast<int64_t>(k) * kT + threadIdx.x;
    v[k] = ld_stream(val + (p < t1 ? p : t0));
  }
  const int32_t i0 = tile_row[tile];
  // the tile's entries of dropped rows
  if (keep_row) {
    for (int32_t r = i0 + static_cast<int32_t>(threadIdx.x); r < m && rp[r] < t1; r += kT) {
      if (keep_row[r]) continue;
      const int64_t a = (rp[r] > t0 ? rp[r] : t0) - t0, b = (rp[r + 1] < t1 ? rp[r + 1] : t1) - t0;
      for (int64_t q = a; q < b;) {
        const int64_t w = q >> 5, e = ((w + 1) << 5) < b ? ((w + 1) << 5) : b;
        const uint32_t below_e = (e - (w << 5)) == 32 ? ~0u : ((1u << (e - (w << 5))) - 1u);
        atomicOr(&s_drop[w], below_e & ~((1u << (q & 31)) - 1u));
        q = e;
      }
    }
  }
  uint2 rank[P];
#pragma unroll
  for (int k = 0; k < P; ++k) rank[k] = col_rank.words ? col_rank.words[j[k] >> 5] : make_uint2(~0u, 0u);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int32_t pos = k * kT + static_cast<int32_t>(threadIdx.x);
    const bool in = t0 + pos < t1 && !((s_drop[pos >> 5] >> (pos & 31)) & 1u);
    const uint32_t bit = 1u << (j[k] & 31);
    const int32_t nj = col_rank.words ? static_cast<int32_t>(rank[k].y) + __popc(rank[k].x & (bit - 1u)) : j[k];
    j[k] = in && (rank[k].x & bit) ? nj : -1;
    const unsigned long long b = __ballot(j[k] >= 0);
    if (lane == 0) s_bal[k * kWaves + wave] = b;
  }
  __syncthreads();
  if (wave == 0) {
    // exclusive prefix of the group counts, kPerLane consecutive groups per lane
    int32_t c[kPerLane];
    int32_t sum = 0;
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
      c[u] = __popcll(s_bal[lane * kPerLane + u]);
      sum += c[u];
    }
    int32_t incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    int32_t run = incl - sum;
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
      s_pre[lane * kPerLane + u] = run;
      run += c[u];
    }
    const int64_t own = __shfl(incl, 63, 64);
    int64_t excl = 0;
    if (tile == 0) {
      if (lane == 0) publish(status, kStatIncl | static_cast<unsigned long long>(own));
    } else {
      if (lane == 0) publish(status + tile, kStatAgg | static_cast<unsigned long long>(own));
      excl = wave_look_back(status, tile, err);
      if (lane == 0) publish(status + tile, kStatIncl | static_cast<unsigned long long>(excl + own));
    }
    if (lane == 0) {
      s_pre[kGroups] = static_cast<int32_t>(own);
      if (tile == tiles - 1) {
        ctl[2] = static_cast<unsigned long long>(excl + own);
        out_rp[m_new] = static_cast<gp_offset>(excl + own);
      }
      s_base = excl;
    }
  }
  __syncthreads();
  const int64_t base = s_base;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    if (j[k] < 0) continue;
    const int g = k * kWaves + wave;
    const int64_t at = base + s_pre[g] + __popcll(s_bal[g] & lt);
    out_col[at] = j[k];
    out_val[at] = v[k];
  }
  // row pointers of the kept rows starting in this tile (trailing empty rows
  // at nnz belong to the last tile; empty rows just before i0 start at t0)
  const bool last = t1 == nnz;
  for (int32_t r = i0 + static_cast<int32_t>(threadIdx.x); r < m && (rp[r] < t1 || (last && rp[r] == nnz)); r += kT) {
    const int64_t a = rp[r];
    if (a < t0 || (keep_row && !keep_row[r])) continue;
    int64_t at = s_pre[kGroups];
    if (a < t1) {
      const int32_t pos = static_cast<int32_t>(a - t0), g = pos >> 6;
      at = s_pre[g] + __popcll(s_bal[g] & ((1ull << (pos & 63)) - 1ull));
    }
    out_rp[row_o2n ? row_o2n[r] : r] = static_cast<gp_offset>(base + at);
  }
  if (threadIdx.x == 0)
    for (int32_t r = i0 - 1; r >= 0 && rp[r] == t0; --r)
      if (!keep_row || keep_row[r]) out_rp[row_o2n ? row_o2n[r] : r] = static_cast<gp_offset>(base);
}
}  // namespace

void compact_csr(const gp_csr& A, const uint8_t* keep_row, const uint8_t* keep_col, const int32_t* row_o2n,
                 int32_t m_new, int32_t n_new, gp_csr& out) {
  const int32_t m = A.rows;
  // column ranks: the kept bits and the kept count before each 32-column word
  const int64_t nw = (static_cast<int64_t>(A.cols) + 31) / 32;
  DVec<uint2> rank_words(static_cast<size_t>(keep_col && nw > 0 ? nw : 1));
  if (keep_col && nw > 0) {
    const PackedMask cols(keep_col, A.cols);
    DVec<int32_t> word_prefix(static_cast<size_t>(nw));
    hipLaunchKernelGGL(k_word_popc, dim3(grid_for(nw)), dim3(kBlock), 0, 0, cols.words.data(), nw, word_prefix.data());
    GP_LAUNCH_CHECK("compact word counts");
    exclusive_scan_i32(word_prefix.data(), nw);
    hipLaunchKernelGGL(k_rank_words, dim3(grid_for(nw)), dim3(kBlock), 0, 0, cols.words.data(), word_prefix.data(), nw,
                       rank_words.data());
    GP_LAUNCH_CHECK("compact rank words");
  }
  const ColRank rank{keep_col && nw > 0 ? rank_words.data() : nullptr};
  // entries per lane: 8 by default (8,192-entry tiles), or the tuned value
  const int32_t per_lane = tuning().compact_per_lane == 4 || tuning().compact_per_lane == 16
                               ? tuning().compact_per_lane
                               : 8;
  const int64_t tile_e = static_cast<int64_t>(kCompactThreads) * per_lane;
  const int64_t tiles = (static_cast<int64_t>(A.nnz) + tile_e - 1) / tile_e;
  out.rows = m_new;
  out.cols = n_new;
  out.row_ptr = pool_alloc<gp_offset>(static_cast<size_t>(m_new) + 1);
  if (tiles == 0) {
    out.nnz = 0;
    out.col_idx = pool_alloc<int32_t>(0);
    out.vals = pool_alloc<double>(0);
    GP_CHECK(hipMemsetAsync(out.row_ptr, 0, sizeof(gp_offset) * (static_cast<size_t>(m_new) + 1), 0));
    return;
  }
  // the kept count is known only after the pass: room for every entry
  out.col_idx = pool_alloc<int32_t>(static_cast<size_t>(A.nnz));
  out.vals = pool_alloc<double>(static_cast<size_t>(A.nnz));
  DVec<int32_t> tile_row(static_cast<size_t>(tiles));
  DVec<unsigned long long> ctl(static_cast<size_t>(tiles) + 3);
  unsigned long long res[3] = {};
  GP_CHECK(hipMemsetAsync(ctl.data(), 0, sizeof(unsigned long long) * (static_cast<size_t>(tiles) + 3), 0));
  hipLaunchKernelGGL(k_tile_rows, dim3(grid_for(m)), dim3(kBlock), 0, 0, A.row_ptr, m, tile_e, tile_row.data());
  {
    // the tile kernel alone
    TimedScope ts(GP_KT_COMPACT, 0.0);
    const dim3 grid(static_cast<unsigned>(tiles)), block(kCompactThreads);
    const int32_t* tr = tile_row.data();
    const int64_t nz = static_cast<int64_t>(A.nnz);
    if (per_lane == 4)
      hipLaunchKernelGGL(k_compact_tiles<4>, grid, block, 0, 0, A.row_ptr, A.col_idx, A.vals, m, nz, tiles, tr,
                         keep_row, row_o2n, rank, m_new, ctl.data(), out.row_ptr, out.col_idx, out.vals);
    else if (per_lane == 16)
      hipLaunchKernelGGL(k_compact_tiles<16>, grid, block, 0, 0, A.row_ptr, A.col_idx, A.vals, m, nz, tiles, tr,
                         keep_row, row_o2n, rank, m_new, ctl.data(), out.row_ptr, out.col_idx, out.vals);
    else
      hipLaunchKernelGGL(k_compact_tiles<8>, grid, block, 0, 0, A.row_ptr, A.col_idx, A.vals, m, nz, tiles, tr,
                         keep_row, row_o2n, rank, m_new, ctl.data(), out.row_ptr, out.col_idx, out.vals);
    GP_LAUNCH_CHECK("compact");
  }
  d2h(res, ctl.data(), 3);
  // read A and the masks, write the kept entries and the new row pointers
  // (the kept count is known after the pass)
  kclock().add_bytes(GP_KT_COMPACT, 4.0 * (m + 1) + 12.0 * A.nnz + (static_cast<double>(m) + A.cols) +
                                        12.0 * static_cast<double>(res[2]) + 4.0 * (m_new + 1));
  if (res[1] != 0) throw std::runtime_error("compaction look-back did not complete");
  if (res[2] > static_cast<unsigned long long>(INT32_MAX)) throw std::length_error("compacted nnz exceeds int32 offsets");
  out.nnz = static_cast<gp_offset>(res[2]);
}


// ---- entry-parallel segment reductions -------------------------------------------------------
namespace {
__global__ void k_seg_tile_first(const gp_offset* rp, int32_t m, int32_t* tile_seg) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int64_t a = rp[r], b = rp[r + 1];
  if (a >= b) return;
  for (int64_t t = (a + kSegTile - 1) / kSegTile; t * kSegTile < b; ++t) tile_seg[t] = static_cast<int32_t>(r);
}

// max |a| over the entries whose row of A is live; nonnegative doubles order
// as their bit patterns
struct AbsMaxOp {
  Bits keep_row;
  using Slot = unsigned long long;
  __device__ Slot value(int32_t, int32_t row, double a) const {
    return keep_row[row] ? static_cast<Slot>(__double_as_longlong(fabs(a))) : 0ull;
  }
  __device__ static Slot merge(Slot a, Slot b) { return a > b ? a : b; }
  __device__ static void combine(Slot* s, Slot v) { atomicMax(s, v); }
};
}  // namespace

void seg_tile_first(const gp_csr& M, int32_t* tile_seg) {
  if (M.rows <= 0) return;
  hipLaunchKernelGGL(k_seg_tile_first, dim3(grid_for(M.rows)), dim3(kBlock), 0, 0, M.row_ptr, M.rows, tile_seg);
  GP_LAUNCH_CHECK("segment tiles");
}

void col_max_abs_tiles(const gp_csr& AT, Bits keep_row, double* vmax) {
  seg_reduce_tiles(AT, AbsMaxOp{keep_row}, reinterpret_cast<unsigned long long*>(vmax));
}

// ---- packed masks ----------------------------------------------------------------------------
namespace {
// one byte per lane; a wavefront's ballot is two 32-bit words
__global__ void k_pack_bits(const uint8_t* mask, int64_t n, uint32_t* words) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const unsigned long long b = __ballot(i < n && mask[i] != 0);
  const int lane = static_cast<int>(__lane_id());
  if ((lane == 0 || lane == 32) && i < n) words[i >> 5] = static_cast<uint32_t>(lane == 0 ? b : b >> 32);
}
}  // namespace

void PackedMask::pack(const uint8_t* mask, int64_t n) {
  null = mask == nullptr;
  if (null) return;
  const int64_t nw = (n + 31) / 32;
  words.resize(static_cast<size_t>(nw ? nw : 1));
  if (nw > 0) {
    hipLaunchKernelGGL(k_pack_bits, dim3(grid_for(nw * 32)), dim3(kBlock), 0, 0, mask, n, words.data());
    GP_LAUNCH_CHECK("pack bits");
  }
}

// ---- rows touching marked columns ------------------------------------------------------------
namespace {
// Two ways to list the live rows that have an entry in a marked column, chosen
// on the device without a host read: k_marked_degree sums the entries of the
// marked columns; below a quarter of nnz the marked columns' entries are
// walked in AT (k_mark_rows_at flags the rows, k_list_rows lists the flags),
// otherwise k_list_rows sweeps A testing one mask bit per entry.  Each kernel
// returns at once when the other way was chosen.  The listed set is the same
// either way; its order is not defined (listed rows are processed
// independently).
__global__ void k_marked_degree(gp_csr AT, const uint8_t* col_mark, unsigned long long* total) {
  unsigned long long acc = 0;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < AT.rows;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x)
    if (col_mark[j]) acc += static_cast<unsigned long long>(AT.row_ptr[j + 1] - AT.row_ptr[j]);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (__lane_id() == 0 && acc) atomicAdd(total, acc);
}

// gate: run only when (*walk < limit) == want_walk
__device__ inline bool gate_open(const unsigned long long* walk, unsigned long long limit, bool want_walk) {
  return !walk || ((*walk < limit) == want_walk);
}

template <int G>
__global__ void __launch_bounds__(kBlock) k_mark_rows_at(gp_csr AT, const uint8_t* col_mark, const uint8_t* keep_row,
                                                          const unsigned long long* walk, unsigned long long limit,
                                                          uint8_t* row_flag) {
  if (!gate_open(walk, limit, true)) return;
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t j = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; j < AT.rows; j += groups) {
    if (!col_mark[j]) continue;
    const gp_offset p1 = AT.row_ptr[j + 1];
    for (gp_offset p = AT.row_ptr[j] + lane; p < p1; p += G) {
      const int32_t r = AT.col_idx[p];
      if (!keep_row || keep_row[r]) row_flag[r] = 1;
    }
  }
}

// row hit: some entry of a live row lies in a marked column (G lanes per row)
struct HitMarkedEntry {
  gp_csr A;
  const uint8_t* keep_row;
  Bits mark;
  template <int G>
  __device__ int eval(int64_t i, int lane) const {
    int hit = 0;
    if (!keep_row || keep_row[i]) {
      const gp_offset p1 = A.row_ptr[i + 1];
      for (gp_offset p = A.row_ptr[i] + lane; p < p1; p += G) hit |= mark[A.col_idx[p]] ? 1 : 0;
    }
    return hit;
  }
};
// row hit: the row's flag is set
struct HitFlag {
  const uint8_t* flag;
  template <int G>
  __device__ int eval(int64_t i, int lane) const {
    return lane == 0 && flag[i] ? 1 : 0;
  }
};

// Listed rows are staged in LDS and flushed with one global atomic per
// workgroup flush, so the list counter is not a hot spot.
template <int G, class Hit>
__global__ void __launch_bounds__(kBlock) k_list_rows(int64_t rows, Hit hit_of, const unsigned long long* walk,
                                                       unsigned long long limit, bool want_walk, int32_t* list,
                                                       int32_t* count) {
  if (!gate_open(walk, limit, want_walk)) return;
  constexpr int kRowsPerIter = kBlock / G;
  constexpr int kStage = 2048;
  __shared__ int32_t s_rows[kStage];
  __shared__ int32_t s_n, s_base;
  const int lane = threadIdx.x % G, grp = threadIdx.x / G;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int64_t b0 = static_cast<int64_t>(blockIdx.x) * kRowsPerIter; b0 < rows;
       b0 += static_cast<int64_t>(gridDim.x) * kRowsPerIter) {
    const int64_t i = b0 + grp;
    int hit = i < rows ? hit_of.template eval<G>(i, lane) : 0;
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) hit |= __shfl_xor(hit, o, G);
    if (lane == 0 && hit) s_rows[atomicAdd(&s_n, 1)] = static_cast<int32_t>(i);
    __syncthreads();
    const int32_t staged = s_n;  // every thread reads the same value before anyone changes it
    __syncthreads();
    const bool last = b0 + static_cast<int64_t>(gridDim.x) * kRowsPerIter >= rows;
    if (staged > kStage - kRowsPerIter || (last && staged > 0)) {  // flush (uniform across the block)
      if (threadIdx.x == 0) s_base = atomicAdd(count, staged);
      __syncthreads();
      for (int k = threadIdx.x; k < staged; k += kBlock) list[s_base + k] = s_rows[k];
      __syncthreads();
      if (threadIdx.x == 0) s_n = 0;
      __syncthreads();
    }
  }
}
static_assert(kBlock % 64 == 0, "row-list staging assumes whole wavefronts");

template <int G>
void launch_rows_touching(const gp_csr& A, const uint8_t* keep_row, Bits mark, const unsigned long long* walk,
                          unsigned long long limit, int32_t* list, int32_t* count) {
  const unsigned grid = grid_capped(static_cast<int64_t>(A.rows) * G, kBlock, 16384);
  hipLaunchKernelGGL((k_list_rows<G, HitMarkedEntry>), dim3(grid), dim3(kBlock), 0, 0, static_cast<int64_t>(A.rows),
                     HitMarkedEntry{A, keep_row, mark}, walk, limit, false, list, count);
}

__global__ void k_shift_listed(gp_csr A, const int32_t* list, const int32_t* count, const uint8_t* sel,
                               const double* val, double* AL, double* AU) {
  const int32_t total = *count;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t i = list[k];
    RowAcc acc;
    const gp_offset p0 = A.row_ptr[i];
    for (gp_offset q = p0; q < A.row_ptr[i + 1]; ++q) {
      acc.pos(q - p0);
      const int32_t j = A.col_idx[q];
      if (sel[j]) acc.add(A.vals[q] * val[j]);
    }
    const double sum = acc.value();
    AL[i] -= sum;
    AU[i] -= sum;
  }
}
}  // namespace

void list_rows_touching(const gp_csr& A, const gp_csr* AT, const uint8_t* keep_row, const uint8_t* col_mark,
                        int32_t* list, int32_t* count) {
  if (A.rows <= 0) return;
  const PackedMask mark(col_mark, A.cols);
  DVec<unsigned long long> walk;
  DVec<uint8_t> row_flag;
  const unsigned long long limit = static_cast<unsigned long long>(A.nnz) / 4;
  const bool can_walk = AT && AT->rows == A.cols && col_mark;
  if (can_walk) {
    walk.resize(1);
    row_flag.resize(static_cast<size_t>(A.rows));
    GP_CHECK(hipMemsetAsync(walk.data(), 0, sizeof(unsigned long long), 0));
    GP_CHECK(hipMemsetAsync(row_flag.data(), 0, static_cast<size_t>(A.rows), 0));
    hipLaunchKernelGGL(k_marked_degree, dim3(grid_capped(AT->rows, kBlock, 1024)), dim3(kBlock), 0, 0, *AT, col_mark,
                       walk.data());
    hipLaunchKernelGGL(k_mark_rows_at<4>, dim3(grid_capped(static_cast<int64_t>(AT->rows) * 4, kBlock, 16384)),
                       dim3(kBlock), 0, 0, *AT, col_mark, keep_row, static_cast<const unsigned long long*>(walk.data()),
                       limit, row_flag.data());
    hipLaunchKernelGGL((k_list_rows<1, HitFlag>), dim3(grid_capped(A.rows, kBlock, 2048)), dim3(kBlock), 0, 0,
                       static_cast<int64_t>(A.rows), HitFlag{row_flag.data()},
                       static_cast<const unsigned long long*>(walk.data()), limit, true, list, count);
    GP_LAUNCH_CHECK("rows touching columns (walk)");
  }
  const unsigned long long* gate = can_walk ? walk.data() : nullptr;
  switch (lanes_for_mean(static_cast<double>(A.nnz) / A.rows)) {
    case 2: launch_rows_touching<2>(A, keep_row, mark.view(), gate, limit, list, count); break;
    case 4: launch_rows_touching<4>(A, keep_row, mark.view(), gate, limit, list, count); break;
    case 8: launch_rows_touching<8>(A, keep_row, mark.view(), gate, limit, list, count); break;
    case 16: launch_rows_touching<16>(A, keep_row, mark.view(), gate, limit, list, count); break;
    default: launch_rows_touching<64>(A, keep_row, mark.view(), gate, limit, list, count); break;
  }
  GP_LAUNCH_CHECK("rows touching columns");
}

void shift_rows_by_columns(const gp_csr& A, const gp_csr* AT, const uint8_t* keep_row, const uint8_t* sel,
                           const double* val, double* AL, double* AU) {
  if (A.rows <= 0) return;
  DVec<int32_t> list(static_cast<size_t>(A.rows)), count(1);
  GP_CHECK(hipMemsetAsync(count.data(), 0, sizeof(int32_t), 0));
  list_rows_touching(A, AT, keep_row, sel, list.data(), count.data());
  hipLaunchKernelGGL(k_shift_listed, dim3(grid_capped(A.rows, kBlock, 4096)), dim3(kBlock), 0, 0, A, list.data(),
                     count.data(), sel, val, AL, AU);
  GP_LAUNCH_CHECK("shift listed rows");
}

// ---- live statistics ---------------------------------------------------------------------
// A group of G lanes per row (G = 4..64 from the mean row length): lanes stride
// the row so each load instruction reads G consecutive entries of 64/G rows.
// Counts are integers and the singleton entry is the largest live position,
// so every result is independent of the lane schedule.
namespace {
template <int G>
__global__ void __launch_bounds__(kBlock) k_live_stats(gp_csr M, const uint8_t* keep_outer,
                                                       Bits keep_inner, int32_t* degree, uint8_t* empty,
                                                       uint8_t* singleton, int32_t* support, double* coefficient,
                                                       unsigned long long* total) {
  __shared__ unsigned long long lds[kBlock / G];
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  unsigned long long acc = 0;
  for (int64_t r = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; r < M.rows; r += groups) {
    int32_t count = 0;
    gp_offset position = -1;
    const bool active = !keep_outer || keep_outer[r];
    if (active) {
      const gp_offset p1 = M.row_ptr[r + 1];
      for (gp_offset p = M.row_ptr[r] + lane; p < p1; p += G) {
        if (!keep_inner[M.col_idx[p]]) continue;
        ++count;
        position = p;
      }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      count += __shfl_xor(count, o, G);
      const gp_offset other = __shfl_xor(position, o, G);
      position = other > position ? other : position;
    }
    if (lane == 0) {
      degree[r] = count;
      empty[r] = active && count == 0;
      singleton[r] = active && count == 1;
      support[r] = count == 1 ? M.col_idx[position] : -1;
      coefficient[r] = count == 1 ? M.vals[position] : 0.0;
      acc += static_cast<unsigned long long>(count);
    }
  }
  if (total) {
    if (lane == 0) lds[threadIdx.x / G] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int k = 0; k < kBlock / G; ++k) t += lds[k];
      if (t) atomicAdd(total, t);
    }
  }
}

template <int G>
void launch_live_stats(const gp_csr& M, const uint8_t* outer, Bits inner, int32_t* degree, uint8_t* empty,
                       uint8_t* singleton, int32_t* support, double* coefficient, unsigned long long* total) {
  const int64_t threads = static_cast<int64_t>(M.rows) * G;
  hipLaunchKernelGGL(k_live_stats<G>, dim3(grid_capped(threads, kBlock, 8192)), dim3(kBlock), 0, 0, M, outer, inner,
                     degree, empty, singleton, support, coefficient, total);
}

void live_stats_side(const gp_csr& M, const uint8_t* outer, Bits inner, int32_t* degree, uint8_t* empty,
                     uint8_t* singleton, int32_t* support, double* coefficient, unsigned long long* total) {
  const double mean = M.rows ? static_cast<double>(M.nnz) / M.rows : 0.0;
  switch (pick_lanes(tuning().lanes_live_stats, mean)) {
    case 2: launch_live_stats<2>(M, outer, inner, degree, empty, singleton, support, coefficient, total); break;
    case 4: launch_live_stats<4>(M, outer, inner, degree, empty, singleton, support, coefficient, total); break;
    case 8: launch_live_stats<8>(M, outer, inner, degree, empty, singleton, support, coefficient, total); break;
    case 16: launch_live_stats<16>(M, outer, inner, degree, empty, singleton, support, coefficient, total); break;
    default: launch_live_stats<64>(M, outer, inner, degree, empty, singleton, support, coefficient, total); break;
  }
}
}  // namespace

void refresh_live_stats(gp_stats& s, const gp_device_lp& lp, const gp_plan& plan, int64_t* device_total) {
  const double bytes = 4.0 * (lp.A.rows + 1) + 4.0 * lp.A.nnz + 4.0 * (lp.AT.rows + 1) + 4.0 * lp.AT.nnz +
                       2.0 * (lp.A.rows + lp.A.cols) + 17.0 * (lp.A.rows + lp.A.cols);
  TimedScope ts(GP_KT_LIVE_STATS, bytes);
  unsigned long long* total = reinterpret_cast<unsigned long long*>(device_total);
  if (total) GP_CHECK(hipMemsetAsync(total, 0, sizeof(unsigned long long), 0));
  const PackedMask rows(plan.keep_row_mask, lp.A.rows), cols(plan.keep_col_mask, lp.A.cols);
  if (lp.A.rows > 0) {
    live_stats_side(lp.A, plan.keep_row_mask, cols.view(), s.row_nnz, s.empty_row_mask, s.singleton_row_mask,
                    s.singleton_row_col, s.singleton_row_val, total);
    GP_LAUNCH_CHECK("live stats rows");
  }
  if (lp.AT.rows > 0) {
    live_stats_side(lp.AT, plan.keep_col_mask, rows.view(), s.col_nnz, s.empty_col_mask,
                    s.singleton_column_mask, s.singleton_column_row, s.singleton_column_val, nullptr);
    GP_LAUNCH_CHECK("live stats cols");
  }
  s.live_masks = 1;
}

// ---- live statistics of a fresh plan ------------------------------------------------------
// With every row and column live (a plan just created over a materialized
// LP) the live statistics are the stored row / column lengths; no mask is
// read and no entry is visited except the one entry of a singleton.
namespace {
__global__ void k_stats_from_lengths(gp_csr M, int32_t* degree, uint8_t* empty, uint8_t* singleton, int32_t* support,
                                     double* coefficient) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= M.rows) return;
  const gp_offset p0 = M.row_ptr[r];
  const int32_t len = static_cast<int32_t>(M.row_ptr[r + 1] - p0);
  degree[r] = len;
  empty[r] = len == 0;
  singleton[r] = len == 1;
  support[r] = len == 1 ? M.col_idx[p0] : -1;
  coefficient[r] = len == 1 ? M.vals[p0] : 0.0;
}
__global__ void k_set_i64(int64_t* p, int64_t v) { *p = v; }
}  // namespace

void live_stats_all_live(gp_stats& s, const gp_device_lp& lp, int64_t* device_total) {
  TimedScope ts(GP_KT_LIVE_STATS, 4.0 * (lp.A.rows + 1) + 4.0 * (lp.AT.rows + 1) + 18.0 * (lp.A.rows + lp.A.cols));
  if (lp.A.rows > 0)
    hipLaunchKernelGGL(k_stats_from_lengths, dim3(grid_for(lp.A.rows)), dim3(kBlock), 0, 0, lp.A, s.row_nnz,
                       s.empty_row_mask, s.singleton_row_mask, s.singleton_row_col, s.singleton_row_val);
  if (lp.AT.rows > 0)
    hipLaunchKernelGGL(k_stats_from_lengths, dim3(grid_for(lp.AT.rows)), dim3(kBlock), 0, 0, lp.AT, s.col_nnz,
                       s.empty_col_mask, s.singleton_column_mask, s.singleton_column_row, s.singleton_column_val);
  if (device_total) hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, 0, device_total, static_cast<int64_t>(lp.A.nnz));
  GP_LAUNCH_CHECK("live stats from lengths");
  s.live_masks = 1;
}

// ---- incremental live statistics ----------------------------------------------------------
// Keep masks only ever go from 1 to 0 inside one plan, so the statistics of
// the current masks follow from those of an earlier snapshot by visiting the
// entries of the rows and columns deleted since then
// (update_live_stats_from_deletions, gpu_presolve_kernels.cu:520-682).  Every
// quantity is an integer count or the single live entry of a row / column,
// so the result equals a full refresh bit for bit.
namespace {
__device__ inline void block_add_u64(unsigned long long v, unsigned long long* total) {
  __shared__ unsigned long long lds[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < kBlock / 64; ++k) t += lds[k];
    if (t) atomicAdd(total, t);
  }
}

// Rows deleted since the snapshot: their formerly live entries leave the
// live total; columns still live lose one entry each.
__global__ void __launch_bounds__(kBlock) k_inc_rows(gp_csr A, const uint8_t* old_rows, const uint8_t* old_cols,
                                                      const uint8_t* keep_rows, const uint8_t* keep_cols,
                                                      int32_t* col_nnz, uint8_t* touched_cols,
                                                      unsigned long long* removed, int32_t* invalid) {
  unsigned long long count = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < A.rows;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    if (!old_rows[i]) {
      if (keep_rows[i]) atomicExch(invalid, 1);
      continue;
    }
    if (keep_rows[i]) continue;
    for (gp_offset p = A.row_ptr[i]; p < A.row_ptr[i + 1]; ++p) {
      const int32_t j = A.col_idx[p];
      if (!old_cols[j]) continue;
      ++count;
      if (keep_cols[j]) {
        atomicSub(col_nnz + j, 1);
        touched_cols[j] = 1;
      }
    }
  }
  block_add_u64(count, removed);
}

// Columns deleted since the snapshot: rows still live lose one entry each.
__global__ void __launch_bounds__(kBlock) k_inc_cols(gp_csr AT, const uint8_t* old_cols, const uint8_t* keep_rows,
                                                      const uint8_t* keep_cols, int32_t* row_nnz,
                                                      uint8_t* touched_rows, unsigned long long* removed,
                                                      int32_t* invalid) {
  unsigned long long count = 0;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < AT.rows;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    if (!old_cols[j]) {
      if (keep_cols[j]) atomicExch(invalid, 1);
      continue;
    }
    if (keep_cols[j]) continue;
    for (gp_offset p = AT.row_ptr[j]; p < AT.row_ptr[j + 1]; ++p) {
      const int32_t i = AT.col_idx[p];
      if (!keep_rows[i]) continue;
      ++count;
      atomicSub(row_nnz + i, 1);
      touched_rows[i] = 1;
    }
  }
  block_add_u64(count, removed);
}

// Flags and the singleton entry of every row (column) deleted or touched
// since the snapshot, as the full refresh writes them.
__device__ inline void finalize_line(const gp_csr& M, int64_t r, bool active, const uint8_t* keep_inner,
                                     int32_t* degree, uint8_t* empty, uint8_t* singleton, int32_t* support,
                                     double* coefficient, int32_t* invalid) {
  const int32_t d = active ? degree[r] : 0;
  if (d < 0) atomicExch(invalid, 1);
  degree[r] = d;
  empty[r] = active && d == 0;
  singleton[r] = active && d == 1;
  int32_t s = -1;
  double v = 0.0;
  if (active && d == 1) {
    gp_offset found = -1;
    for (gp_offset p = M.row_ptr[r]; p < M.row_ptr[r + 1]; ++p)
      if (keep_inner[M.col_idx[p]]) found = p;
    if (found < 0) {
      atomicExch(invalid, 1);
    } else {
      s = M.col_idx[found];
      v = M.vals[found];
    }
  }
  support[r] = s;
  coefficient[r] = v;
}

// With `sync`, line k's snapshot byte takes its current mask and its touched
// byte is cleared: the kernel is the last to read either at index k.
__global__ void k_inc_finalize(gp_csr A, gp_csr AT, uint8_t* old_rows, uint8_t* old_cols, const uint8_t* keep_rows,
                               const uint8_t* keep_cols, uint8_t* touched_rows, uint8_t* touched_cols, gp_stats s,
                               int32_t* invalid, int sync) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < A.rows) {
    const uint8_t was = old_rows[k], now = keep_rows[k], t = touched_rows[k];
    if ((was && !now) || t)
      finalize_line(A, k, now != 0, keep_cols, s.row_nnz, s.empty_row_mask, s.singleton_row_mask,
                    s.singleton_row_col, s.singleton_row_val, invalid);
    if (sync) {
      if (was != now) old_rows[k] = now;
      if (t) touched_rows[k] = 0;
    }
  }
  if (k < AT.rows) {
    const uint8_t was = old_cols[k], now = keep_cols[k], t = touched_cols[k];
    if ((was && !now) || t)
      finalize_line(AT, k, now != 0, keep_rows, s.col_nnz, s.empty_col_mask, s.singleton_column_mask,
                    s.singleton_column_row, s.singleton_column_val, invalid);
    if (sync) {
      if (was != now) old_cols[k] = now;
      if (t) touched_cols[k] = 0;
    }
  }
}

// a rejected update leaves the live total as it was
__global__ void k_inc_total(int64_t* total, const unsigned long long* removed, int32_t* invalid) {
  if (*invalid) return;
  if (static_cast<unsigned long long>(*total) < *removed) {
    *invalid = 1;
    return;
  }
  *total -= static_cast<int64_t>(*removed);
}
}  // namespace

bool update_live_stats(gp_stats& s, const gp_device_lp& lp, const gp_plan& plan, uint8_t* old_rows,
                       uint8_t* old_cols, uint8_t* touched_rows, uint8_t* touched_cols, int64_t* device_total,
                       int32_t* scratch, bool sync) {
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (!plan.keep_row_mask || !plan.keep_col_mask || !device_total) return false;
  TimedScope ts(GP_KT_LIVE_STATS_INC, 2.0 * (m + n) + 2.0 * (m + n) + 4.0 * (m + n));
  if (!sync) {
    if (m) GP_CHECK(hipMemsetAsync(touched_rows, 0, static_cast<size_t>(m), 0));
    if (n) GP_CHECK(hipMemsetAsync(touched_cols, 0, static_cast<size_t>(n), 0));
  }
  GP_CHECK(hipMemsetAsync(scratch, 0, 4 * sizeof(int32_t), 0));
  unsigned long long* removed = reinterpret_cast<unsigned long long*>(scratch);
  int32_t* invalid = scratch + 2;
  if (m)
    hipLaunchKernelGGL(k_inc_rows, dim3(grid_capped(m, kBlock, 4096)), dim3(kBlock), 0, 0, lp.A, old_rows, old_cols,
                       plan.keep_row_mask, plan.keep_col_mask, s.col_nnz, touched_cols, removed, invalid);
  if (n)
    hipLaunchKernelGGL(k_inc_cols, dim3(grid_capped(n, kBlock, 4096)), dim3(kBlock), 0, 0, lp.AT, old_cols,
                       plan.keep_row_mask, plan.keep_col_mask, s.row_nnz, touched_rows, removed, invalid);
  const int64_t most = m > n ? m : n;
  if (most)
    hipLaunchKernelGGL(k_inc_finalize, dim3(grid_for(most)), dim3(kBlock), 0, 0, lp.A, lp.AT, old_rows, old_cols,
                       plan.keep_row_mask, plan.keep_col_mask, touched_rows, touched_cols, s, invalid, sync ? 1 : 0);
  hipLaunchKernelGGL(k_inc_total, dim3(1), dim3(1), 0, 0, device_total, removed, invalid);
  GP_LAUNCH_CHECK("incremental live stats");
  return read1(invalid) == 0;
}

// ---- interleaved bounds ----------------------------------------------------------------
namespace {
__global__ void k_pack_bounds(const double* l, const double* u, int32_t n, double2* lu, double* l_copy,
                              double* u_copy) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double lj = l[j], uj = u[j];
  lu[j] = make_double2(lj, uj);
  if (l_copy) {
    l_copy[j] = lj;
    u_copy[j] = uj;
  }
}
}  // namespace

PackedBounds::PackedBounds(const double* l, const double* u, int32_t n, double* l_copy, double* u_copy)
    : lu(static_cast<size_t>(n > 0 ? n : 1)) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pack_bounds, dim3(grid_for(n)), dim3(kBlock), 0, 0, l, u, n, lu.data(), l_copy, u_copy);
  GP_LAUNCH_CHECK("pack bounds");
}

namespace {
// bit j: column j is live (keep null: every column) and its pair equals the snap (bitwise)
__global__ void k_pairs_current(const double2* lu, const double2* snap, const uint8_t* keep, int32_t n,
                                uint32_t* words) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool c = false;
  if (j < n) {
    const double2 a = lu[j], b = snap[j];
    c = __double_as_longlong(a.x) == __double_as_longlong(b.x) && __double_as_longlong(a.y) == __double_as_longlong(b.y) &&
        (!keep || keep[j]);
  }
  const unsigned long long bal = __ballot(c);
  const int lane = static_cast<int>(__lane_id());
  if ((lane == 0 || lane == 32) && j < n) words[j >> 5] = static_cast<uint32_t>(lane == 0 ? bal : bal >> 32);
}
}  // namespace

StoreView store_view(EntryBounds* eb, const gp_csr& A, const double2* lu, const uint8_t* keep_col, double zero_tol,
                     bool writer, DVec<uint32_t>& current) {
  StoreView v;
  if (!eb || A.nnz <= 0 || A.cols <= 0) return v;
  const int32_t n = A.cols;
  if (eb->holds(A, zero_tol)) {
    // current pairs, the snap, the mask, a bit out per column
    TimedScope ts(GP_KT_ENTRY_BOUNDS, (keep_col ? 33.0 : 32.0) * n + n / 8.0);
    current.resize(static_cast<size_t>((n + 31) / 32));
    hipLaunchKernelGGL(k_pairs_current, dim3(grid_for((static_cast<int64_t>(n) + 31) / 32 * 32)), dim3(kBlock), 0, 0,
                       lu, eb->snap.data(), keep_col, n, current.data());
    GP_LAUNCH_CHECK("current pairs");
    v.current = Bits{current.data()};
  } else if (writer) {
    eb->drop();
    eb->pairs.resize(static_cast<size_t>(A.nnz));
    v.current = Bits{nullptr};  // no column: gather and store
  } else {
    return v;
  }
  v.pairs = eb->pairs.data();
  v.zero_tol = zero_tol;
  v.write = writer;
  return v;
}

void store_commit(EntryBounds* eb, const gp_csr& A, const double2* lu, double zero_tol) {
  if (!eb || A.nnz <= 0 || A.cols <= 0) return;
  {
    TimedScope ts(GP_KT_ENTRY_BOUNDS, 32.0 * A.cols);
    eb->snap.resize(static_cast<size_t>(A.cols));
    d2d(eb->snap.data(), lu, static_cast<size_t>(A.cols));
  }
  eb->col_idx = A.col_idx;
  eb->nnz = static_cast<int64_t>(A.nnz);
  eb->n = A.cols;
  eb->zero_tol = zero_tol;
  eb->valid = true;
}

void store_commit(EntryBounds* eb, const gp_csr& A, PackedBounds& lu, double zero_tol) {
  if (!eb || A.nnz <= 0 || A.cols <= 0) return;
  std::swap(eb->snap, lu.lu);
  lu.lu.reset();
  eb->col_idx = A.col_idx;
  eb->nnz = static_cast<int64_t>(A.nnz);
  eb->n = A.cols;
  eb->zero_tol = zero_tol;
  eb->valid = true;
}

// ---- row activity summary ------------------------------------------------------------------
// Per live row: the finite parts of the minimum / maximum activity (blocked
// Neumaier sums in CSR order) and the counts of -inf / +inf terms; entries
// of deleted columns and with |a| <= zero_tol are skipped.
namespace {
struct ActivityFold {
  const uint8_t* keep_row;
  double zero_tol;
  double* min_fin;
  double* max_fin;
  int32_t* min_ninf;
  int32_t* max_pinf;
  BsRow* rec;  // optional
  const double* AL;
  const double* AU;
  const int32_t* row_nnz;
  struct State {
    RowAcc lo, hi;
    int32_t ni, pi;
  };
  __device__ bool active(int64_t i) const { return !keep_row || keep_row[i]; }
  __device__ bool use(double a) const { return fabs(a) > zero_tol; }
  __device__ void init(State& s) const {
    s.lo = RowAcc{};
    s.hi = RowAcc{};
    s.ni = s.pi = 0;
  }
  __device__ void add(State& s, int64_t k, double lo, double hi, bool skip) const {
    s.lo.pos(k);
    s.hi.pos(k);
    if (skip) return;
    if (isfinite(lo)) s.lo.add(lo);
    else if (lo < 0.0) ++s.ni;
    if (isfinite(hi)) s.hi.add(hi);
    else if (hi > 0.0) ++s.pi;
  }
  __device__ void finish(int64_t i, const State& s, bool) const {
    const double lo = s.lo.value(), hi = s.hi.value();
    min_fin[i] = lo;
    max_fin[i] = hi;
    min_ninf[i] = s.ni;
    max_pinf[i] = s.pi;
    if (rec) rec[i] = BsRow{lo, hi, AL[i], AU[i], s.ni, s.pi, row_nnz[i], active(i) ? 1 : 0};
  }
};
}  // namespace

void row_activity_summary(const gp_csr& A, const double2* lu, double zero_tol, const uint8_t* keep_row,
                          const uint8_t* keep_col, double* min_fin, double* max_fin, int32_t* min_ninf,
                          int32_t* max_pinf, BsRow* rec, const double* AL, const double* AU,
                          const int32_t* row_nnz, const StoreView& store) {
  if (A.rows <= 0) return;
  const PackedMask cols(keep_col, A.cols);
  // compulsory bytes of the sweep: row_ptr, col_idx, vals; one (l, u) pair per column; the row and
  // column masks (one byte each); 24 bytes out per row.  The optional stored outputs (16 bytes per
  // entry, a 64-byte record per row) are intermediates for later rules, not counted.
  const double bytes = 4.0 * (A.rows + 1) + 12.0 * A.nnz + 16.0 * A.cols + (static_cast<double>(A.rows) + A.cols) +
                       24.0 * A.rows;
  TimedScope ts(GP_KT_ACTIVITY, bytes);
  launch_row_fold(A, lu, cols.view(),
                  ActivityFold{keep_row, zero_tol, min_fin, max_fin, min_ninf, max_pinf, rec, AL, AU, row_nnz},
                  store);
  GP_LAUNCH_CHECK("activity summary");
}

void row_activity_summary(const gp_csr& A, const double* l, const double* u, double zero_tol,
                          const uint8_t* keep_row, const uint8_t* keep_col, double* min_fin, double* max_fin,
                          int32_t* min_ninf, int32_t* max_pinf) {
  if (A.rows <= 0) return;
  const PackedBounds lu(l, u, A.cols);
  row_activity_summary(A, lu.data(), zero_tol, keep_row, keep_col, min_fin, max_fin, min_ninf, max_pinf);
}

}  // namespace gp

// ---- row activity summary without allocation ---------------------------------------------
// For callers that cannot allocate device memory (code under stream capture,
// tight serving loops).  The two scratch buffers of the activity sweep -- the
// interleaved (l, u) pairs and the packed column-mask bits -- live in one
// file-scope device array of fixed size instead of pool blocks.  Every call on
// a device shares that device's array, so calls that overlap on one device
// race; each call rewrites every scratch word it reads.
namespace gp {
namespace {
constexpr int64_t kNoAllocMaxCols = int64_t{1} << 21;  // 2,097,152 columns
constexpr size_t kScratchBoundsBytes = static_cast<size_t>(kNoAllocMaxCols) * sizeof(double2);
constexpr size_t kScratchBytes = kScratchBoundsBytes + static_cast<size_t>(kNoAllocMaxCols / 32) * sizeof(uint32_t);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device
// with that device current.  The host-side name of the array is not a device
// address; only this pointer is passed to kernels.
unsigned char* scratch_on_current_device() {
  int dev = 0;
  GP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(GP_ERR_RUNTIME, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    GP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(GP_ERR_RUNTIME, "no device address for the scratch array");
    g_scratch_addr[dev] = p;
  }
  return static_cast<unsigned char*>(g_scratch_addr[dev]);
}
}  // namespace
}  // namespace gp

extern "C" int gp_row_activity_summary_noalloc(const gp_csr* A, const double* l, const double* u, double zero_tol,
                                               const uint8_t* keep_row, const uint8_t* keep_col, double* min_fin,
                                               double* max_fin, int32_t* min_ninf, int32_t* max_pinf) {
  using namespace gp;
  return guarded([&] {
    if (!A || !l || !u || !keep_row || !min_fin || !max_fin || !min_ninf || !max_pinf)
      throw std::invalid_argument("null argument");
    if (A->rows < 0 || A->cols < 0) throw std::invalid_argument("negative matrix dimension");
    if (A->cols > kNoAllocMaxCols)
      throw std::invalid_argument("gp_row_activity_summary_noalloc: more than 2097152 columns (use "
                                  "gp_row_activity_summary)");
    if (A->rows == 0) return;
    unsigned char* base = scratch_on_current_device();
    double2* lu = reinterpret_cast<double2*>(base);
    uint32_t* words = reinterpret_cast<uint32_t*>(base + kScratchBoundsBytes);
    const int32_t n = A->cols;
    if (n > 0) {
      hipLaunchKernelGGL(k_pack_bounds, dim3(grid_for(n)), dim3(kBlock), 0, 0, l, u, n, lu, nullptr, nullptr);
      if (keep_col)
        hipLaunchKernelGGL(k_pack_bits, dim3(grid_for(((static_cast<int64_t>(n) + 31) / 32) * 32)), dim3(kBlock), 0, 0,
                           keep_col, static_cast<int64_t>(n), words);
      GP_LAUNCH_CHECK("pack activity scratch");
    }
    launch_row_fold(*A, lu, Bits{keep_col ? words : nullptr},
                    ActivityFold{keep_row, zero_tol, min_fin, max_fin, min_ninf, max_pinf});
    GP_LAUNCH_CHECK("activity summary (no allocation)");
    GP_CHECK(hipDeviceSynchronize());
  });
}
