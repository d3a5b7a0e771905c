// gp_common.hpp -- shared host/device infrastructure of the MI355X presolve
// engine: error handling, a size-class device memory pool, device vectors,
// launch helpers and the numerics shared by every kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpupre.h"

namespace gp {

// ---- errors -----------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw Error(e == hipErrorOutOfMemory ? GP_ERR_ALLOCATION : GP_ERR_RUNTIME,
                std::string(what) + ": " + hipGetErrorString(e));
  }
}

#define GP_CHECK(expr) ::gp::check((expr), #expr)
#define GP_LAUNCH_CHECK(name) ::gp::check(hipGetLastError(), name)

void set_last_error(const std::string& msg);

template <class F>
int guarded(F&& body) {
  try {
    body();
    return GP_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::invalid_argument& e) {
    set_last_error(e.what());
    return GP_ERR_INVALID_ARGUMENT;
  } catch (const std::length_error& e) {
    set_last_error(e.what());
    return GP_ERR_LENGTH;
  } catch (const std::bad_alloc& e) {
    set_last_error("host allocation failed");
    return GP_ERR_ALLOCATION;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return GP_ERR_RUNTIME;
  }
}

// ---- device memory pool -------------------------------------------------------
// Size-class (power of two, >= 256 B) cache of hipMalloc blocks.  Presolve
// rounds allocate and release plans, maps and scratch many times; the pool
// turns those into free-list operations after the first round.
class Pool {
 public:
  static Pool& instance();
  void* get(size_t bytes);
  void put(void* p);
  void release_cached();  // hipFree every cached block (no live block touched)
  size_t cached_bytes() const { return cached_; }
  ~Pool();

 private:
  std::map<size_t, std::vector<void*>> free_;
  std::map<void*, size_t> size_of_;
  size_t cached_ = 0;
};

inline size_t size_class(size_t bytes) {
  size_t s = 256;
  while (s < bytes) s <<= 1;
  return s;
}

template <class T>
struct DVec {
  T* p = nullptr;
  size_t n = 0;    // logical size
  size_t cap = 0;  // elements available
  DVec() = default;
  explicit DVec(size_t count) { resize(count); }
  DVec(const DVec&) = delete;
  DVec& operator=(const DVec&) = delete;
  DVec(DVec&& o) noexcept : p(o.p), n(o.n), cap(o.cap) { o.p = nullptr; o.n = o.cap = 0; }
  DVec& operator=(DVec&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p; n = o.n; cap = o.cap;
      o.p = nullptr; o.n = o.cap = 0;
    }
    return *this;
  }
  ~DVec() { reset(); }
  void reset() {
    if (p) Pool::instance().put(p);
    p = nullptr;
    n = cap = 0;
  }
  // Discard contents and make room for `count` elements.
  void resize(size_t count) {
    if (count > cap) {
      reset();
      const size_t bytes = size_class(sizeof(T) * (count ? count : 1));
      p = static_cast<T*>(Pool::instance().get(bytes));
      cap = bytes / sizeof(T);
    }
    n = count;
  }
  // Grow keeping the first `keep` elements.
  void grow_keep(size_t count, size_t keep) {
    if (count <= cap) {
      n = count;
      return;
    }
    const size_t bytes = size_class(sizeof(T) * count);
    T* q = static_cast<T*>(Pool::instance().get(bytes));
    if (keep) GP_CHECK(hipMemcpyAsync(q, p, sizeof(T) * keep, hipMemcpyDeviceToDevice, 0));
    if (p) Pool::instance().put(p);
    p = q;
    cap = bytes / sizeof(T);
    n = count;
  }
  // Hand the block to an owner outside the pool's RAII (freed via pool_free).
  T* release() {
    T* q = p;
    p = nullptr;
    n = cap = 0;
    return q;
  }
  T* data() const { return p; }
  size_t size() const { return n; }
};

inline void pool_free(void* p) {
  if (p) Pool::instance().put(p);
}

template <class T>
T* pool_alloc(size_t count) {
  return static_cast<T*>(Pool::instance().get(size_class(sizeof(T) * (count ? count : 1))));
}

template <class T>
void d2h(T* dst, const T* src, size_t count) {
  if (count) GP_CHECK(hipMemcpy(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost));
}
template <class T>
void h2d(T* dst, const T* src, size_t count) {
  if (count) GP_CHECK(hipMemcpy(dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
}
template <class T>
void d2d(T* dst, const T* src, size_t count) {
  if (count) GP_CHECK(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToDevice, 0));
}
template <class T>
T read1(const T* src) {
  T v{};
  GP_CHECK(hipMemcpy(&v, src, sizeof(T), hipMemcpyDeviceToHost));
  return v;
}

// ---- launch helpers -----------------------------------------------------------
constexpr int kBlock = 256;
inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}
// Grid for grid-stride kernels that publish one atomic per block: enough
// blocks to fill 256 CUs several times over, few enough to keep the
// same-address atomics cheap.
inline unsigned grid_capped(int64_t n, int block = kBlock, int64_t cap = 4096) {
  const int64_t g = grid_for(n, block);
  return static_cast<unsigned>(g < cap ? g : cap);
}

// ---- kernel-family timing ---------------------------------------------------------
// When enabled, every launch family is bracketed by HIP events on the null
// stream; elapsed time and algorithmic bytes accumulate per family.
// Scopes nest: an inner scope pauses the enclosing one, so each family's
// time is exclusive of the families timed inside it.
struct KernelClock {
  bool enabled = false;
  double ms[GP_KT_COUNT] = {};
  int64_t launches[GP_KT_COUNT] = {};
  double bytes[GP_KT_COUNT] = {};
  std::vector<hipEvent_t> pool;
  struct Segment { int family; hipEvent_t a, b; };
  std::vector<Segment> segments;  // recorded, not yet read
  std::vector<int> open;          // families of the open scopes, innermost last
  hipEvent_t take();
  void start_segment(int family);
  void stop_segment();
  int begin(int family);                // returns a token (or -1)
  void end(int token, double algorithmic_bytes);
  // bytes of a family whose size is known only after its scope closed
  void add_bytes(int family, double b) {
    if (enabled) bytes[family] += b;
  }
  void flush();                          // sync + accumulate
  ~KernelClock();
};
KernelClock& kclock();

struct TimedScope {
  int token;
  double bytes;
  TimedScope(int family, double b) : token(kclock().begin(family)), bytes(b) {}
  ~TimedScope() { kclock().end(token, bytes); }
};

// ---- numerics shared by all kernels (mirrored by the CPU restatement) --------------
constexpr int kSumBlock = 256;  // stored entries per Neumaier block in row sums

struct Neu {
  double s, c;
};

__host__ __device__ inline void neu_add(Neu& a, double x) {
  const double t = a.s + x;
  if (fabs(a.s) >= fabs(x))
    a.c += (a.s - t) + x;
  else
    a.c += (x - t) + a.s;
  a.s = t;
}
__host__ __device__ inline void neu_merge(Neu& a, Neu b) {
  neu_add(a, b.s);
  a.c += b.c;
}
__host__ __device__ inline double neu_value(Neu a) { return a.s + a.c; }

// Row accumulator: pos(k) for every stored position k of the row before add().
struct RowAcc {
  Neu total{0.0, 0.0}, blk{0.0, 0.0};
  __host__ __device__ inline void pos(int64_t k) {
    if (k > 0 && k % kSumBlock == 0) {
      neu_merge(total, blk);
      blk.s = blk.c = 0.0;
    }
  }
  __host__ __device__ inline void add(double x) { neu_add(blk, x); }
  __host__ __device__ inline double value() const {
    Neu t = total;
    neu_merge(t, blk);
    return neu_value(t);
  }
};

// fmax / fmin with C library semantics: ties return the first argument.
__host__ __device__ inline double cmax(double x, double y) { return (x >= y || y != y) ? x : y; }
__host__ __device__ inline double cmin(double x, double y) { return (x <= y || y != y) ? x : y; }

// Max / min under the IEEE total order of non-NaN values (-0 < +0).
__host__ __device__ inline bool total_greater(double a, double b) {
  return a > b || (a == b && std::signbit(b) && !std::signbit(a));
}

__device__ inline void atomic_total_max(double* slot, double v) {
  unsigned long long* addr = reinterpret_cast<unsigned long long*>(slot);
  unsigned long long old = __atomic_load_n(addr, __ATOMIC_RELAXED);
  while (total_greater(v, __longlong_as_double(static_cast<long long>(old)))) {
    const unsigned long long want = static_cast<unsigned long long>(__double_as_longlong(v));
    if (__atomic_compare_exchange_n(addr, &old, want, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      break;
  }
}
__device__ inline void atomic_total_min(double* slot, double v) {
  unsigned long long* addr = reinterpret_cast<unsigned long long*>(slot);
  unsigned long long old = __atomic_load_n(addr, __ATOMIC_RELAXED);
  while (total_greater(__longlong_as_double(static_cast<long long>(old)), v)) {
    const unsigned long long want = static_cast<unsigned long long>(__double_as_longlong(v));
    if (__atomic_compare_exchange_n(addr, &old, want, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      break;
  }
}

// ---- packed keep masks ----------------------------------------------------------------
// Bit-packed copy of a 0/1 byte mask for the gather-heavy kernels: an 8M-entry
// mask becomes 1 MB and stays resident in every XCD's L2, so the random
// keep[col] lookups of a CSR sweep stop costing one HBM line each.  A null
// view means "everything live".
struct Bits {
  const uint32_t* w;
  __device__ inline bool operator[](int64_t i) const { return !w || ((w[i >> 5] >> (i & 31)) & 1u); }
  __host__ __device__ explicit operator bool() const { return w != nullptr; }
};

struct PackedMask {
  DVec<uint32_t> words;
  bool null = true;
  PackedMask() = default;
  PackedMask(const uint8_t* mask, int64_t n) { pack(mask, n); }
  void pack(const uint8_t* mask, int64_t n);  // gp_primitives.hip
  Bits view() const { return Bits{null ? nullptr : words.data()}; }
};

// ---- wave-aggregated flag / counter updates ------------------------------------------
// One atomic per wavefront instead of one per lane: the lowest active lane
// publishes for every lane that reached the call.
__device__ inline bool wave_leader() {
  const unsigned long long active = __ballot(1);
  return __lane_id() == static_cast<unsigned>(__ffsll(static_cast<unsigned long long>(active)) - 1);
}
// flag = max(flag, v); every lane reaching one call site passes the same v.
// A plain read first: once set, later waves skip the atomic entirely.
__device__ inline void wave_flag(int32_t* flag, int32_t v) {
  if (wave_leader() && __atomic_load_n(flag, __ATOMIC_RELAXED) < v) atomicMax(flag, v);
}
// counter += number of lanes reaching the call.
__device__ inline void wave_count(int32_t* counter) {
  const unsigned long long active = __ballot(1);
  if (__lane_id() == static_cast<unsigned>(__ffsll(active) - 1)) atomicAdd(counter, static_cast<int32_t>(__popcll(active)));
}

// ---- flag words -------------------------------------------------------------------
// Eight zeroed int32 device words for a rule's flags and counters.  Slots come
// from two pre-zeroed buffers of kFlagSlots slots used in turn; a buffer is
// zeroed again (one memset, stream-ordered) when its turn comes back, so a slot
// stays valid while fewer than kFlagSlots further slots are taken -- a rule
// takes a handful.  Saves a memset launch per rule call.
constexpr int kFlagSlots = 4096;
int32_t* flag_slot();
struct Flags {
  struct Slot {
    int32_t* p;
    int32_t* data() const { return p; }
  } d;
  int32_t h[8] = {};
  Flags() : d{flag_slot()} {}
  const int32_t* read() {
    d2h(h, d.p, 8);
    return h;
  }
};

// ---- primitives (gp_primitives.hip) ---------------------------------------------
// In-place inclusive / exclusive scans; `total` receives the sum when non-null.
void inclusive_scan_i32(int32_t* data, int64_t n);
void exclusive_scan_i64(int64_t* data, int64_t n, int64_t* total);
int64_t exclusive_scan_i32_total(int32_t* data, int64_t n);  // returns total
void exclusive_scan_i32(int32_t* data, int64_t n);            // no host read
// exclusive scan whose total is written to *dev_total (device), no host read
void exclusive_scan_i64_device_total(int64_t* data, int64_t n, int64_t* dev_total);
// the same for three arrays of length n at once; totals to dev_totals[0..2]
void exclusive_scan3_i64_device_total(int64_t* a0, int64_t* a1, int64_t* a2, int64_t n, int64_t* dev_totals);
// old_to_new[i] = keep[i] ? rank : -1 ; returns number kept.
int32_t build_old_to_new(const uint8_t* keep, int32_t n, int32_t* old_to_new);
// Sum in the fixed pairwise order (chunks of 1024, stride halving, recursive).
double det_sum(const double* x, int64_t n);
void fill_u8(uint8_t* p, uint8_t v, int64_t n);
// What one row fold sees of the entry-bounds store (struct EntryBounds
// below): an entry with |a| > zero_tol whose column is marked current reads
// its stored pair; every other entry gathers lu[col], and a refreshing
// (writer) fold stores the gathered pair of such an entry back.  A current
// column is also live under the fold's column mask, so its entries skip the
// mask lookup.
struct StoreView {
  double2* pairs = nullptr;  // null: no store, every pair gathered
  Bits current{nullptr};     // live columns whose pair equals the store's snap (null: none)
  double zero_tol = 0.0;
  bool write = false;
};

// ---- LDS-staged row folds --------------------------------------------------------
// A row fold visits every stored entry of a row in CSR order and folds the
// entry's bound terms (a*l, a*u oriented by the sign of a) into per-row state
// sequentially, so order-dependent sums come out as in a one-lane-per-row
// sweep.  Rows go 256 to a block; their entries stream through LDS in chunks:
// column indices, values and the entries' (l, u) pairs are loaded
// entry-parallel -- every load of a lane issued before any is used -- with
// the pair read from the entry-bounds store where the store holds it
// (coalesced) and gathered from lu[col] otherwise; then each lane folds its
// own row's staged terms.  The policy P supplies
//   bool active(row)                        -- fold this row?
//   bool use(a)                             -- entry contributes (else "skipped")
//   void init(State&)
//   void add(State&, pos_in_row, lo, hi, skipped)   -- called for every stored entry
//   void finish(row, const State&, active)  -- called for every row of the grid
// Read-once matrix stream: the non-temporal hint asks L2 and the Infinity
// Cache not to keep these lines, so the streamed CSR does not evict the
// gathered bound and mask arrays.
template <class T>
__device__ inline T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Rows per workgroup (kFoldRows by default, or the tuned value) and the LDS
// chunk of 8 entries per row.  The fold of a row is independent of both.
constexpr int kFoldRows = 256;

template <class P, int kFoldRows, int kFoldChunk = 8 * kFoldRows>
__global__ void __launch_bounds__(kFoldRows) k_row_fold(gp_csr A, const double2* lu, Bits keep_col, P pol,
                                                        StoreView store) {
  constexpr int kPer = kFoldChunk / kFoldRows;  // staged entries per lane and chunk
  __shared__ double s_lo[kFoldChunk];
  __shared__ double s_hi[kFoldChunk];
  __shared__ uint8_t s_skip[kFoldChunk];
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kFoldRows;
  const int64_t r1 = r0 + kFoldRows < A.rows ? r0 + kFoldRows : A.rows;
  const int64_t row = r0 + threadIdx.x;
  const bool mine = row < r1 && pol.active(row);
  typename P::State st;
  pol.init(st);
  if (__syncthreads_or(mine)) {
    const gp_offset e0 = A.row_ptr[r0], e1 = A.row_ptr[r1];
    const gp_offset p0 = mine ? A.row_ptr[row] : 0, p1 = mine ? A.row_ptr[row + 1] : 0;
    for (gp_offset c0 = e0; c0 < e1; c0 += kFoldChunk) {
      const gp_offset c1 = e1 - c0 > kFoldChunk ? c0 + kFoldChunk : e1;
      int32_t j[kPer];
      double a[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const gp_offset q = c0 + k * kFoldRows + static_cast<gp_offset>(threadIdx.x);
        const gp_offset qq = q < c1 ? q : c0;
        j[k] = ld_stream(A.col_idx + qq);
        a[k] = ld_stream(A.vals + qq);
      }
      bool used[kPer], stored[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const gp_offset q = c0 + k * kFoldRows + static_cast<gp_offset>(threadIdx.x);
        const bool cur = store.pairs && store.current.w && store.current[j[k]];
        used[k] = q < c1 && (cur || keep_col[j[k]]) && pol.use(a[k]);
        stored[k] = cur && fabs(a[k]) > store.zero_tol;
      }
      double2 b[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const gp_offset q = c0 + k * kFoldRows + static_cast<gp_offset>(threadIdx.x);
        const double2* src = stored[k] ? store.pairs + q : lu + j[k];
        b[k] = used[k] ? *src : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const gp_offset q = c0 + k * kFoldRows + static_cast<gp_offset>(threadIdx.x);
        if (store.write && used[k] && !stored[k] && store.pairs && fabs(a[k]) > store.zero_tol) store.pairs[q] = b[k];
        if (q < c1) {
          const double lo = used[k] ? (a[k] >= 0.0 ? a[k] * b[k].x : a[k] * b[k].y) : 0.0;
          const double hi = used[k] ? (a[k] >= 0.0 ? a[k] * b[k].y : a[k] * b[k].x) : 0.0;
          s_lo[q - c0] = lo;
          s_hi[q - c0] = hi;
          s_skip[q - c0] = used[k] ? 0 : 1;
        }
      }
      __syncthreads();
      if (mine) {
        const gp_offset a0 = p0 > c0 ? p0 : c0, a1 = p1 < c1 ? p1 : c1;
        for (gp_offset q = a0; q < a1; ++q) pol.add(st, q - p0, s_lo[q - c0], s_hi[q - c0], s_skip[q - c0] != 0);
      }
      __syncthreads();
    }
  }
  if (row < r1) pol.finish(row, st, mine);
}

// Launch-schedule knobs; every choice gives identical results.  lanes_*: lanes
// per row of the grouped sweeps (64 / lanes rows per wavefront); fold_rows:
// rows per workgroup of the row folds; compact_per_lane: entries per lane of
// the compaction tiles (4, 8 or 16 with 1,024 lanes per tile).  0 = derived
// from the matrix (compaction: 8).
struct Tuning {
  int32_t lanes_candidates = 0, lanes_col_max = 0, compact_per_lane = 0, lanes_live_stats = 0, fold_rows = 0;
};
Tuning& tuning();  // process-wide (gp_primitives.hip)

template <class P, int R>
void launch_row_fold_rows(const gp_csr& A, const double2* lu, Bits keep_col, const P& pol, const StoreView& store) {
  const int64_t blocks = (static_cast<int64_t>(A.rows) + R - 1) / R;
  hipLaunchKernelGGL((k_row_fold<P, R>), dim3(static_cast<unsigned>(blocks)), dim3(R), 0, 0, A, lu, keep_col, pol,
                     store);
}

template <class P>
void launch_row_fold(const gp_csr& A, const double2* lu, Bits keep_col, const P& pol,
                     const StoreView& store = StoreView{}) {
  if (A.rows <= 0) return;
  switch (tuning().fold_rows) {
    case 128: launch_row_fold_rows<P, 128>(A, lu, keep_col, pol, store); break;
    case 512: launch_row_fold_rows<P, 512>(A, lu, keep_col, pol, store); break;
    default: launch_row_fold_rows<P, kFoldRows>(A, lu, keep_col, pol, store); break;
  }
}

// (l[j], u[j]) interleaved, so a sweep gathers both bounds with one load.
struct PackedBounds {
  DVec<double2> lu;
  // also copies l and u into l_copy / u_copy when given (gp_primitives.hip)
  PackedBounds(const double* l, const double* u, int32_t n, double* l_copy = nullptr, double* u_copy = nullptr);
  const double2* data() const { return lu.data(); }
};

// ---- the entry-bounds store ---------------------------------------------------------
// The (l, u) pair of each stored entry's column of A, kept at the entry's
// position across the rules of a plan, so the row folds read the pairs
// coalesced instead of gathering them from the 16-byte-per-column array (a
// gather pulls a whole cache line per entry).  snap[j] is the pair that
// every stored entry of column j holds.  The store covers the entries with
// |a| > zero_tol and a live column in every row that was live at the last
// full-coverage fold (a fold whose active rows are all live rows).  Such a
// fold refreshes the entries of the columns whose current pair differs from
// snap, then takes the current pairs as snap.  Rows and columns only die
// within a plan, so no fold reads an entry the store does not cover.  A new
// matrix (materialize, rewritten A) starts the store over.
struct EntryBounds {
  DVec<double2> pairs;               // one per entry of the matrix it belongs to
  DVec<double2> snap;                // one per column
  const int32_t* col_idx = nullptr;  // identity of that matrix
  int64_t nnz = -1;
  int32_t n = -1;
  double zero_tol = 0.0;
  bool valid = false;
  bool holds(const gp_csr& A, double zt) const {
    return valid && col_idx == A.col_idx && nnz == static_cast<int64_t>(A.nnz) && n == A.cols && zero_tol == zt;
  }
  void drop() {
    valid = false;
    col_idx = nullptr;
  }
};
// The view of `eb` (may be null) for one fold over A with the current pairs
// lu and the fold's column mask keep_col (null: every column live);
// `current` receives the current-column bits.  A refreshing (writer) fold
// starts the store over when it holds nothing for A; a reader then gathers.
StoreView store_view(EntryBounds* eb, const gp_csr& A, const double2* lu, const uint8_t* keep_col, double zero_tol,
                     bool writer, DVec<uint32_t>& current);  // gp_primitives.hip
// After a writer fold: the current pairs become the snap.
void store_commit(EntryBounds* eb, const gp_csr& A, const double2* lu, double zero_tol);
// The same, taking over lu's buffer as the snap instead of copying it (lu is left empty).
void store_commit(EntryBounds* eb, const gp_csr& A, PackedBounds& lu, double zero_tol);

// Stable ascending sort of (key, value) pairs on bits [0, end_bit) of the key,
// in place (gp_sort.hip).
void sort_keys_u64(uint64_t* keys, int64_t n, int end_bit = 64);
void sort_pairs_u64_i32(uint64_t* keys, int32_t* vals, int64_t n, int end_bit = 64);
void sort_pairs_u64_f64(uint64_t* keys, double* vals, int64_t n, int end_bit = 64);
// (row, value) payload of the transpose sort
struct RowVal {
  int32_t row;
  int32_t pad;
  double val;
};
// The sorted pairs are left where the last pass put them: in keys / vals or
// in k2 / v2 (which receive the scratch); *keys_out / *vals_out point there.
void sort_pairs_u32_rowval(uint32_t* keys, RowVal* vals, int64_t n, int end_bit, DVec<uint32_t>& k2,
                           DVec<RowVal>& v2, const uint32_t** keys_out, const RowVal** vals_out);
void gather_compact(double* dst, const double* src, const int32_t* old_to_new, int32_t n_old);
void gather_compact_i32(int32_t* dst, const int32_t* src, const int32_t* old_to_new, int32_t n_old);
int32_t count_u8(const uint8_t* mask, int32_t n);

// ---- entry-parallel segment reductions ------------------------------------------
// An order-free reduction over every row ("segment") of a CSR matrix M --
// typically AT, whose rows are the columns of A -- that reads M entry-parallel
// instead of giving each segment its own lanes, so short and long segments
// cost the same per entry.  Tiles of kSegTile consecutive entries:
//  * each segment met by the tile marks its first tile position in a
//    kSegTile-bit head mask (one LDS OR per segment) and records whether it
//    lies wholly in the tile;
//  * an entry's slot is the last head at or before it: the highest set bit
//    of its mask word up to it, else the carry of the words before (one
//    wavefront scans the 64 words);
//  * each wavefront reduces runs of equal slots with shuffles (its 64 lanes
//    hold 64 consecutive entries), and the first lane of each run combines
//    the run into the slot with an LDS atomic;
//  * a segment wholly in the tile is stored, one that crosses a tile edge is
//    combined into the output with a global atomic.
// The output starts zero-filled, so 0 must be the identity of the reduction
// and the result of a segment with no entries.  Op supplies
//   using Slot (a 32- or 64-bit unsigned integer)
//   Slot value(int32_t segment, int32_t inner, double a)
//   static Slot merge(Slot a, Slot b)      -- the reduction of two values
//   static void combine(Slot* s, Slot v)   -- atomic merge, LDS or global memory
constexpr int kSegTile = 2048;
void seg_tile_first(const gp_csr& M, int32_t* tile_seg);  // tile -> segment of its first entry

template <class Op>
__global__ void __launch_bounds__(256) k_seg_reduce_tiles(gp_csr M, const int32_t* tile_seg, Op op,
                                                          typename Op::Slot* out) {
  using Slot = typename Op::Slot;
  constexpr int kWords = kSegTile / 32;
  static_assert(kWords == 64, "one wavefront scans the head words");
  __shared__ int32_t seg_at[kSegTile];  // segment whose first entry in the tile is here
  __shared__ uint8_t whole[kSegTile];   // that segment lies wholly in the tile
  __shared__ Slot acc[kSegTile];
  __shared__ uint32_t head[kWords];
  __shared__ int32_t carry[kWords];     // last head before each word, -1 for none
  constexpr int kPerLane = kSegTile / 256;
  const int lane = static_cast<int>(__lane_id());
  const gp_offset* rp = M.row_ptr;
  const int64_t nnz = M.nnz;
  for (int q = threadIdx.x; q < kSegTile; q += 256) acc[q] = Slot(0);
  if (threadIdx.x < kWords) head[threadIdx.x] = 0u;
  __syncthreads();
  for (int64_t tile = blockIdx.x; tile * kSegTile < nnz; tile += gridDim.x) {
    const int64_t t0 = tile * kSegTile;
    const int64_t t1 = t0 + kSegTile < nnz ? t0 + kSegTile : nnz;
    int32_t inner[kPerLane];
    double val[kPerLane];
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
      const int64_t q = t0 + static_cast<int64_t>(k) * 256 + threadIdx.x;
      const int64_t at = q < t1 ? q : t0;
      inner[k] = ld_stream(M.col_idx + at);
      val[k] = ld_stream(M.vals + at);
    }
    for (int32_t r = tile_seg[tile] + static_cast<int32_t>(threadIdx.x); r < M.rows; r += 256) {
      const int64_t ra = rp[r];
      if (ra >= t1) break;
      const int64_t rb = rp[r + 1];
      const int64_t a = ra > t0 ? ra : t0, b = rb < t1 ? rb : t1;
      if (a >= b) continue;
      const int32_t slot = static_cast<int32_t>(a - t0);
      seg_at[slot] = r;
      whole[slot] = ra >= t0 && rb <= t1 ? 1 : 0;
      atomicOr(&head[slot >> 5], 1u << (slot & 31));
    }
    __syncthreads();
    if (threadIdx.x < kWords) {
      const uint32_t w = head[threadIdx.x];
      int32_t last = w ? static_cast<int32_t>(threadIdx.x) * 32 + 31 - __clz(static_cast<int>(w)) : -1;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int32_t up = __shfl_up(last, o, 64);
        if (lane >= o && up > last) last = up;
      }
      const int32_t before = __shfl_up(last, 1, 64);
      carry[threadIdx.x] = lane == 0 ? -1 : before;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
      const int32_t off = k * 256 + static_cast<int32_t>(threadIdx.x);
      const bool in = t0 + off < t1;
      const uint32_t upto = head[off >> 5] & (0xffffffffu >> (31 - (off & 31)));
      const int32_t slot = !in ? -1 - lane : (upto ? (off & ~31) + 31 - __clz(static_cast<int>(upto)) : carry[off >> 5]);
      Slot v = op.value(seg_at[in ? slot : 0], inner[k], val[k]);  // slot 0 always holds a head
      if (!in) v = Slot(0);
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const Slot y = __shfl_down(v, o, 64);
        const int32_t ys = __shfl_down(slot, o, 64);
        if (lane + o < 64 && ys == slot) v = Op::merge(v, y);
      }
      const int32_t prev = __shfl_up(slot, 1, 64);
      if (in && (lane == 0 || prev != slot) && v) Op::combine(&acc[slot], v);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kSegTile; q += 256) {
      if (!((head[q >> 5] >> (q & 31)) & 1u)) continue;
      const int32_t r = seg_at[q];
      const Slot v = acc[q];
      if (whole[q]) out[r] = v;
      else if (v) Op::combine(out + r, v);
      acc[q] = Slot(0);
    }
    __syncthreads();
    if (threadIdx.x < kWords) head[threadIdx.x] = 0u;
    __syncthreads();
  }
}

// out[r] (M.rows slots) = the reduction of row r of M; zero-filled first.
template <class Op>
void seg_reduce_tiles(const gp_csr& M, const Op& op, typename Op::Slot* out) {
  if (M.rows <= 0) return;
  GP_CHECK(hipMemsetAsync(out, 0, sizeof(typename Op::Slot) * static_cast<size_t>(M.rows), 0));
  const int64_t tiles = (static_cast<int64_t>(M.nnz) + kSegTile - 1) / kSegTile;
  if (tiles == 0) return;
  DVec<int32_t> tile_seg(static_cast<size_t>(tiles));
  seg_tile_first(M, tile_seg.data());
  const unsigned grid = static_cast<unsigned>(tiles < 16384 ? tiles : 16384);
  hipLaunchKernelGGL(k_seg_reduce_tiles<Op>, dim3(grid), dim3(256), 0, 0, M,
                     static_cast<const int32_t*>(tile_seg.data()), op, out);
  GP_LAUNCH_CHECK("segment reduce tiles");
}

// vmax[j] = max |a| over the entries of row j of AT whose row of A is live
// (0 when there is none).
void col_max_abs_tiles(const gp_csr& AT, Bits keep_row, double* vmax);
// Matrices.
void transpose_csr(const gp_csr& A, gp_csr& AT);  // AT arrays from the pool
int bits_for(int64_t v);  // bits needed for the values 0..v-1 (at least 1)
// the histogram / atomic-scatter / segment-sort transpose (for matrices with more
// entries than one radix sort takes, 2^31 - 1)
void transpose_csr_scatter(const gp_csr& A, gp_csr& AT);
// Stable compaction by keep masks (null = keep all); row_old_to_new is the
// exclusive prefix of keep_row (null with keep_row null).  Column indices are
// renumbered to their rank among the kept columns.
void compact_csr(const gp_csr& A, const uint8_t* keep_row, const uint8_t* keep_col, const int32_t* row_old_to_new,
                 int32_t m_new, int32_t n_new, gp_csr& out);
void csr_pool_free(gp_csr& A);

// Statistics.
void refresh_live_stats(gp_stats& s, const gp_device_lp& lp, const gp_plan& plan,
                        int64_t* device_total);
// Live statistics when every row and column is live (all-ones masks).
void live_stats_all_live(gp_stats& s, const gp_device_lp& lp, int64_t* device_total);
// Bring live statistics from the masks (old_rows, old_cols) they were last
// computed for to the plan's current masks, visiting only deleted lines.
// `scratch` holds 4 int32.  Returns false (stats then need a full refresh)
// when a mask went from 0 to 1 or a count went negative.  With `sync`, the
// touched bytes arrive all zero and are left all zero, and old_rows / old_cols
// are the caller's snapshot, left equal to the plan's masks (no copy, no fill).
bool update_live_stats(gp_stats& s, const gp_device_lp& lp, const gp_plan& plan, uint8_t* old_rows,
                       uint8_t* old_cols, uint8_t* touched_rows, uint8_t* touched_cols, int64_t* device_total,
                       int32_t* scratch, bool sync = false);
void row_activity_summary(const gp_csr& A, const double* l, const double* u, double zero_tol,
                          const uint8_t* keep_row, const uint8_t* keep_col, double* min_fin,
                          double* max_fin, int32_t* min_ninf, int32_t* max_pinf);
// Everything the support-row search of bound strengthening reads about a row,
// in one 64-byte record (one cache line per gather instead of eight).
struct alignas(64) BsRow {
  double min_fin, max_fin, al, au;
  int32_t min_ninf, max_pinf, row_nnz, live;
};
// rec (optional): the summary of each row with its AL, AU, row_nnz and
// liveness, written by the same sweep.  entry_lu (optional, nnz pairs): the
// (l, u) pair gathered for each used entry (live column, |a| > zero_tol) of an
// active row, stored at the entry's position; other entries are left as they are.
void row_activity_summary(const gp_csr& A, const double2* lu, double zero_tol, const uint8_t* keep_row,
                          const uint8_t* keep_col, double* min_fin, double* max_fin, int32_t* min_ninf,
                          int32_t* max_pinf, BsRow* rec = nullptr, const double* AL = nullptr,
                          const double* AU = nullptr, const int32_t* row_nnz = nullptr,
                          const StoreView& store = StoreView{});

// Kept rows of A with an entry in a column whose col_mark byte is set, appended
// to list (any order); *count (device, caller-zeroed) receives their number.
// With AT (the transpose of A, or null), when the marked columns hold less
// than a quarter of the entries, their entries are walked in AT; otherwise G
// lanes per row of A test one bit per entry of a packed copy of col_mark.  The
// choice is made on the device.
void list_rows_touching(const gp_csr& A, const gp_csr* AT, const uint8_t* keep_row, const uint8_t* col_mark,
                        int32_t* list, int32_t* count);
// AL/AU -= the blocked CSR-order sum of a * val[col] over the entries of each
// kept row whose column has sel[col] set.  Rows without such an entry are left
// as they are (x - (+0) == x), so only the rows listed by list_rows_touching
// are folded.
void shift_rows_by_columns(const gp_csr& A, const gp_csr* AT, const uint8_t* keep_row, const uint8_t* sel,
                           const double* val, double* AL, double* AU);

// ---- batched presolve (gp_batch.hip) --------------------------------------------------------
// A batch is presolved as the block-diagonal direct sum of its instances.  The
// rules need no change for that (no reduction couples two blocks); only the
// objective constant is kept per instance as well.  While a batch runs, the
// thread's BatchContext points at the physical -> original maps of the
// working LP (updated at every materialization) and the instances' original
// row / column offsets.
struct BatchContext {
  int32_t count = 0;
  const DVec<int32_t>* row_map = nullptr;  // physical -> original row of the direct sum
  const DVec<int32_t>* col_map = nullptr;
  DVec<int32_t> row_off, col_off;          // [count+1] original offsets (device)
  DVec<int32_t> marks;                     // [count] infeasible / unbounded marks (device)
  std::vector<double> obj;                 // per-instance objective deltas
};
BatchContext*& batch_context();
// Instances a rule finds infeasible (bit 1) or unbounded (bit 2) in a batched
// presolve: the kernels that raise those flags also mark the instance that
// holds the row or column (current line -> original line -> instance), so the
// batch can retire exactly those instances.  Inactive (marks null) outside a
// batch.
struct BatchMarks {
  const int32_t* row_map = nullptr;  // current -> original row of the direct sum
  const int32_t* col_map = nullptr;
  const int32_t* row_off = nullptr;  // [count+1] original offsets
  const int32_t* col_off = nullptr;
  int32_t count = 0;
  int32_t* marks = nullptr;  // [count]
  __device__ void row(int64_t i, int32_t what) const {
    if (marks) put(row_off, row_map[i], what);
  }
  __device__ void col(int64_t j, int32_t what) const {
    if (marks) put(col_off, col_map[j], what);
  }
  // the last instance whose first line is at or before `line` (empty
  // instances share their offset with the next one)
  __device__ void put(const int32_t* off, int32_t line, int32_t what) const {
    int32_t lo = 0, hi = count;
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) >> 1;
      if (off[mid] <= line) lo = mid;
      else hi = mid;
    }
    atomicOr(marks + lo, what);
  }
};
BatchMarks batch_marks();  // gp_batch.hip: the current batch's marks, inactive outside a batch
// Objective delta of a rule: the fixed-order det_sum of contrib[0..n).  In a
// batch the contributions are also summed per instance, in item order: item i
// belongs to physical line item_line[i] (or i when item_line is null), a row
// when item_rows is set, else a column; items must be in increasing line order.
double obj_delta(const double* contrib, int64_t n, const int32_t* item_line = nullptr, bool item_rows = false);

// ---- x87 extended-precision accumulation ------------------------------------------------
// The reference sums the objective of the columns bound strengthening fixes on
// the host in long double, sequentially in column order
// (fixed_variable_tape.cu:305-325): each product c * v and each partial sum is
// rounded to a 64-bit significand (x86 extended precision), and the total is
// rounded to binary64 at the end.  X80 reproduces that arithmetic exactly in
// integer code: value = (-1)^neg * m * 2^e with m normalized (bit 63 set),
// round-to-nearest-even at every step.  Products of binary64 values and their
// sums stay far inside the extended exponent range, so only the final
// conversion handles overflow and subnormal results.
struct X80 {
  unsigned long long m;
  int32_t e;
  uint32_t neg;
  uint32_t cls;  // 0 zero, 1 finite, 2 infinity, 3 NaN
};

__device__ inline X80 x80_zero() { return X80{0ull, 0, 0u, 0u}; }

__device__ inline X80 x80_from_double(double d) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(d));
  X80 r{0ull, 0, static_cast<uint32_t>(b >> 63), 0u};
  const int ef = static_cast<int>((b >> 52) & 0x7ff);
  const unsigned long long f = b & ((1ull << 52) - 1);
  if (ef == 0x7ff) {
    r.cls = f ? 3u : 2u;
    return r;
  }
  unsigned long long mant = ef ? (f | (1ull << 52)) : f;
  if (mant == 0) return r;  // signed zero
  const int ex = ef ? ef - 1075 : -1074;
  const int sh = __clzll(static_cast<long long>(mant));
  r.m = mant << sh;
  r.e = ex - sh;
  r.cls = 1u;
  return r;
}

// round the 128-bit significand (hi:lo, bit 127 set) to 64 bits; value = hi:lo * 2^(e-64)
__device__ inline X80 x80_round128(unsigned long long hi, unsigned long long lo, int32_t e, uint32_t neg) {
  const bool round = (lo >> 63) != 0;
  const bool sticky = (lo & ~(1ull << 63)) != 0;
  if (round && (sticky || (hi & 1ull))) {
    ++hi;
    if (hi == 0) {
      hi = 1ull << 63;
      ++e;
    }
  }
  return X80{hi, e, neg, 1u};
}

__device__ inline X80 x80_mul(const X80& a, const X80& b) {
  const uint32_t neg = a.neg ^ b.neg;
  if (a.cls == 3u || b.cls == 3u || (a.cls == 2u && b.cls == 0u) || (a.cls == 0u && b.cls == 2u))
    return X80{0ull, 0, 1u, 3u};
  if (a.cls == 2u || b.cls == 2u) return X80{0ull, 0, neg, 2u};
  if (a.cls == 0u || b.cls == 0u) return X80{0ull, 0, neg, 0u};
  unsigned long long hi = __umul64hi(a.m, b.m);
  unsigned long long lo = a.m * b.m;
  int32_t e = a.e + b.e + 64;
  if (!(hi >> 63)) {  // product in [2^126, 2^127): one normalizing shift
    hi = (hi << 1) | (lo >> 63);
    lo <<= 1;
    e -= 1;
  }
  return x80_round128(hi, lo, e, neg);
}

__device__ inline X80 x80_add(const X80& a, const X80& b) {
  if (a.cls == 3u || b.cls == 3u) return X80{0ull, 0, 1u, 3u};
  if (a.cls == 2u || b.cls == 2u) {
    if (a.cls == 2u && b.cls == 2u && a.neg != b.neg) return X80{0ull, 0, 1u, 3u};
    return a.cls == 2u ? a : b;
  }
  if (a.cls == 0u && b.cls == 0u) return X80{0ull, 0, a.neg & b.neg, 0u};  // -0 only for -0 + -0
  if (a.cls == 0u) return b;
  if (b.cls == 0u) return a;
  // big >= small in magnitude
  const bool a_big = a.e > b.e || (a.e == b.e && a.m >= b.m);
  const X80& big = a_big ? a : b;
  const X80& sml = a_big ? b : a;
  const uint32_t d = static_cast<uint32_t>(big.e - sml.e);
  // 128-bit significands: big = (big.m : 0), small shifted right by d with sticky
  unsigned long long bh = big.m, bl = 0ull, sh, sl;
  if (d == 0) {
    sh = sml.m;
    sl = 0ull;
  } else if (d < 64) {
    sh = sml.m >> d;
    sl = sml.m << (64 - d);
  } else if (d < 128) {
    sh = 0ull;
    sl = d == 64 ? sml.m : (sml.m >> (d - 64));
    const unsigned long long lost = d == 64 ? 0ull : (sml.m << (128 - d));
    if (lost) sl |= 1ull;
  } else {
    sh = 0ull;
    sl = 1ull;  // sticky only
  }
  int32_t e = big.e;
  if (a.neg == b.neg) {
    unsigned long long rl = bl + sl;
    unsigned long long carry = rl < bl ? 1ull : 0ull;
    unsigned long long rh = bh + sh + carry;
    const bool out = rh < bh || (rh == bh && (sh | carry) != 0ull);
    if (out) {  // carry beyond bit 127: shift right one, keep the sticky bit
      const unsigned long long st = rl & 1ull;
      rl = (rl >> 1) | (rh << 63) | st;
      rh = (rh >> 1) | (1ull << 63);
      e += 1;
    }
    return x80_round128(rh, rl, e, big.neg);
  }
  // magnitude difference (big - small >= 0)
  unsigned long long rl = bl - sl;
  const unsigned long long borrow = bl < sl ? 1ull : 0ull;
  unsigned long long rh = bh - sh - borrow;
  if (rh == 0ull && rl == 0ull) return X80{0ull, 0, 0u, 0u};  // exact cancellation: +0
  int z;
  if (rh) {
    z = __clzll(static_cast<long long>(rh));
  } else {
    z = 64 + __clzll(static_cast<long long>(rl));
  }
  if (z >= 64) {
    rh = rl << (z - 64);
    rl = 0ull;
  } else if (z > 0) {
    rh = (rh << z) | (rl >> (64 - z));
    rl <<= z;
  }
  e -= z;
  return x80_round128(rh, rl, e, big.neg);
}

// round to binary64, nearest-even (subnormal results and overflow included)
__device__ inline double x80_to_double(const X80& x) {
  const unsigned long long sign = static_cast<unsigned long long>(x.neg) << 63;
  if (x.cls == 0u) return __longlong_as_double(static_cast<long long>(sign));
  if (x.cls == 2u) return __longlong_as_double(static_cast<long long>(sign | 0x7ff0000000000000ull));
  if (x.cls == 3u) return __longlong_as_double(static_cast<long long>(sign | 0x7ff8000000000000ull));
  const int32_t ue = x.e + 63;  // value in [2^ue, 2^(ue+1))
  if (ue > 1023) return __longlong_as_double(static_cast<long long>(sign | 0x7ff0000000000000ull));
  int32_t shift = ue >= -1022 ? 11 : 11 + (-1022 - ue);
  unsigned long long q;
  bool rb, st;
  if (shift >= 65) {
    q = 0ull;
    rb = false;
    st = true;
  } else if (shift == 64) {
    q = 0ull;
    rb = (x.m >> 63) != 0ull;
    st = (x.m & ~(1ull << 63)) != 0ull;
  } else {
    q = x.m >> shift;
    rb = ((x.m >> (shift - 1)) & 1ull) != 0ull;
    st = (x.m & ((1ull << (shift - 1)) - 1ull)) != 0ull;
  }
  if (rb && (st || (q & 1ull))) ++q;
  unsigned long long bits;
  if (ue >= -1022) {
    // q in [2^52, 2^53]; a carry to 2^53 moves to the next binade
    int32_t be = ue + 1023;
    if (q >> 53) {
      q >>= 1;
      ++be;
    }
    if (be >= 0x7ff) return __longlong_as_double(static_cast<long long>(sign | 0x7ff0000000000000ull));
    bits = (static_cast<unsigned long long>(be) << 52) | (q & ((1ull << 52) - 1ull));
  } else {
    bits = q;  // subnormal; a carry to 2^52 is the smallest normal, encoded alike
  }
  return __longlong_as_double(static_cast<long long>(sign | bits));
}

// Objective delta of fixed columns the reference's way: sum over the columns
// listed in `cols` (increasing), in order, of c[j] * v[j] in x87 extended
// precision, rounded to binary64 once; in a batch also per instance.
double obj_delta_x87(const double* c, const double* v, const int32_t* cols, int64_t count);

// ---- learned pass scheduler (gp_learned.hip) ---------------------------------------------
int rule_group(int32_t rule);
void schedule_features(const gp_round_trace* hist, int32_t done, int32_t phase, int64_t nnz_now, int64_t nnz0,
                       int32_t m0, int32_t n0, int32_t bound_only, int32_t bound_limit, float* features);
// The network on the device plus the per-group logit thresholds; decide()
// returns the mask of rule groups (bit GP_GROUP_*) to run in the next round.
struct LearnedGate {
  DVec<float> params, io;
  float threshold[GP_MLP_OUT] = {};
  bool ready = false;
  void init(const float* blob);
  uint32_t decide(const float* features, float* logits = nullptr);
};

// Lanes per row for a group-per-row sweep of rows with this mean length.
inline int lanes_for_mean(double mean) {
  if (mean <= 3.0) return 2;
  if (mean <= 10.0) return 4;
  if (mean <= 24.0) return 8;
  if (mean <= 64.0) return 16;
  return 64;
}
// Launch K<G>::kernel with G lanes per row (G in {2, 4, 8, 16, 64}) over `rows`
// rows in a grid-stride loop.
template <template <int> class K, class... Args>
void launch_grouped(int lanes, int64_t rows, Args... args) {
  const unsigned grid = grid_capped(rows * lanes, kBlock, 16384);
  switch (lanes) {
    case 2: hipLaunchKernelGGL(K<2>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
    case 4: hipLaunchKernelGGL(K<4>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
    case 8: hipLaunchKernelGGL(K<8>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
    case 16: hipLaunchKernelGGL(K<16>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
    default: hipLaunchKernelGGL(K<64>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
  }
}

// the tuned lane count when one is set, else the mean-length rule
inline int pick_lanes(int32_t knob, double mean) { return knob > 0 ? knob : lanes_for_mean(mean); }

}  // namespace gp
