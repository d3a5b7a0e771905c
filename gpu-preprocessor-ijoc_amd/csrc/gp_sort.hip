// gp_sort.hip -- stable LSD radix sort of (key, value) pairs on the device:
// groups equal row / column hashes, orders the transpose's column keys and
// the postsolve access lists.
//
// One pass per 8-bit digit, three launches per pass:
//  (1) k_radix_hist: per 4,096-item tile, the count of every digit (per-wave
//      LDS histograms, summed), stored digit-major: hist[digit * tiles + tile];
//  (2) an exclusive scan of that array (on the device, no host read) gives,
//      for every (digit, tile), the output position of the tile's first item
//      with that digit;
//  (3) k_radix_scatter: the tile ranks its items stably by digit -- every
//      wave owns a contiguous quarter of the tile; wave ballots over the digit
//      bits give each lane's rank among its step's lanes with the same digit,
//      a per-wave running count per digit extends it across the wave's steps
//      without block barriers, and one scan of the per-wave counts places the
//      waves -- writes the keys in digit order into LDS, then streams them out
//      digit run by digit run, each value read once from the tile's window of
//      the input.
// Items keep their input order within a digit, so the passes compose into a
// stable sort on bits [0, end_bit), the unique result of any stable sort.
#include "gp_common.hpp"

namespace gp {

namespace {

constexpr int kRadixBits = 8;
constexpr int kDigits = 1 << kRadixBits;
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;  // 4,096 items
constexpr int kSortWaves = kSortBlock / 64;
static_assert(kDigits == kSortBlock, "one thread per digit");

struct NoValue {};

template <class K>
__device__ inline uint32_t digit_of(K key, int shift, uint32_t mask) {
  return static_cast<uint32_t>(key >> shift) & mask;
}

template <class K>
__global__ void __launch_bounds__(kSortBlock) k_radix_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                            uint32_t mask, int64_t tiles, int32_t* __restrict__ hist) {
  __shared__ int32_t s_h[kSortWaves][kDigits];
  const int tid = threadIdx.x, wave = tid >> 6;
  for (int w = 0; w < kSortWaves; ++w) s_h[w][tid] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
  const int64_t cnt = n - base < kSortTile ? n - base : kSortTile;
#pragma unroll 4
  for (int it = 0; it < kSortItems; ++it) {
    const int64_t t = static_cast<int64_t>(it) * kSortBlock + tid;
    if (t < cnt) atomicAdd(&s_h[wave][digit_of(keys[base + t], shift, mask)], 1);
  }
  __syncthreads();
  int32_t total = 0;
  for (int w = 0; w < kSortWaves; ++w) total += s_h[w][tid];
  hist[static_cast<int64_t>(tid) * tiles + blockIdx.x] = total;
}

template <class K, class V, bool kHasValues>
__global__ void __launch_bounds__(kSortBlock) k_radix_scatter(const K* __restrict__ kin, K* __restrict__ kout,
                                                               const V* __restrict__ vin, V* __restrict__ vout,
                                                               int64_t n, int shift, uint32_t mask, int64_t tiles,
                                                               const int32_t* __restrict__ offsets) {
  __shared__ K s_key[kSortTile];
  __shared__ uint16_t s_src[kSortTile];
  __shared__ int32_t s_wrun[kSortWaves][kDigits];  // per wave: running count of each digit
  __shared__ int32_t s_start[kDigits];
  __shared__ int32_t s_goff[kDigits];
  __shared__ int32_t s_wsum[kSortWaves];
  const int tid = threadIdx.x, wave = tid >> 6;
  const int lane = static_cast<int>(__lane_id());
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
  const int cnt = static_cast<int>(n - base < kSortTile ? n - base : kSortTile);
  for (int w = 0; w < kSortWaves; ++w) s_wrun[w][tid] = 0;
  s_goff[tid] = offsets[static_cast<int64_t>(tid) * tiles + blockIdx.x];
  __syncthreads();

  // Wave w owns the tile's items [w * 1024, (w + 1) * 1024), 64 at a time in
  // order, so tile order is (wave, step, lane).  A lane's rank within its wave
  // and digit: its wave's running count of the digit before this step plus
  // the lower lanes of this step with the same digit.  LDS accesses of one
  // wavefront complete in issue order, so every lane reads the running count
  // before the digit's lowest lane writes it back.
  K key[kSortItems];
  int32_t rank[kSortItems];
  const int wbase = wave * (kSortTile / kSortWaves);
  for (int it = 0; it < kSortItems; ++it) {
    const int t = wbase + it * 64 + lane;
    const bool valid = t < cnt;
    key[it] = valid ? kin[base + t] : K{};
    const uint32_t d = valid ? digit_of(key[it], shift, mask) : 0u;
    unsigned long long eq = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const unsigned long long bal = __ballot((d >> b) & 1u);
      eq &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const int in_wave = __popcll(eq & below);
    const int32_t run = valid ? s_wrun[wave][d] : 0;
    rank[it] = run + in_wave;
    __builtin_amdgcn_wave_barrier();
    if (valid && in_wave == 0) s_wrun[wave][d] = run + __popcll(eq);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // digit d's items: waves in order, so the start of (wave w, digit d) in the
  // tile is the tile's start of d plus the counts of d in waves before w
  int32_t before_w[kSortWaves];
  int32_t total = 0;
  for (int w = 0; w < kSortWaves; ++w) {
    before_w[w] = total;
    total += s_wrun[w][tid];
  }
  {
    // exclusive scan of the tile's digit totals (one digit per thread)
    int32_t x = total;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    int32_t prior = 0;
    for (int w = 0; w < wave; ++w) prior += s_wsum[w];
    const int32_t start = prior + x - total;
    s_start[tid] = start;
    // reuse s_wrun for the start of (wave, digit)
    for (int w = 0; w < kSortWaves; ++w) s_wrun[w][tid] = start + before_w[w];
    __syncthreads();
  }
  for (int it = 0; it < kSortItems; ++it) {
    const int t = wbase + it * 64 + lane;
    if (t < cnt) {
      const int pos = s_wrun[wave][digit_of(key[it], shift, mask)] + rank[it];
      s_key[pos] = key[it];
      s_src[pos] = static_cast<uint16_t>(t);
    }
  }
  __syncthreads();
  for (int p = tid; p < cnt; p += kSortBlock) {
    const K k = s_key[p];
    const uint32_t d = digit_of(k, shift, mask);
    const int64_t g = static_cast<int64_t>(s_goff[d]) + (p - s_start[d]);
    kout[g] = k;
    if (kHasValues) vout[g] = vin[base + s_src[p]];
  }
}

// With `out`, the sorted data stays where the last pass put it: out->keys /
// out->vals point there and out holds the scratch buffers (no copy back).
template <class K, class V>
struct SortOut {
  DVec<K> k2;
  DVec<V> v2;
  const K* keys = nullptr;
  const V* vals = nullptr;
};

template <class K, class V, bool kHasValues>
void radix_sort(K* keys, V* vals, int64_t n, int end_bit, SortOut<K, V>* out = nullptr) {
  if (out) {
    out->keys = keys;
    out->vals = vals;
  }
  if (n <= 1 || end_bit <= 0) return;
  if (n > INT32_MAX) throw std::length_error("radix sort input exceeds int32 items");
  const int key_bits = static_cast<int>(8 * sizeof(K));
  if (end_bit > key_bits) end_bit = key_bits;
  const int64_t tiles = (n + kSortTile - 1) / kSortTile;
  DVec<K> k2(static_cast<size_t>(n));
  DVec<V> v2(kHasValues ? static_cast<size_t>(n) : 1);
  DVec<int32_t> offsets(static_cast<size_t>(kDigits * tiles));
  K* kin = keys;
  K* kout = k2.data();
  V* vin = vals;
  V* vout = v2.data();
  for (int shift = 0; shift < end_bit; shift += kRadixBits) {
    const int bits = end_bit - shift < kRadixBits ? end_bit - shift : kRadixBits;
    const uint32_t mask = (1u << bits) - 1u;
    hipLaunchKernelGGL(k_radix_hist<K>, dim3(static_cast<unsigned>(tiles)), dim3(kSortBlock), 0, 0,
                       static_cast<const K*>(kin), n, shift, mask, tiles, offsets.data());
    GP_LAUNCH_CHECK("radix histogram");
    exclusive_scan_i32(offsets.data(), kDigits * tiles);
    hipLaunchKernelGGL((k_radix_scatter<K, V, kHasValues>), dim3(static_cast<unsigned>(tiles)), dim3(kSortBlock), 0,
                       0, static_cast<const K*>(kin), kout, static_cast<const V*>(vin), vout, n, shift, mask, tiles,
                       static_cast<const int32_t*>(offsets.data()));
    GP_LAUNCH_CHECK("radix scatter");
    std::swap(kin, kout);
    if (kHasValues) std::swap(vin, vout);
  }
  if (out) {
    out->keys = kin;
    out->vals = vin;
    out->k2 = std::move(k2);
    out->v2 = std::move(v2);
    return;
  }
  // an odd number of passes leaves the result in the scratch buffers
  if (kin != keys) d2d(keys, kin, static_cast<size_t>(n));
  if (kHasValues && vin != vals) d2d(vals, vin, static_cast<size_t>(n));
}

}  // namespace

void sort_keys_u64(uint64_t* keys, int64_t n, int end_bit) {
  radix_sort<uint64_t, NoValue, false>(keys, nullptr, n, end_bit);
}

void sort_pairs_u64_i32(uint64_t* keys, int32_t* vals, int64_t n, int end_bit) {
  radix_sort<uint64_t, int32_t, true>(keys, vals, n, end_bit);
}

void sort_pairs_u64_f64(uint64_t* keys, double* vals, int64_t n, int end_bit) {
  radix_sort<uint64_t, double, true>(keys, vals, n, end_bit);
}

void sort_pairs_u32_rowval(uint32_t* keys, RowVal* vals, int64_t n, int end_bit, DVec<uint32_t>& k2,
                           DVec<RowVal>& v2, const uint32_t** keys_out, const RowVal** vals_out) {
  SortOut<uint32_t, RowVal> out;
  radix_sort<uint32_t, RowVal, true>(keys, vals, n, end_bit, &out);
  k2 = std::move(out.k2);
  v2 = std::move(out.v2);
  *keys_out = out.keys;
  *vals_out = out.vals;
}

}  // namespace gp
