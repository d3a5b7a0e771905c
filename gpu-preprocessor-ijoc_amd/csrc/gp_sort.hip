// gp_sort.hip -- stable key/value radix sorts on the device (rocPRIM through
// hipCUB), used to group equal row / column hashes and to order delta keys.
// Kept in its own translation unit: the sort templates dominate compile time.
#include <hipcub/hipcub.hpp>

#include "gp_common.hpp"

namespace gp {

namespace {
template <class K, class V>
void sort_pairs(K* keys, V* vals, int64_t n, int end_bit) {
  if (n <= 1) return;
  if (n > INT32_MAX) throw std::length_error("radix sort input exceeds int32 items");
  DVec<K> k2(static_cast<size_t>(n));
  DVec<V> v2(static_cast<size_t>(n));
  hipcub::DoubleBuffer<K> kb(keys, k2.data());
  hipcub::DoubleBuffer<V> vb(vals, v2.data());
  size_t bytes = 0;
  GP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kb, vb, static_cast<int>(n), 0, end_bit, 0));
  DVec<unsigned char> tmp(bytes ? bytes : 1);
  GP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.data(), bytes, kb, vb, static_cast<int>(n), 0, end_bit, 0));
  if (kb.Current() != keys) d2d(keys, kb.Current(), static_cast<size_t>(n));
  if (vb.Current() != vals) d2d(vals, vb.Current(), static_cast<size_t>(n));
}
}  // namespace

void sort_keys_u64(uint64_t* keys, int64_t n, int end_bit) {
  if (n <= 1) return;
  if (n > INT32_MAX) throw std::length_error("radix sort input exceeds int32 items");
  DVec<uint64_t> k2(static_cast<size_t>(n));
  hipcub::DoubleBuffer<uint64_t> kb(keys, k2.data());
  size_t bytes = 0;
  GP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, kb, static_cast<int>(n), 0, end_bit, 0));
  DVec<unsigned char> tmp(bytes ? bytes : 1);
  GP_CHECK(hipcub::DeviceRadixSort::SortKeys(tmp.data(), bytes, kb, static_cast<int>(n), 0, end_bit, 0));
  if (kb.Current() != keys) d2d(keys, kb.Current(), static_cast<size_t>(n));
}

void sort_pairs_u64_i32(uint64_t* keys, int32_t* vals, int64_t n, int end_bit) {
  sort_pairs<uint64_t, int32_t>(keys, vals, n, end_bit);
}

void sort_pairs_u64_f64(uint64_t* keys, double* vals, int64_t n, int end_bit) {
  sort_pairs<uint64_t, double>(keys, vals, n, end_bit);
}

void sort_pairs_u32_rowval(uint32_t* keys, RowVal* vals, int64_t n, int end_bit) {
  sort_pairs<uint32_t, RowVal>(keys, vals, n, end_bit);
}

}  // namespace gp
