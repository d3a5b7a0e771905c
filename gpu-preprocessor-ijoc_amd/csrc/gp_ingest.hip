// gp_ingest.hip -- device assembly of an LP from matrix entries
// (gp_lp_upload_entries, gp_mps_upload).
//
// The reference builds the CSC on the host -- per column a stable sort of the
// entries by row, duplicates summed left to right from 0.0
// (cpp/mps_reader/src/io/mpsreader.cpp:625-648) -- converts it to CSR with a
// counting pass (cpp/mps_reader/src/core/sparse_matrix.cpp:111-135) and
// copies both to the device (cpp/tools/gpu_preprocessor_tool_common.hpp:108-141).
// Here the entries go to HBM as they are (16 B each) and the device does the
// rest:
//  (1) key = column << row_bits | row, order = entry number;
//  (2) stable radix sort of (key, order) over row_bits + col_bits bits --
//      equal keys keep file order, which is the reference's stable sort;
//  (3) run heads (first entry of each key) and their exclusive scan give
//      every distinct (row, column) its slot in A^T;
//  (4) one thread per run sums the run's values in file order from 0.0;
//  (5) column starts of A^T from the sorted distinct keys;
//  (6) A = transpose(A^T) (stable, so columns increase within every row).
// The result is bit-identical to the reference's CSC / CSR, explicit zeros
// and sums that cancel to zero included.
#include "gp_common.hpp"

namespace gp {

namespace {

__global__ void k_entry_keys(const int32_t* __restrict__ row, const int32_t* __restrict__ col, int64_t entries,
                             int32_t m, int32_t n, int row_bits, uint64_t* __restrict__ key,
                             int32_t* __restrict__ order, int32_t* __restrict__ bad) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < entries; k += stride) {
    const int32_t i = row[k], j = col[k];
    const bool ok = i >= 0 && i < m && j >= 0 && j < n;
    if (!ok) *bad = 1;
    key[k] = ok ? (static_cast<uint64_t>(j) << row_bits) | static_cast<uint64_t>(i) : 0;
    order[k] = static_cast<int32_t>(k);
  }
}

// slot[k] = 1 when entry k starts a run of equal keys; slot[entries] = 0
__global__ void k_run_heads(const uint64_t* __restrict__ key, int64_t entries, int32_t* __restrict__ slot) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k > entries) return;
  slot[k] = k < entries && (k == 0 || key[k] != key[k - 1]) ? 1 : 0;
}

// One thread per run head: the run's values summed in file order.
__global__ void k_sum_runs(const uint64_t* __restrict__ key, const int32_t* __restrict__ order,
                           const double* __restrict__ val, const int32_t* __restrict__ slot, int64_t entries,
                           int row_bits, int32_t* __restrict__ at_col_idx, double* __restrict__ at_val,
                           int32_t* __restrict__ at_row_of) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= entries) return;
  const uint64_t kk = key[k];
  if (k > 0 && key[k - 1] == kk) return;
  double s = 0.0;
  for (int64_t t = k; t < entries && key[t] == kk; ++t) s += val[order[t]];
  const int32_t p = slot[k];
  at_col_idx[p] = static_cast<int32_t>(kk & ((uint64_t{1} << row_bits) - 1));
  at_val[p] = s;
  at_row_of[p] = static_cast<int32_t>(kk >> row_bits);
}

// A^T row pointers: position p (0 <= p <= nnz) opens the columns
// (column of p-1, column of p]; p = nnz closes through column n.
__global__ void k_column_starts(const int32_t* __restrict__ at_row_of, int64_t nnz, int32_t n,
                                gp_offset* __restrict__ at_row_ptr) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p > nnz) return;
  const int64_t cur = p < nnz ? at_row_of[p] : n;
  const int64_t prev = p > 0 ? at_row_of[p - 1] : -1;
  for (int64_t j = prev + 1; j <= cur; ++j) at_row_ptr[j] = static_cast<gp_offset>(p);
}

template <class T>
T* upload_vector(const T* host, int64_t count) {
  T* d = pool_alloc<T>(static_cast<size_t>(count));
  h2d(d, host, static_cast<size_t>(count));
  return d;
}

void assemble(int32_t m, int32_t n, int64_t entries, const int32_t* entry_row, const int32_t* entry_col,
              const double* entry_val, gp_csr& A, gp_csr& AT) {
  if (entries > INT32_MAX) throw std::length_error("matrix entries exceed the int32 offset range");
  const int row_bits = bits_for(m), col_bits = bits_for(n);
  const size_t e = static_cast<size_t>(entries);
  DVec<int32_t> row(e), col(e), order(e), slot(e + 1);
  DVec<double> val(e);
  DVec<uint64_t> key(e);
  h2d(row.data(), entry_row, e);
  h2d(col.data(), entry_col, e);
  h2d(val.data(), entry_val, e);
  int64_t nnz = 0;
  AT.rows = n;
  AT.cols = m;
  AT.row_ptr = pool_alloc<gp_offset>(static_cast<size_t>(n) + 1);
  DVec<int32_t> row_of;
  if (entries > 0) {
    DVec<int32_t> bad(1);
    GP_CHECK(hipMemsetAsync(bad.data(), 0, sizeof(int32_t), 0));
    hipLaunchKernelGGL(k_entry_keys, dim3(grid_capped(entries)), dim3(kBlock), 0, 0, row.data(), col.data(),
                       entries, m, n, row_bits, key.data(), order.data(), bad.data());
    GP_LAUNCH_CHECK("entry keys");
    if (read1(bad.data())) throw std::invalid_argument("matrix entry outside the m x n matrix");
    sort_pairs_u64_i32(key.data(), order.data(), entries, row_bits + col_bits);
    hipLaunchKernelGGL(k_run_heads, dim3(grid_for(entries + 1)), dim3(kBlock), 0, 0, key.data(), entries,
                       slot.data());
    GP_LAUNCH_CHECK("run heads");
    nnz = exclusive_scan_i32_total(slot.data(), entries + 1);
  }
  AT.nnz = static_cast<gp_offset>(nnz);
  AT.col_idx = pool_alloc<int32_t>(static_cast<size_t>(nnz));
  AT.vals = pool_alloc<double>(static_cast<size_t>(nnz));
  row_of.resize(static_cast<size_t>(nnz));
  if (entries > 0) {
    hipLaunchKernelGGL(k_sum_runs, dim3(grid_for(entries)), dim3(kBlock), 0, 0, key.data(), order.data(),
                       val.data(), slot.data(), entries, row_bits, AT.col_idx, AT.vals, row_of.data());
    GP_LAUNCH_CHECK("sum runs");
  }
  hipLaunchKernelGGL(k_column_starts, dim3(grid_for(nnz + 1)), dim3(kBlock), 0, 0, row_of.data(), nnz, n,
                     AT.row_ptr);
  GP_LAUNCH_CHECK("column starts");
  transpose_csr(AT, A);
}

void upload_entries(int32_t m, int32_t n, int64_t entries, const int32_t* entry_row, const int32_t* entry_col,
                    const double* entry_val, const double* c, const double* AL, const double* AU, const double* l,
                    const double* u, double obj_constant, gp_device_lp* out) {
  if (!out || m < 0 || n < 0 || entries < 0) throw std::invalid_argument("bad LP arguments");
  if (entries > 0 && (!entry_row || !entry_col || !entry_val)) throw std::invalid_argument("null matrix entries");
  if ((n > 0 && (!c || !l || !u)) || (m > 0 && (!AL || !AU))) throw std::invalid_argument("null LP vector");
  std::memset(out, 0, sizeof(*out));
  try {
    assemble(m, n, entries, entry_row, entry_col, entry_val, out->A, out->AT);
    out->c = upload_vector(c, n);
    out->l = upload_vector(l, n);
    out->u = upload_vector(u, n);
    out->AL = upload_vector(AL, m);
    out->AU = upload_vector(AU, m);
    out->obj_constant = obj_constant;
    GP_CHECK(hipDeviceSynchronize());
  } catch (...) {
    gp_lp_free(out);
    throw;
  }
}

}  // namespace

}  // namespace gp

extern "C" int gp_lp_upload_entries(int32_t m, int32_t n, int64_t entries, const int32_t* entry_row,
                                    const int32_t* entry_col, const double* entry_val, const double* c,
                                    const double* AL, const double* AU, const double* l, const double* u,
                                    double obj_constant, gp_device_lp* out) {
  return gp::guarded([&] {
    gp::upload_entries(m, n, entries, entry_row, entry_col, entry_val, c, AL, AU, l, u, obj_constant, out);
  });
}

extern "C" int gp_mps_upload(const gp_mps_model* model, gp_device_lp* out) {
  return gp::guarded([&] {
    if (!model) throw std::invalid_argument("gp_mps_upload: null model");
    gp::upload_entries(model->rows, model->cols, model->entries, model->entry_row, model->entry_col,
                       model->entry_val, model->c, model->AL, model->AU, model->l, model->u, model->obj_constant,
                       out);
  });
}
