// gp_fold.hip -- LP folding by colour refinement (gp_fold, and the folding
// stage of gp_preprocess).
//
// The method of cpp/src/folding/ (folding.cu:69-136, folding_refinement.cu,
// folding_kernels.cu, folding_reduce.cu, folding_map.cu):
//  * classes start from the row sides (AL, then AU) and the column data
//    (c, then l, then u): within a class, lines whose value, rounded to
//    float, exceeds its predecessor's by more than the tolerance start a new
//    class (ids are the 1-based ranks of the (class, value) order);
//  * every class draws a signature, a standard normal from a SplitMix64
//    hash of (round, side, class) through Box-Muller; a round propagates
//    the row signatures to the columns (A^T times them) and refines the
//    column classes by the result, then the column signatures to the rows
//    (A times them), until neither side gains a class;
//  * the refinement gives up when both sides keep more than 80% of their
//    lines; no fold either when every line is its own class;
//  * the folded LP has one row per row class and one column per column
//    class: entry (R, C) sums a_ij / |R| over i in R, j in C, entries with
//    |sum| <= tolerance dropped; c sums over the class; l, u, AL, AU are
//    class averages; sums run over members in increasing index order.
//
// MI355X design: every class update is one stable radix sort of packed
// (class, orderable float) keys with the line index as payload, a start
// flag pass, a device scan and one scatter that also writes the new
// signatures -- no comparison sort, no host loop over lines.  Signature
// products are one thread per line in CSR order, so a run is deterministic
// and equals a sequential sum.  Class member lists (for the ordered sums)
// are one more radix sort; the folded matrix is a radix sort of
// (row class, column class) keys, segment sums in sorted order and the
// stable transpose.
#include "gp_fold.hpp"

#include <chrono>

namespace gp {

namespace {

constexpr double kMaxFoldKeep = 0.8;  // reduction_too_small (folding_refinement.cu:24-37)

// SplitMix64 finaliser (Steele, Lea and Flood; constants of its published form)
__host__ __device__ inline uint64_t splitmix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

uint64_t round_seed(int32_t round, uint64_t side) {
  return splitmix(static_cast<uint64_t>(round) ^ (side * 0x9e3779b97f4a7c15ULL));
}

// a double in (0, 1): the top 53 bits of the hash, 0 moved to 2^-53
__device__ inline double unit53(uint64_t x) {
  const double u = static_cast<double>(splitmix(x) >> 11) * 0x1.0p-53;
  return u > 0.0 ? u : 0x1.0p-53;
}

// signature of class `cls` for one (round, side) seed: Box-Muller normal
__device__ inline double class_signature(uint64_t seed, int32_t cls) {
  const uint64_t c = static_cast<uint64_t>(cls);
  const double u1 = unit53(seed ^ (c * 0xbf58476d1ce4e5b9ULL));
  const double u2 = unit53((seed + 0x9e3779b97f4a7c15ULL) ^ (c * 0x94d049bb133111ebULL));
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

// float bits mapped to an unsigned order (negative values reversed)
__device__ inline uint32_t float_order(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float order_float(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void k_class_keys(int32_t n, const int32_t* __restrict__ cls, const double* __restrict__ value,
                             uint64_t* __restrict__ key, int32_t* __restrict__ idx) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = (static_cast<uint64_t>(static_cast<uint32_t>(cls[i])) << 32) |
           float_order(static_cast<float>(value[i]));
  idx[i] = static_cast<int32_t>(i);
}

__device__ inline int32_t class_start(const uint64_t* key, int64_t k, double tol) {
  if (k == 0) return 1;
  const uint64_t a = key[k - 1], b = key[k];
  if ((a >> 32) != (b >> 32)) return 1;
  const double fa = static_cast<double>(order_float(static_cast<uint32_t>(a)));
  const double fb = static_cast<double>(order_float(static_cast<uint32_t>(b)));
  return fb > fa + tol ? 1 : 0;
}

__global__ void k_class_starts(int32_t n, const uint64_t* __restrict__ key, double tol, int32_t* __restrict__ pos) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) pos[k] = class_start(key, k, tol);
}

// new class of the line at sorted position k: its group's rank (pos holds the
// exclusive scan of the start flags); with sig, the new class's signature
__global__ void k_class_scatter(int32_t n, const uint64_t* __restrict__ key, const int32_t* __restrict__ idx,
                                const int32_t* __restrict__ pos, double tol, int32_t* __restrict__ cls,
                                double* __restrict__ sig, uint64_t seed) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t c = pos[k] + class_start(key, k, tol);
  const int32_t i = idx[k];
  cls[i] = c;
  if (sig) sig[i] = class_signature(seed, c);
}

__global__ void k_write_signatures(int32_t n, const int32_t* __restrict__ cls, uint64_t seed,
                                   double* __restrict__ sig) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) sig[i] = class_signature(seed, cls[i]);
}

// out[i] = sum over line i of vals * sig[col], left to right from 0
__global__ void k_signature_product(gp_csr M, const double* __restrict__ sig, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= M.rows) return;
  double s = 0.0;
  for (gp_offset p = M.row_ptr[i]; p < M.row_ptr[i + 1]; ++p) s += M.vals[p] * sig[M.col_idx[p]];
  out[i] = s;
}

__global__ void k_fill_i32(int64_t n, int32_t v, int32_t* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}

// One refinement of a side's classes by a value per line.
class ClassRefiner {
 public:
  explicit ClassRefiner(int32_t n) : n_(n), key_(static_cast<size_t>(n)), idx_(static_cast<size_t>(n)),
                                     pos_(static_cast<size_t>(n)) {}
  // returns the new number of classes
  int32_t refine(DVec<int32_t>& cls, int32_t classes, const double* value, double tol, double* sig, uint64_t seed) {
    if (n_ == 0) return 0;
    hipLaunchKernelGGL(k_class_keys, dim3(grid_for(n_)), dim3(kBlock), 0, 0, n_,
                       static_cast<const int32_t*>(cls.data()), value, key_.data(), idx_.data());
    GP_LAUNCH_CHECK("class keys");
    sort_pairs_u64_i32(key_.data(), idx_.data(), n_, 32 + bits_for(static_cast<int64_t>(classes) + 1));
    hipLaunchKernelGGL(k_class_starts, dim3(grid_for(n_)), dim3(kBlock), 0, 0, n_,
                       static_cast<const uint64_t*>(key_.data()), tol, pos_.data());
    GP_LAUNCH_CHECK("class starts");
    const int64_t total = exclusive_scan_i32_total(pos_.data(), n_);
    hipLaunchKernelGGL(k_class_scatter, dim3(grid_for(n_)), dim3(kBlock), 0, 0, n_,
                       static_cast<const uint64_t*>(key_.data()), static_cast<const int32_t*>(idx_.data()),
                       static_cast<const int32_t*>(pos_.data()), tol, cls.data(), sig, seed);
    GP_LAUNCH_CHECK("class scatter");
    return static_cast<int32_t>(total);
  }

 private:
  int32_t n_;
  DVec<uint64_t> key_;
  DVec<int32_t> idx_, pos_;
};

bool too_little(int32_t row_classes, int32_t col_classes, int32_t m, int32_t n) {
  return static_cast<double>(row_classes) > kMaxFoldKeep * m && static_cast<double>(col_classes) > kMaxFoldKeep * n;
}

// ---- the folded LP -------------------------------------------------------------------
__global__ void k_class_count(int32_t n, const int32_t* __restrict__ cls, int32_t* __restrict__ count) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(count + (cls[i] - 1), 1);
}

__global__ void k_member_keys(int32_t n, const int32_t* __restrict__ cls, uint64_t* __restrict__ key) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) key[i] = (static_cast<uint64_t>(static_cast<uint32_t>(cls[i] - 1)) << 32) | static_cast<uint32_t>(i);
}

// out[k] = sum of v over class k's members in increasing index (divided by
// the class size when `average`), from 0
__global__ void k_class_sums(int32_t classes, const uint64_t* __restrict__ members, const int32_t* __restrict__ start,
                             const int32_t* __restrict__ count, const double* __restrict__ v, int average,
                             double* __restrict__ out) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= classes) return;
  double s = 0.0;
  const int32_t b = start[k], c = count[k];
  for (int32_t p = b; p < b + c; ++p) s += v[static_cast<uint32_t>(members[p])];
  out[k] = average ? s / static_cast<double>(c) : s;
}

__global__ void k_scale_of(int32_t n, const int32_t* __restrict__ cls, const int32_t* __restrict__ count,
                           double* __restrict__ scale) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) scale[i] = 1.0 / static_cast<double>(count[cls[i] - 1]);
}

// entry keys (row class - 1) * col_classes + (col class - 1), values a / |row class|
__global__ void k_fold_entries(gp_csr A, const int32_t* __restrict__ row_cls, const int32_t* __restrict__ col_cls,
                               const int32_t* __restrict__ row_count, int32_t col_classes,
                               uint64_t* __restrict__ key, int32_t* __restrict__ order, double* __restrict__ val) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  const int32_t r = row_cls[i] - 1;
  const double w = 1.0 / static_cast<double>(row_count[r]);
  for (gp_offset p = A.row_ptr[i]; p < A.row_ptr[i + 1]; ++p) {
    key[p] = static_cast<uint64_t>(r) * static_cast<uint64_t>(col_classes) +
             static_cast<uint64_t>(col_cls[A.col_idx[p]] - 1);
    order[p] = static_cast<int32_t>(p);
    val[p] = A.vals[p] * w;
  }
}

// head[k] = 1 at the first entry of a key that survives (|sum| > tol)
__global__ void k_fold_heads(int64_t nnz, const uint64_t* __restrict__ key, const int32_t* __restrict__ order,
                             const double* __restrict__ val, double tol, int32_t* __restrict__ keep,
                             double* __restrict__ sum) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k > nnz) return;
  if (k == nnz) {
    keep[k] = 0;
    return;
  }
  const uint64_t kk = key[k];
  if (k > 0 && key[k - 1] == kk) {
    keep[k] = 0;
    return;
  }
  double s = 0.0;
  for (int64_t t = k; t < nnz && key[t] == kk; ++t) s += val[order[t]];
  sum[k] = s;
  keep[k] = fabs(s) > tol ? 1 : 0;
}

__global__ void k_fold_emit(int64_t nnz, const uint64_t* __restrict__ key, const int32_t* __restrict__ keep_pos,
                            const double* __restrict__ sum, double tol, int32_t col_classes,
                            int32_t* __restrict__ row_of, int32_t* __restrict__ col_idx, double* __restrict__ vals) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  if (k > 0 && key[k - 1] == key[k]) return;
  if (!(fabs(sum[k]) > tol)) return;
  const int32_t q = keep_pos[k];
  row_of[q] = static_cast<int32_t>(key[k] / static_cast<uint64_t>(col_classes));
  col_idx[q] = static_cast<int32_t>(key[k] % static_cast<uint64_t>(col_classes));
  vals[q] = sum[k];
}

// row pointers of a row-sorted entry list: position p opens rows
// (row of p-1, row of p]; p = nnz closes through `rows`
__global__ void k_rows_from_sorted(const int32_t* __restrict__ row_of, int64_t nnz, int32_t rows,
                                   gp_offset* __restrict__ row_ptr) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p > nnz) return;
  const int64_t cur = p < nnz ? row_of[p] : rows;
  const int64_t prev = p > 0 ? row_of[p - 1] : -1;
  for (int64_t r = prev + 1; r <= cur; ++r) row_ptr[r] = static_cast<gp_offset>(p);
}

struct ClassLists {
  DVec<int32_t> count, start;
  DVec<uint64_t> members;
  void build(int32_t n, int32_t classes, const int32_t* cls) {
    count.resize(static_cast<size_t>(classes));
    start.resize(static_cast<size_t>(classes));
    members.resize(static_cast<size_t>(n));
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(classes)), dim3(kBlock), 0, 0, static_cast<int64_t>(classes), 0,
                       count.data());
    hipLaunchKernelGGL(k_class_count, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, cls, count.data());
    hipLaunchKernelGGL(k_member_keys, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, cls, members.data());
    GP_LAUNCH_CHECK("class members");
    sort_keys_u64(members.data(), n, 32 + bits_for(classes));
    d2d(start.data(), count.data(), static_cast<size_t>(classes));
    exclusive_scan_i32(start.data(), classes);
  }
  void sums(int32_t classes, const double* v, bool average, double* out) const {
    hipLaunchKernelGGL(k_class_sums, dim3(grid_for(classes)), dim3(kBlock), 0, 0, classes,
                       static_cast<const uint64_t*>(members.data()), static_cast<const int32_t*>(start.data()),
                       static_cast<const int32_t*>(count.data()), v, average ? 1 : 0, out);
    GP_LAUNCH_CHECK("class sums");
  }
};

void build_folded_lp(const gp_device_lp& lp, double tol, FoldState& f) {
  const int32_t m = lp.A.rows, n = lp.A.cols, mf = f.mf, nf = f.nf;
  ClassLists rows, cols;
  rows.build(m, mf, f.row_class.data());
  cols.build(n, nf, f.col_class.data());
  gp_device_lp& out = f.lp;
  out = gp_device_lp{};
  out.c = pool_alloc<double>(static_cast<size_t>(nf));
  out.l = pool_alloc<double>(static_cast<size_t>(nf));
  out.u = pool_alloc<double>(static_cast<size_t>(nf));
  out.AL = pool_alloc<double>(static_cast<size_t>(mf));
  out.AU = pool_alloc<double>(static_cast<size_t>(mf));
  out.obj_constant = lp.obj_constant;
  cols.sums(nf, lp.c, false, out.c);
  cols.sums(nf, lp.l, true, out.l);
  cols.sums(nf, lp.u, true, out.u);
  rows.sums(mf, lp.AL, true, out.AL);
  rows.sums(mf, lp.AU, true, out.AU);
  f.row_scale.resize(static_cast<size_t>(m));
  f.col_scale.resize(static_cast<size_t>(n));
  hipLaunchKernelGGL(k_scale_of, dim3(grid_for(m)), dim3(kBlock), 0, 0, m,
                     static_cast<const int32_t*>(f.row_class.data()), static_cast<const int32_t*>(rows.count.data()),
                     f.row_scale.data());
  hipLaunchKernelGGL(k_scale_of, dim3(grid_for(n)), dim3(kBlock), 0, 0, n,
                     static_cast<const int32_t*>(f.col_class.data()), static_cast<const int32_t*>(cols.count.data()),
                     f.col_scale.data());
  GP_LAUNCH_CHECK("fold scales");

  // matrix
  const int64_t nnz = lp.A.nnz;
  out.A.rows = mf;
  out.A.cols = nf;
  out.A.row_ptr = pool_alloc<gp_offset>(static_cast<size_t>(mf) + 1);
  int64_t kept = 0;
  DVec<int32_t> row_of;
  if (nnz > 0) {
    DVec<uint64_t> key(static_cast<size_t>(nnz));
    DVec<int32_t> order(static_cast<size_t>(nnz)), keep(static_cast<size_t>(nnz) + 1);
    DVec<double> val(static_cast<size_t>(nnz)), sum(static_cast<size_t>(nnz));
    hipLaunchKernelGGL(k_fold_entries, dim3(grid_for(m)), dim3(kBlock), 0, 0, lp.A,
                       static_cast<const int32_t*>(f.row_class.data()), static_cast<const int32_t*>(f.col_class.data()),
                       static_cast<const int32_t*>(rows.count.data()), nf, key.data(), order.data(), val.data());
    GP_LAUNCH_CHECK("fold entries");
    sort_pairs_u64_i32(key.data(), order.data(), nnz, bits_for(static_cast<int64_t>(mf) * nf));
    hipLaunchKernelGGL(k_fold_heads, dim3(grid_for(nnz + 1)), dim3(kBlock), 0, 0, nnz,
                       static_cast<const uint64_t*>(key.data()), static_cast<const int32_t*>(order.data()),
                       static_cast<const double*>(val.data()), tol, keep.data(), sum.data());
    GP_LAUNCH_CHECK("fold heads");
    kept = exclusive_scan_i32_total(keep.data(), nnz + 1);
    out.A.col_idx = pool_alloc<int32_t>(static_cast<size_t>(kept));
    out.A.vals = pool_alloc<double>(static_cast<size_t>(kept));
    row_of.resize(static_cast<size_t>(kept));
    hipLaunchKernelGGL(k_fold_emit, dim3(grid_for(nnz)), dim3(kBlock), 0, 0, nnz,
                       static_cast<const uint64_t*>(key.data()), static_cast<const int32_t*>(keep.data()),
                       static_cast<const double*>(sum.data()), tol, nf, row_of.data(), out.A.col_idx, out.A.vals);
    GP_LAUNCH_CHECK("fold emit");
  } else {
    out.A.col_idx = pool_alloc<int32_t>(0);
    out.A.vals = pool_alloc<double>(0);
  }
  out.A.nnz = static_cast<gp_offset>(kept);
  hipLaunchKernelGGL(k_rows_from_sorted, dim3(grid_for(kept + 1)), dim3(kBlock), 0, 0,
                     static_cast<const int32_t*>(row_of.data()), kept, mf, out.A.row_ptr);
  GP_LAUNCH_CHECK("fold row pointers");
  transpose_csr(out.A, out.AT);
}

__global__ void k_unfold_cols(int32_t n, const int32_t* __restrict__ cls, const double* __restrict__ scale,
                              const double* __restrict__ x_red, const double* __restrict__ z_red,
                              double* __restrict__ x, double* __restrict__ z) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int32_t c = cls[j] - 1;
  if (x && x_red) x[j] = x_red[c];
  if (z && z_red) z[j] = z_red[c] * scale[j];
}

__global__ void k_unfold_rows(int32_t m, const int32_t* __restrict__ cls, const double* __restrict__ scale,
                              const double* __restrict__ y_red, double* __restrict__ y) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) y[i] = y_red[cls[i] - 1] * scale[i];
}

}  // namespace

FoldState::~FoldState() {
  if (!applied) return;
  csr_pool_free(lp.A);
  csr_pool_free(lp.AT);
  pool_free(lp.c);
  pool_free(lp.l);
  pool_free(lp.u);
  pool_free(lp.AL);
  pool_free(lp.AU);
}

void FoldState::fill(gp_fold_result& r) const {
  std::memset(&r, 0, sizeof(r));
  r.applied = applied ? 1 : 0;
  r.original_rows = m;
  r.original_cols = n;
  r.folded_rows = applied ? mf : m;
  r.folded_cols = applied ? nf : n;
  r.rounds = rounds;
  r.refinement_ok = refinement_ok ? 1 : 0;
  r.final_row_colors = mf;
  r.final_col_colors = nf;
  r.seconds = seconds;
  if (applied) {
    r.folded_lp = lp;
    r.row_color_id = const_cast<int32_t*>(row_class.data());
    r.col_color_id = const_cast<int32_t*>(col_class.data());
    r.row_scale = const_cast<double*>(row_scale.data());
    r.col_scale = const_cast<double*>(col_scale.data());
  }
}

std::unique_ptr<FoldState> fold_lp(const gp_device_lp& lp, double tol) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!std::isfinite(tol) || tol < 0.0) throw std::invalid_argument("folding tolerance must be finite and nonnegative");
  auto f = std::make_unique<FoldState>();
  const int32_t m = lp.A.rows, n = lp.A.cols;
  f->m = f->mf = m;
  f->n = f->nf = n;
  auto done = [&] {
    GP_CHECK(hipDeviceSynchronize());
    f->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return std::move(f);
  };
  if (m == 0 || n == 0) return done();
  f->row_class.resize(static_cast<size_t>(m));
  f->col_class.resize(static_cast<size_t>(n));
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(m)), dim3(kBlock), 0, 0, static_cast<int64_t>(m), 1,
                     f->row_class.data());
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(kBlock), 0, 0, static_cast<int64_t>(n), 1,
                     f->col_class.data());
  DVec<double> row_sig(static_cast<size_t>(m)), col_sig(static_cast<size_t>(n));
  ClassRefiner rref(m), cref(n);
  // initial classes (folding_refinement.cu:235-262)
  int32_t mc = rref.refine(f->row_class, 1, lp.AL, tol, nullptr, 0);
  mc = rref.refine(f->row_class, mc, lp.AU, tol, nullptr, 0);
  int32_t nc = cref.refine(f->col_class, 1, lp.c, tol, nullptr, 0);
  nc = cref.refine(f->col_class, nc, lp.l, tol, nullptr, 0);
  nc = cref.refine(f->col_class, nc, lp.u, tol, nullptr, 0);
  hipLaunchKernelGGL(k_write_signatures, dim3(grid_for(m)), dim3(kBlock), 0, 0, m,
                     static_cast<const int32_t*>(f->row_class.data()), round_seed(0, 2), row_sig.data());
  hipLaunchKernelGGL(k_write_signatures, dim3(grid_for(n)), dim3(kBlock), 0, 0, n,
                     static_cast<const int32_t*>(f->col_class.data()), round_seed(0, 5), col_sig.data());
  GP_LAUNCH_CHECK("initial signatures");
  f->mf = mc;
  f->nf = nc;
  if (too_little(mc, nc, m, n)) {
    f->refinement_ok = false;
    return done();
  }
  DVec<double> prod_c(static_cast<size_t>(n)), prod_r(static_cast<size_t>(m));
  const int32_t max_round = std::max<int32_t>(1, m + n);
  for (int32_t round = 1; round <= max_round; ++round) {
    f->rounds = round;
    hipLaunchKernelGGL(k_signature_product, dim3(grid_for(n)), dim3(kBlock), 0, 0, lp.AT,
                       static_cast<const double*>(row_sig.data()), prod_c.data());
    GP_LAUNCH_CHECK("column signature product");
    const int32_t nc_new = cref.refine(f->col_class, nc, prod_c.data(), tol, col_sig.data(), round_seed(round, 1));
    hipLaunchKernelGGL(k_signature_product, dim3(grid_for(m)), dim3(kBlock), 0, 0, lp.A,
                       static_cast<const double*>(col_sig.data()), prod_r.data());
    GP_LAUNCH_CHECK("row signature product");
    const int32_t mc_new = rref.refine(f->row_class, mc, prod_r.data(), tol, row_sig.data(), round_seed(round, 2));
    const bool changed = nc_new > nc || mc_new > mc;
    nc = f->nf = nc_new;
    mc = f->mf = mc_new;
    if (too_little(mc, nc, m, n)) {
      f->refinement_ok = false;
      return done();
    }
    if (!changed) break;
  }
  if (mc == m && nc == n) return done();
  build_folded_lp(lp, tol, *f);
  f->applied = true;
  return done();
}

void unfold(const FoldState& f, const double* x_red, const double* y_red, const double* z_red, double* x, double* y,
            double* z) {
  if (!f.applied) throw std::invalid_argument("unfold: no folding was applied");
  if (f.n > 0 && (x || z))
    hipLaunchKernelGGL(k_unfold_cols, dim3(grid_for(f.n)), dim3(kBlock), 0, 0, f.n,
                       static_cast<const int32_t*>(f.col_class.data()), static_cast<const double*>(f.col_scale.data()),
                       x_red, z_red, x, z);
  if (f.m > 0 && y && y_red)
    hipLaunchKernelGGL(k_unfold_rows, dim3(grid_for(f.m)), dim3(kBlock), 0, 0, f.m,
                       static_cast<const int32_t*>(f.row_class.data()), static_cast<const double*>(f.row_scale.data()),
                       y_red, y);
  GP_LAUNCH_CHECK("unfold");
}

}  // namespace gp

extern "C" int gp_fold(const gp_device_lp* lp, double tolerance, gp_fold_result* out) {
  return gp::guarded([&] {
    if (!lp || !out) throw std::invalid_argument("gp_fold: null argument");
    std::memset(out, 0, sizeof(*out));
    std::unique_ptr<gp::FoldState> f = gp::fold_lp(*lp, tolerance);
    f->fill(*out);
    out->state = f.release();
  });
}

extern "C" int gp_unfold(const gp_fold_result* fold, const double* x_red, const double* y_red, const double* z_red,
                         double* x, double* y, double* z) {
  return gp::guarded([&] {
    if (!fold || !fold->state) throw std::invalid_argument("gp_unfold: no fold");
    const gp::FoldState& f = *static_cast<const gp::FoldState*>(fold->state);
    gp::DVec<double> dx(static_cast<size_t>(f.n ? f.n : 1)), dy(static_cast<size_t>(f.m ? f.m : 1)),
        dz(static_cast<size_t>(f.n ? f.n : 1));
    gp::unfold(f, x_red, y_red, z_red, x ? dx.data() : nullptr, y ? dy.data() : nullptr, z ? dz.data() : nullptr);
    if (x && x_red) gp::d2h(x, dx.data(), static_cast<size_t>(f.n));
    if (y && y_red) gp::d2h(y, dy.data(), static_cast<size_t>(f.m));
    if (z && z_red) gp::d2h(z, dz.data(), static_cast<size_t>(f.n));
  });
}

extern "C" void gp_free_fold(gp_fold_result* fold) {
  if (!fold) return;
  delete static_cast<gp::FoldState*>(fold->state);
  std::memset(fold, 0, sizeof(*fold));
}
