// gp_lpsolve.hip -- FP64 first-order LP solver on the device (gp_lp_solve):
// restarted Halpern iteration of the Peaceman-Rachford / PDHG operator, the
// method of HPR-LP (third_party/hprlp: solver/solve.cu, the x and y updates of
// cuda_kernels/backends/simple/simple_update_kernels.cu:3-95, Halpern factors
// 1/(k+2) of cuda_kernels/shared/halpern_kernels.cu:3-11), for
//
//   min c'x + obj_constant   s.t.  AL <= A x <= AU,  l <= x <= u.
//
// One iteration, with primal step sigma and dual step 1/(sigma lambda),
// lambda >= ||A||_2^2:
//   x_bar = clamp(x + sigma (A'y - c), l, u),   x_hat = 2 x_bar - x
//   v     = A x_hat - sigma lambda y,           y_bar = (clamp(v, AL, AU) - v) / (sigma lambda)
//   y_hat = 2 y_bar - y
//   (x, y) <- (k+1)/(k+2) (x_hat, y_hat) + 1/(k+2) (x0, y0)        (Halpern)
// The anchor (x0, y0) and k restart at (x_bar, y_bar) when the fixed-point
// residual has dropped enough; sigma is rebalanced at every restart from the
// primal and dual movement since the last one.  The problem is first scaled
// (Ruiz equilibration, Pock-Chambolle, then bound and cost norms) and the
// termination test is on the unscaled problem: relative primal residual,
// dual residual and duality gap all <= tol.
//
// MI355X design: the two sparse products are fused with their updates --
// one launch over the columns (rows of A^T: gather of y, dot product, clamp,
// reflection, Halpern step) and one over the rows of A (gather of x_hat,
// projection onto [AL, AU], reflection, Halpern step) -- so an iteration
// reads each matrix once and each vector a small number of times.  Lanes
// per row follow the mean row length; every reduction is a two-stage block
// sum (fixed order, so a run is reproducible).
#include "gp_common.hpp"

#include <chrono>

namespace gp {

namespace {

constexpr int kRedBlocks = 1024;  // partial-sum blocks of the check reductions
constexpr int kRedWidth = 4;      // sums per reduction

template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

// Block reduction of kRedWidth values into partial[blockIdx.x * kRedWidth + t].
__device__ inline void block_partials(double (&v)[kRedWidth], double* partial) {
  __shared__ double s[kBlock / 64][kRedWidth];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < kRedWidth; ++t) {
    double x = v[t];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) s[wave][t] = x;
  }
  __syncthreads();
  if (threadIdx.x < kRedWidth) {
    double x = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) x += s[w][threadIdx.x];
    partial[blockIdx.x * kRedWidth + threadIdx.x] = x;
  }
}

// out[t] = sum over blocks of partial[b * kRedWidth + t]; one block.
__global__ void __launch_bounds__(kBlock) k_finish_sums(const double* __restrict__ partial, int blocks,
                                                        double* __restrict__ out) {
  __shared__ double s[kBlock][kRedWidth];
  double v[kRedWidth] = {};
  for (int b = threadIdx.x; b < blocks; b += kBlock)
    for (int t = 0; t < kRedWidth; ++t) v[t] += partial[b * kRedWidth + t];
  for (int t = 0; t < kRedWidth; ++t) s[threadIdx.x][t] = v[t];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int t = 0; t < kRedWidth; ++t) s[threadIdx.x][t] += s[threadIdx.x + w][t];
    __syncthreads();
  }
  if (threadIdx.x < kRedWidth) out[threadIdx.x] = s[0][threadIdx.x];
}

// ---- the iteration ------------------------------------------------------------------
struct Step {
  double sigma;     // primal step
  double inv_dual;  // sigma * lambda (the inverse of the dual step)
  double f_anchor;  // 1 / (k + 2)
  double f_point;   // 1 - f_anchor
};

// Columns j (rows of A^T): x_bar, x_hat, Halpern step of x.  With x_bar
// non-null (check iterations) also stores x_bar and sums (x_bar - x)^2.
template <int G>
__global__ void __launch_bounds__(kBlock) k_primal(gp_csr AT, const double* __restrict__ y,
                                                   const double* __restrict__ c, const double* __restrict__ l,
                                                   const double* __restrict__ u, double* __restrict__ x,
                                                   const double* __restrict__ x0, double* __restrict__ x_hat,
                                                   double* __restrict__ x_bar, Step st, double* partial) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  double acc[kRedWidth] = {};
  for (int64_t j = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; j < AT.rows; j += groups) {
    double aty = 0.0;
    const gp_offset p1 = AT.row_ptr[j + 1];
    for (gp_offset p = AT.row_ptr[j] + lane; p < p1; p += G) aty += AT.vals[p] * y[AT.col_idx[p]];
    aty = group_sum<G>(aty);
    if (lane == 0) {
      const double xj = x[j];
      const double xb = fmin(u[j], fmax(l[j], xj + st.sigma * (aty - c[j])));
      const double xh = 2.0 * xb - xj;
      x_hat[j] = xh;
      x[j] = st.f_point * xh + st.f_anchor * x0[j];
      if (x_bar) {
        x_bar[j] = xb;
        acc[0] += (xb - xj) * (xb - xj);
      }
    }
  }
  if (partial) block_partials(acc, partial);
}

// Rows i of A: y_bar, Halpern step of y; on check iterations also stores
// y_bar and sums (y_bar - y)^2.
template <int G>
__global__ void __launch_bounds__(kBlock) k_dual(gp_csr A, const double* __restrict__ x_hat,
                                                 const double* __restrict__ AL, const double* __restrict__ AU,
                                                 double* __restrict__ y, const double* __restrict__ y0,
                                                 double* __restrict__ y_bar, Step st, double* partial) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  double acc[kRedWidth] = {};
  for (int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; i < A.rows; i += groups) {
    double ax = 0.0;
    const gp_offset p1 = A.row_ptr[i + 1];
    for (gp_offset p = A.row_ptr[i] + lane; p < p1; p += G) ax += A.vals[p] * x_hat[A.col_idx[p]];
    ax = group_sum<G>(ax);
    if (lane == 0) {
      const double yi = y[i];
      const double v = ax - st.inv_dual * yi;
      const double yb = (fmin(AU[i], fmax(AL[i], v)) - v) / st.inv_dual;
      const double yh = 2.0 * yb - yi;
      y[i] = st.f_point * yh + st.f_anchor * y0[i];
      if (y_bar) {
        y_bar[i] = yb;
        acc[0] += (yb - yi) * (yb - yi);
      }
    }
  }
  if (partial) block_partials(acc, partial);
}

// The normal (non-check) iterations walk the lines of one length bin: the
// bin's line list picks the line, and every line of a launch gets the same
// lanes, so a long column does not hold up the short ones sharing its wave.
template <int G>
__global__ void __launch_bounds__(kBlock) k_primal_bin(gp_csr AT, const int32_t* __restrict__ lines, int32_t count,
                                                       const double* __restrict__ y, const double* __restrict__ c,
                                                       const double* __restrict__ l, const double* __restrict__ u,
                                                       double* __restrict__ x, const double* __restrict__ x0,
                                                       double* __restrict__ x_hat, Step st) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t g = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; g < count; g += groups) {
    const int32_t j = lines[g];
    double aty = 0.0;
    const gp_offset p1 = AT.row_ptr[j + 1];
    for (gp_offset p = AT.row_ptr[j] + lane; p < p1; p += G)
      aty += AT.vals[p] * y[AT.col_idx[p]];
    aty = group_sum<G>(aty);
    if (lane == 0) {
      const double xj = x[j];
      const double xb = fmin(u[j], fmax(l[j], xj + st.sigma * (aty - c[j])));
      const double xh = 2.0 * xb - xj;
      x_hat[j] = xh;
      x[j] = st.f_point * xh + st.f_anchor * x0[j];
    }
  }
}

template <int G>
__global__ void __launch_bounds__(kBlock) k_dual_bin(gp_csr A, const int32_t* __restrict__ lines, int32_t count,
                                                     const double* __restrict__ x_hat, const double* __restrict__ AL,
                                                     const double* __restrict__ AU, double* __restrict__ y,
                                                     const double* __restrict__ y0, Step st) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t g = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; g < count; g += groups) {
    const int32_t i = lines[g];
    double ax = 0.0;
    const gp_offset p1 = A.row_ptr[i + 1];
    for (gp_offset p = A.row_ptr[i] + lane; p < p1; p += G)
      ax += A.vals[p] * x_hat[A.col_idx[p]];
    ax = group_sum<G>(ax);
    if (lane == 0) {
      const double yi = y[i];
      const double v = ax - st.inv_dual * yi;
      const double yb = (fmin(AU[i], fmax(AL[i], v)) - v) / st.inv_dual;
      y[i] = st.f_point * (2.0 * yb - yi) + st.f_anchor * y0[i];
    }
  }
}

template <int G>
struct PrimalBinK {
  static constexpr auto kernel = k_primal_bin<G>;
};
template <int G>
struct DualBinK {
  static constexpr auto kernel = k_dual_bin<G>;
};

// Lines of a matrix in at most two bins: lines of up to 64 entries with the
// lanes their mean length calls for, and longer lines with a whole wave each
// (a heavy-tailed column degree otherwise holds up every short line sharing
// the long one's wave).  Each bin is a device list of line indices in
// increasing order.  A matrix without long lines has one bin.
struct LineBins {
  static constexpr int kBins = 2;
  static constexpr int kLong = 64;
  int lanes[kBins] = {4, 64};
  int32_t count[kBins] = {};
  DVec<int32_t> lines[kBins];

  void build(const gp_csr& M) {
    std::vector<gp_offset> ptr(static_cast<size_t>(M.rows) + 1);
    d2h(ptr.data(), M.row_ptr, ptr.size());
    std::vector<int32_t> host[kBins];
    int64_t short_nnz = 0;
    for (int32_t i = 0; i < M.rows; ++i) {
      const gp_offset len = ptr[static_cast<size_t>(i) + 1] - ptr[static_cast<size_t>(i)];
      host[len > kLong ? 1 : 0].push_back(i);
      if (len <= kLong) short_nnz += len;
    }
    lanes[0] = lanes_for_mean(host[0].empty() ? 0.0 : static_cast<double>(short_nnz) / host[0].size());
    for (int b = 0; b < kBins; ++b) {
      count[b] = static_cast<int32_t>(host[b].size());
      lines[b].resize(host[b].size());
      h2d(lines[b].data(), host[b].data(), host[b].size());
    }
  }
};

template <int G>
struct PrimalK {
  static constexpr auto kernel = k_primal<G>;
};
template <int G>
struct DualK {
  static constexpr auto kernel = k_dual<G>;
};

// ---- termination measures (scaled quantities, unscaled by dr / dc) -------------------
// Rows: (A x_bar - proj(A x_bar))_i in original units, the rows' share of the
// dual objective, and the dual movement since the anchor.
template <int G>
__global__ void __launch_bounds__(kBlock) k_check_rows(gp_csr A, const double* __restrict__ x_bar,
                                                       const double* __restrict__ y_bar,
                                                       const double* __restrict__ y0,
                                                       const double* __restrict__ AL, const double* __restrict__ AU,
                                                       const double* __restrict__ dr, double b_scale,
                                                       double* partial) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  double acc[kRedWidth] = {};
  for (int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; i < A.rows; i += groups) {
    double ax = 0.0;
    const gp_offset p1 = A.row_ptr[i + 1];
    for (gp_offset p = A.row_ptr[i] + lane; p < p1; p += G) ax += A.vals[p] * x_bar[A.col_idx[p]];
    ax = group_sum<G>(ax);
    if (lane == 0) {
      const double viol = (ax - fmin(AU[i], fmax(AL[i], ax))) * (b_scale / dr[i]);
      const double yb = y_bar[i];
      acc[0] += viol * viol;
      if (yb > 0.0) acc[1] += yb * AL[i];
      else if (yb < 0.0) acc[1] += yb * AU[i];
      acc[2] += (yb - y0[i]) * (yb - y0[i]);
    }
  }
  block_partials(acc, partial);
}

// Columns: dual residual c - A'y - z (z the projection of the reduced cost
// onto the signs the bounds allow) in original units, the columns' share of
// the dual objective, the primal objective and the primal movement.
template <int G>
__global__ void __launch_bounds__(kBlock) k_check_cols(gp_csr AT, const double* __restrict__ y_bar,
                                                       const double* __restrict__ x_bar,
                                                       const double* __restrict__ x0, const double* __restrict__ c,
                                                       const double* __restrict__ l, const double* __restrict__ u,
                                                       const double* __restrict__ dc, double c_scale,
                                                       double* partial) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  double acc[kRedWidth] = {};
  for (int64_t j = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; j < AT.rows; j += groups) {
    double aty = 0.0;
    const gp_offset p1 = AT.row_ptr[j + 1];
    for (gp_offset p = AT.row_ptr[j] + lane; p < p1; p += G) aty += AT.vals[p] * y_bar[AT.col_idx[p]];
    aty = group_sum<G>(aty);
    if (lane == 0) {
      const double g = c[j] - aty;
      double z = 0.0;
      if (g > 0.0 && isfinite(l[j])) z = g;
      if (g < 0.0 && isfinite(u[j])) z = g;
      const double r = (g - z) * (c_scale / dc[j]);
      acc[0] += r * r;
      if (z > 0.0) acc[1] += z * l[j];
      else if (z < 0.0) acc[1] += z * u[j];
      const double xb = x_bar[j];
      acc[2] += c[j] * xb;
      acc[3] += (xb - x0[j]) * (xb - x0[j]);
    }
  }
  block_partials(acc, partial);
}

// Unscaled solution: x = b_scale dc x_bar, y = c_scale dr y_bar,
// z = c_scale / dc (c - A'y_bar projected).
template <int G>
__global__ void __launch_bounds__(kBlock) k_unscale_cols(gp_csr AT, const double* __restrict__ y_bar,
                                                         const double* __restrict__ x_bar,
                                                         const double* __restrict__ c, const double* __restrict__ l,
                                                         const double* __restrict__ u, const double* __restrict__ dc,
                                                         double b_scale, double c_scale, double* __restrict__ x_out,
                                                         double* __restrict__ z_out) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t j = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; j < AT.rows; j += groups) {
    double aty = 0.0;
    const gp_offset p1 = AT.row_ptr[j + 1];
    for (gp_offset p = AT.row_ptr[j] + lane; p < p1; p += G) aty += AT.vals[p] * y_bar[AT.col_idx[p]];
    aty = group_sum<G>(aty);
    if (lane == 0) {
      const double g = c[j] - aty;
      double z = 0.0;
      if (g > 0.0 && isfinite(l[j])) z = g;
      if (g < 0.0 && isfinite(u[j])) z = g;
      if (x_out) x_out[j] = b_scale * dc[j] * x_bar[j];
      if (z_out) z_out[j] = c_scale / dc[j] * z;
    }
  }
}

__global__ void k_unscale_rows(int32_t m, const double* __restrict__ y_bar, const double* __restrict__ dr,
                               double c_scale, double* __restrict__ y_out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) y_out[i] = c_scale * dr[i] * y_bar[i];
}

template <int G>
struct CheckRowsK {
  static constexpr auto kernel = k_check_rows<G>;
};
template <int G>
struct CheckColsK {
  static constexpr auto kernel = k_check_cols<G>;
};
template <int G>
struct UnscaleColsK {
  static constexpr auto kernel = k_unscale_cols<G>;
};

// ---- scaling ----------------------------------------------------------------------------
// norm_kind 0: max |a| over the line; 1: sum |a| over the line.
template <int G>
__global__ void __launch_bounds__(kBlock) k_line_norm(gp_csr M, int norm_kind, double* __restrict__ out) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  for (int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; i < M.rows; i += groups) {
    double v = 0.0;
    const gp_offset p1 = M.row_ptr[i + 1];
    for (gp_offset p = M.row_ptr[i] + lane; p < p1; p += G) {
      const double a = fabs(M.vals[p]);
      v = norm_kind ? v + a : fmax(v, a);
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      const double w = __shfl_xor(v, o, G);
      v = norm_kind ? v + w : fmax(v, w);
    }
    if (lane == 0) out[i] = v;
  }
}
template <int G>
struct LineNormK {
  static constexpr auto kernel = k_line_norm<G>;
};

// f = 1 / sqrt(norm) (1 for an empty line); d *= f
__global__ void k_scale_factor(int64_t count, double* __restrict__ norm_to_factor, double* __restrict__ d) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double v = norm_to_factor[i];
  const double f = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
  norm_to_factor[i] = f;
  d[i] *= f;
}

// M.vals[p] *= f_line[line] * f_other[col_idx[p]] (the same product for the
// entry in A and in A^T, so the two stay exact transposes)
__global__ void k_scale_entries(gp_csr M, const double* __restrict__ f_line, const double* __restrict__ f_other,
                                int64_t nnz, const int32_t* __restrict__ line_of) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  M.vals[p] *= f_line[line_of[p]] * f_other[M.col_idx[p]];
}

__global__ void k_line_of(gp_csr M, int32_t* __restrict__ line_of) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= M.rows) return;
  for (gp_offset p = M.row_ptr[i]; p < M.row_ptr[i + 1]; ++p) line_of[p] = static_cast<int32_t>(i);
}

__global__ void k_fill(int64_t count, double v, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) out[i] = v;
}

// scaled vectors: c_s = dc c / c_scale, l_s = l / (dc b_scale), AL_s = dr AL / b_scale, ...
__global__ void k_scale_cols(int32_t n, const double* c, const double* l, const double* u, const double* dc,
                             double b_scale, double c_scale, double* cs, double* ls, double* us) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  cs[j] = dc[j] * c[j] / c_scale;
  ls[j] = l[j] / (dc[j] * b_scale);
  us[j] = u[j] / (dc[j] * b_scale);
}
__global__ void k_scale_rows(int32_t m, const double* AL, const double* AU, const double* dr, double b_scale,
                             double* ALs, double* AUs) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  ALs[i] = dr[i] * AL[i] / b_scale;
  AUs[i] = dr[i] * AU[i] / b_scale;
}

// squares of: |c_j| and the finite part of the row sides (max of |AL|, |AU|
// over finite sides), for the norms of the relative tolerances and scaling
__global__ void __launch_bounds__(kBlock) k_vec_norms(int32_t m, int32_t n, const double* c, const double* AL,
                                                      const double* AU, double* partial) {
  double acc[kRedWidth] = {};
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < static_cast<int64_t>(m) + n;
       t += stride) {
    if (t < n) {
      acc[0] += c[t] * c[t];
    } else {
      const int64_t i = t - n;
      double b = 0.0;
      if (isfinite(AL[i])) b = fabs(AL[i]);
      if (isfinite(AU[i])) b = fmax(b, fabs(AU[i]));
      acc[1] += b * b;
    }
  }
  block_partials(acc, partial);
}

// power iteration on A'A: w = A v; v_next = A' w, with sums of |A v|^2 and
// |v_next|^2 for the Rayleigh quotient and the normalisation
template <int G>
__global__ void __launch_bounds__(kBlock) k_spmv_sq(gp_csr M, const double* __restrict__ v, double scale,
                                                    double* __restrict__ out, double* partial) {
  const int lane = threadIdx.x % G;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (blockDim.x / G);
  double acc[kRedWidth] = {};
  for (int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G; i < M.rows; i += groups) {
    double s = 0.0;
    const gp_offset p1 = M.row_ptr[i + 1];
    for (gp_offset p = M.row_ptr[i] + lane; p < p1; p += G) s += M.vals[p] * v[M.col_idx[p]];
    s = group_sum<G>(s) * scale;
    if (lane == 0) {
      out[i] = s;
      acc[0] += s * s;
    }
  }
  block_partials(acc, partial);
}
template <int G>
struct SpmvSqK {
  static constexpr auto kernel = k_spmv_sq<G>;
};

__global__ void k_power_start(int32_t n, double* v) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  // deterministic, not orthogonal to any particular direction
  uint32_t h = static_cast<uint32_t>(j) * 2654435761u;
  h ^= h >> 15;
  v[j] = 0.5 + static_cast<double>(h & 0xffffu) / 65536.0;
}

// ---- the solver ---------------------------------------------------------------------------
// Partial-sum slots: a reduction kernel of `blocks` blocks fills blocks x
// kRedWidth partials of its slot; finish() sums exactly those.
struct Sums {
  DVec<double> partial{static_cast<size_t>(kRedBlocks) * kRedWidth * 4};
  DVec<double> out{static_cast<size_t>(kRedWidth) * 4};
  int blocks[4] = {};
  double h[kRedWidth * 4] = {};
  double* slot(int k, unsigned used) {
    blocks[k] = static_cast<int>(used);
    return partial.data() + static_cast<size_t>(k) * kRedBlocks * kRedWidth;
  }
  void finish(int k) {
    hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(kBlock), 0, 0,
                       partial.data() + static_cast<size_t>(k) * kRedBlocks * kRedWidth, blocks[k],
                       out.data() + k * kRedWidth);
    GP_LAUNCH_CHECK("finish sums");
  }
  const double* read(int k) {
    d2h(h, out.data(), static_cast<size_t>(k) * kRedWidth);
    return h;
  }
};

unsigned red_grid(int64_t lines, int lanes) {
  const int64_t g = grid_for(lines * lanes, kBlock);
  return static_cast<unsigned>(g < kRedBlocks ? g : kRedBlocks);
}

template <template <int> class K, class... Args>
void launch_reduce(int lanes, int64_t lines, Args... args) {
  const unsigned grid = red_grid(lines, lanes);
  switch (lanes) {
    case 2: hipLaunchKernelGGL(K<2>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
    case 4: hipLaunchKernelGGL(K<4>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
    case 8: hipLaunchKernelGGL(K<8>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
    case 16: hipLaunchKernelGGL(K<16>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
    default: hipLaunchKernelGGL(K<64>::kernel, dim3(grid), dim3(kBlock), 0, 0, args...); break;
  }
}

class LpSolver {
 public:
  LpSolver(const gp_device_lp& lp, const gp_lp_solve_params& p) : lp_(lp), p_(p), m_(lp.A.rows), n_(lp.A.cols) {}

  void run(double* x_out, double* y_out, double* z_out, gp_lp_solve_info* info);

 private:
  void scale();
  double power_lambda();
  void check(double& r_fixed, double& dx_anchor, double& dy_anchor, gp_lp_solve_info* info);

  const gp_device_lp& lp_;
  gp_lp_solve_params p_;
  int32_t m_, n_;
  int lanes_a_ = 4, lanes_at_ = 4;
  LineBins bins_a_, bins_at_;
  gp_csr A_{}, AT_{};  // scaled values, the input's structure
  DVec<double> a_vals_, at_vals_;
  DVec<double> dr_, dc_, c_, l_, u_, AL_, AU_;
  double b_scale_ = 1.0, c_scale_ = 1.0, norm_b_ = 0.0, norm_c_ = 0.0, lambda_ = 1.0;
  DVec<double> x_, y_, x0_, y0_, x_hat_, x_bar_, y_bar_;
  Sums sums_;
  double sigma_ = 1.0;
};

void LpSolver::scale() {
  const int64_t nnz = lp_.A.nnz;
  a_vals_.resize(static_cast<size_t>(nnz));
  at_vals_.resize(static_cast<size_t>(nnz));
  d2d(a_vals_.data(), lp_.A.vals, static_cast<size_t>(nnz));
  d2d(at_vals_.data(), lp_.AT.vals, static_cast<size_t>(nnz));
  A_ = lp_.A;
  A_.vals = a_vals_.data();
  AT_ = lp_.AT;
  AT_.vals = at_vals_.data();
  dr_.resize(static_cast<size_t>(m_));
  dc_.resize(static_cast<size_t>(n_));
  hipLaunchKernelGGL(k_fill, dim3(grid_for(m_)), dim3(kBlock), 0, 0, static_cast<int64_t>(m_), 1.0, dr_.data());
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n_)), dim3(kBlock), 0, 0, static_cast<int64_t>(n_), 1.0, dc_.data());
  DVec<double> fr(static_cast<size_t>(m_)), fc(static_cast<size_t>(n_));
  DVec<int32_t> row_of(static_cast<size_t>(nnz)), col_of(static_cast<size_t>(nnz));
  hipLaunchKernelGGL(k_line_of, dim3(grid_for(m_)), dim3(kBlock), 0, 0, A_, row_of.data());
  hipLaunchKernelGGL(k_line_of, dim3(grid_for(n_)), dim3(kBlock), 0, 0, AT_, col_of.data());
  GP_LAUNCH_CHECK("line of");
  auto pass = [&](int kind) {
    launch_grouped<LineNormK>(lanes_a_, m_, A_, kind, fr.data());
    launch_grouped<LineNormK>(lanes_at_, n_, AT_, kind, fc.data());
    hipLaunchKernelGGL(k_scale_factor, dim3(grid_for(m_)), dim3(kBlock), 0, 0, static_cast<int64_t>(m_), fr.data(),
                       dr_.data());
    hipLaunchKernelGGL(k_scale_factor, dim3(grid_for(n_)), dim3(kBlock), 0, 0, static_cast<int64_t>(n_), fc.data(),
                       dc_.data());
    if (nnz > 0) {
      hipLaunchKernelGGL(k_scale_entries, dim3(grid_for(nnz)), dim3(kBlock), 0, 0, A_,
                         static_cast<const double*>(fr.data()), static_cast<const double*>(fc.data()), nnz,
                         static_cast<const int32_t*>(row_of.data()));
      hipLaunchKernelGGL(k_scale_entries, dim3(grid_for(nnz)), dim3(kBlock), 0, 0, AT_,
                         static_cast<const double*>(fc.data()), static_cast<const double*>(fr.data()), nnz,
                         static_cast<const int32_t*>(col_of.data()));
    }
    GP_LAUNCH_CHECK("scaling pass");
  };
  for (int it = 0; it < p_.ruiz_iterations; ++it) pass(0);
  if (p_.pock_chambolle) pass(1);

  // norms of the original cost and row sides (relative tolerances), then of
  // the equilibrated ones (bound / cost scaling)
  Sums& s = sums_;
  hipLaunchKernelGGL(k_vec_norms, dim3(kRedBlocks), dim3(kBlock), 0, 0, m_, n_, lp_.c, lp_.AL, lp_.AU,
                     s.slot(0, kRedBlocks));
  s.finish(0);
  c_.resize(static_cast<size_t>(n_));
  l_.resize(static_cast<size_t>(n_));
  u_.resize(static_cast<size_t>(n_));
  AL_.resize(static_cast<size_t>(m_));
  AU_.resize(static_cast<size_t>(m_));
  hipLaunchKernelGGL(k_scale_cols, dim3(grid_for(n_)), dim3(kBlock), 0, 0, n_, lp_.c, lp_.l, lp_.u,
                     static_cast<const double*>(dc_.data()), 1.0, 1.0, c_.data(), l_.data(), u_.data());
  hipLaunchKernelGGL(k_scale_rows, dim3(grid_for(m_)), dim3(kBlock), 0, 0, m_, lp_.AL, lp_.AU,
                     static_cast<const double*>(dr_.data()), 1.0, AL_.data(), AU_.data());
  hipLaunchKernelGGL(k_vec_norms, dim3(kRedBlocks), dim3(kBlock), 0, 0, m_, n_,
                     static_cast<const double*>(c_.data()), static_cast<const double*>(AL_.data()),
                     static_cast<const double*>(AU_.data()), s.slot(1, kRedBlocks));
  s.finish(1);
  const double* h = s.read(2);
  norm_c_ = std::sqrt(h[0]);
  norm_b_ = std::sqrt(h[1]);
  const double cs = std::sqrt(h[kRedWidth + 0]), bs = std::sqrt(h[kRedWidth + 1]);
  if (p_.bound_cost_scaling) {
    b_scale_ = 1.0 + bs;
    c_scale_ = 1.0 + cs;
  }
  hipLaunchKernelGGL(k_scale_cols, dim3(grid_for(n_)), dim3(kBlock), 0, 0, n_, lp_.c, lp_.l, lp_.u,
                     static_cast<const double*>(dc_.data()), b_scale_, c_scale_, c_.data(), l_.data(), u_.data());
  hipLaunchKernelGGL(k_scale_rows, dim3(grid_for(m_)), dim3(kBlock), 0, 0, m_, lp_.AL, lp_.AU,
                     static_cast<const double*>(dr_.data()), b_scale_, AL_.data(), AU_.data());
  GP_LAUNCH_CHECK("scaled vectors");
  // sigma starts at ||b|| / ||c|| of the scaled problem
  const double bn = bs / b_scale_, cn = cs / c_scale_;
  sigma_ = (bn > 1e-10 && cn > 1e-10) ? bn / cn : 1.0;
}

double LpSolver::power_lambda() {
  if (lp_.A.nnz == 0) return 1.0;
  DVec<double> v(static_cast<size_t>(n_)), w(static_cast<size_t>(m_));
  hipLaunchKernelGGL(k_power_start, dim3(grid_for(n_)), dim3(kBlock), 0, 0, n_, v.data());
  Sums& s = sums_;
  // |v|^2 of the start vector (k_vec_norms with v in the cost slot)
  hipLaunchKernelGGL(k_vec_norms, dim3(kRedBlocks), dim3(kBlock), 0, 0, 0, n_, static_cast<const double*>(v.data()),
                     static_cast<const double*>(nullptr), static_cast<const double*>(nullptr),
                     s.slot(0, kRedBlocks));
  s.finish(0);
  const double vnorm2 = s.read(1)[0];
  double lambda = 0.0;
  double scale = 1.0 / std::sqrt(vnorm2);
  for (int it = 0; it < p_.power_iterations; ++it) {
    // w = A (v scale); v = A' w; lambda = |w|^2 (|v scale| = 1)
    launch_reduce<SpmvSqK>(lanes_a_, m_, A_, static_cast<const double*>(v.data()), scale, w.data(),
                           s.slot(0, red_grid(m_, lanes_a_)));
    launch_reduce<SpmvSqK>(lanes_at_, n_, AT_, static_cast<const double*>(w.data()), 1.0, v.data(),
                           s.slot(1, red_grid(n_, lanes_at_)));
    s.finish(0);
    s.finish(1);
    const double* h = s.read(2);
    const double next = h[0];
    const double vn2 = h[kRedWidth];
    if (!(vn2 > 0.0)) break;
    scale = 1.0 / std::sqrt(vn2);
    const bool done = it > 5 && std::fabs(next - lambda) <= p_.power_tolerance * next;
    lambda = next;
    if (done) break;
  }
  return lambda > 0.0 ? lambda * 1.01 : 1.0;
}

void LpSolver::check(double& r_fixed, double& dx_anchor, double& dy_anchor, gp_lp_solve_info* info) {
  Sums& s = sums_;
  // slots 0, 1 were filled by the check iteration's k_primal / k_dual
  launch_reduce<CheckRowsK>(lanes_a_, m_, A_, static_cast<const double*>(x_bar_.data()),
                            static_cast<const double*>(y_bar_.data()), static_cast<const double*>(y0_.data()),
                            static_cast<const double*>(AL_.data()), static_cast<const double*>(AU_.data()),
                            static_cast<const double*>(dr_.data()), b_scale_, s.slot(2, red_grid(m_, lanes_a_)));
  launch_reduce<CheckColsK>(lanes_at_, n_, AT_, static_cast<const double*>(y_bar_.data()),
                            static_cast<const double*>(x_bar_.data()), static_cast<const double*>(x0_.data()),
                            static_cast<const double*>(c_.data()), static_cast<const double*>(l_.data()),
                            static_cast<const double*>(u_.data()), static_cast<const double*>(dc_.data()), c_scale_,
                            s.slot(3, red_grid(n_, lanes_at_)));
  GP_LAUNCH_CHECK("check sums");
  for (int k = 0; k < 4; ++k) s.finish(k);
  const double* h = s.read(4);
  const double dx2 = h[0], dy2 = h[kRedWidth];
  const double viol2 = h[2 * kRedWidth + 0], dual_rows = h[2 * kRedWidth + 1], dya2 = h[2 * kRedWidth + 2];
  const double res2 = h[3 * kRedWidth + 0], dual_cols = h[3 * kRedWidth + 1], pobj = h[3 * kRedWidth + 2];
  const double dxa2 = h[3 * kRedWidth + 3];
  const double obj_scale = b_scale_ * c_scale_;
  info->primal_objective = obj_scale * pobj + lp_.obj_constant;
  info->dual_objective = obj_scale * (dual_rows + dual_cols) + lp_.obj_constant;
  info->primal_residual = std::sqrt(viol2) / (1.0 + norm_b_);
  info->dual_residual = std::sqrt(res2) / (1.0 + norm_c_);
  info->gap = std::fabs(info->primal_objective - info->dual_objective) /
              (1.0 + std::fabs(info->primal_objective) + std::fabs(info->dual_objective));
  // fixed-point residual in the iteration's metric
  r_fixed = std::sqrt(dx2 / sigma_ + sigma_ * lambda_ * dy2);
  dx_anchor = std::sqrt(dxa2);
  dy_anchor = std::sqrt(dya2);
}

void LpSolver::run(double* x_out, double* y_out, double* z_out, gp_lp_solve_info* info) {
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(info, 0, sizeof(*info));
  info->status = GP_LP_ITERATION_LIMIT;
  lanes_a_ = lanes_for_mean(m_ ? static_cast<double>(lp_.A.nnz) / m_ : 0.0);
  lanes_at_ = lanes_for_mean(n_ ? static_cast<double>(lp_.A.nnz) / n_ : 0.0);
  scale();
  lambda_ = power_lambda();
  bins_a_.build(A_);
  bins_at_.build(AT_);
  info->lambda = lambda_;
  x_.resize(static_cast<size_t>(n_));
  x0_.resize(static_cast<size_t>(n_));
  x_hat_.resize(static_cast<size_t>(n_));
  x_bar_.resize(static_cast<size_t>(n_));
  y_.resize(static_cast<size_t>(m_));
  y0_.resize(static_cast<size_t>(m_));
  y_bar_.resize(static_cast<size_t>(m_));
  // start at the projection of 0 onto the bounds, y = 0
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n_)), dim3(kBlock), 0, 0, static_cast<int64_t>(n_), 0.0, x_.data());
  hipLaunchKernelGGL(k_fill, dim3(grid_for(m_)), dim3(kBlock), 0, 0, static_cast<int64_t>(m_), 0.0, y_.data());
  d2d(x0_.data(), x_.data(), static_cast<size_t>(n_));
  d2d(y0_.data(), y_.data(), static_cast<size_t>(m_));
  GP_LAUNCH_CHECK("solver start");

  int64_t total = 0, inner = 0;
  double r0 = -1.0, r_last = INFINITY;
  const int every = p_.check_every > 0 ? p_.check_every : 64;
  bool done = false;
  while (!done) {
    for (int it = 0; it < every; ++it) {
      const bool last = it == every - 1;
      Step st;
      st.sigma = sigma_;
      st.inv_dual = sigma_ * lambda_;
      st.f_anchor = 1.0 / (static_cast<double>(inner) + 2.0);
      st.f_point = 1.0 - st.f_anchor;
      // check iterations reduce into at most kRedBlocks partials; the others
      // take the full grid
      const double* yc = y_.data();
      const double* xh = x_hat_.data();
      if (last) {
        launch_reduce<PrimalK>(lanes_at_, n_, AT_, yc, c_.data(), l_.data(), u_.data(), x_.data(), x0_.data(),
                               x_hat_.data(), x_bar_.data(), st, sums_.slot(0, red_grid(n_, lanes_at_)));
        launch_reduce<DualK>(lanes_a_, m_, A_, xh, AL_.data(), AU_.data(), y_.data(), y0_.data(), y_bar_.data(), st,
                             sums_.slot(1, red_grid(m_, lanes_a_)));
      } else {
        for (int b = 0; b < LineBins::kBins; ++b)
          if (bins_at_.count[b])
            launch_grouped<PrimalBinK>(bins_at_.lanes[b], bins_at_.count[b], AT_,
                                       static_cast<const int32_t*>(bins_at_.lines[b].data()), bins_at_.count[b], yc,
                                       static_cast<const double*>(c_.data()), static_cast<const double*>(l_.data()),
                                       static_cast<const double*>(u_.data()), x_.data(),
                                       static_cast<const double*>(x0_.data()), x_hat_.data(), st);
        for (int b = 0; b < LineBins::kBins; ++b)
          if (bins_a_.count[b])
            launch_grouped<DualBinK>(bins_a_.lanes[b], bins_a_.count[b], A_,
                                     static_cast<const int32_t*>(bins_a_.lines[b].data()), bins_a_.count[b], xh,
                                     static_cast<const double*>(AL_.data()), static_cast<const double*>(AU_.data()),
                                     y_.data(), static_cast<const double*>(y0_.data()), st);
      }
      GP_LAUNCH_CHECK("solver iteration");
      ++inner;
      ++total;
    }
    double r = 0.0, dxa = 0.0, dya = 0.0;
    check(r, dxa, dya, info);
    info->iterations = static_cast<int32_t>(total);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!std::isfinite(info->primal_objective) || !std::isfinite(info->dual_residual) || !std::isfinite(r)) {
      info->status = GP_LP_NUMERICAL_ERROR;
      break;
    }
    if (info->primal_residual <= p_.tolerance && info->dual_residual <= p_.tolerance && info->gap <= p_.tolerance) {
      info->status = GP_LP_OPTIMAL;
      break;
    }
    if (total >= p_.max_iterations) {
      info->status = GP_LP_ITERATION_LIMIT;
      break;
    }
    if (p_.time_limit > 0.0 && secs >= p_.time_limit) {
      info->status = GP_LP_TIME_LIMIT;
      break;
    }
    if (r0 < 0.0) r0 = r;
    const bool restart = r <= p_.restart_sufficient * r0 || (r <= p_.restart_necessary * r0 && r > r_last) ||
                         static_cast<double>(inner) >= p_.restart_artificial * static_cast<double>(total);
    r_last = r;
    if (restart) {
      // new anchor at (x_bar, y_bar); sigma set from the primal and dual
      // movement since the last restart (the step ratio that balances them)
      if (dxa > 1e-16 && dya > 1e-16 && dxa < 1e12 && dya < 1e12) sigma_ = dxa / (dya * std::sqrt(lambda_));
      if (p_.residual_balance != 0.0 && info->primal_residual > 0.0 && info->dual_residual > 0.0) {
        const double ratio = info->primal_residual / info->dual_residual;
        if (ratio > p_.residual_balance_band || ratio * p_.residual_balance_band < 1.0)
          sigma_ *= std::fmin(10.0, std::fmax(0.1, std::pow(ratio, p_.residual_balance)));
      }
      d2d(x_.data(), x_bar_.data(), static_cast<size_t>(n_));
      d2d(x0_.data(), x_bar_.data(), static_cast<size_t>(n_));
      d2d(y_.data(), y_bar_.data(), static_cast<size_t>(m_));
      d2d(y0_.data(), y_bar_.data(), static_cast<size_t>(m_));
      inner = 0;
      r0 = r;
      r_last = INFINITY;
      ++info->restarts;
    }
  }
  info->sigma = sigma_;
  // the solution of the last check, unscaled
  if (x_out || z_out)
    launch_grouped<UnscaleColsK>(lanes_at_, n_, AT_, static_cast<const double*>(y_bar_.data()),
                                 static_cast<const double*>(x_bar_.data()), static_cast<const double*>(c_.data()),
                                 static_cast<const double*>(l_.data()), static_cast<const double*>(u_.data()),
                                 static_cast<const double*>(dc_.data()), b_scale_, c_scale_, x_out, z_out);
  if (y_out && m_ > 0)
    hipLaunchKernelGGL(k_unscale_rows, dim3(grid_for(m_)), dim3(kBlock), 0, 0, m_,
                       static_cast<const double*>(y_bar_.data()), static_cast<const double*>(dr_.data()), c_scale_,
                       y_out);
  GP_LAUNCH_CHECK("unscale");
  GP_CHECK(hipDeviceSynchronize());
  info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // algorithmic bytes of one iteration: both matrices once (12 B per entry
  // and the row pointers), one gathered 8 B vector value per entry, and the
  // per-line vectors (x, x0, c, l, u read, x, x_hat written; y, y0, AL, AU
  // read, y written)
  info->bytes_per_iteration = 2.0 * (20.0 * static_cast<double>(lp_.A.nnz)) + 4.0 * (m_ + 1) + 4.0 * (n_ + 1) +
                              56.0 * n_ + 40.0 * m_;
}

// No matrix entries: every column on its own (the bound its cost points to,
// 0 clamped to the bounds at zero cost), every row's activity 0.
__global__ void k_separable(int32_t n, const double* c, const double* l, const double* u, double* x, double* z,
                            double* partial) {
  double acc[kRedWidth] = {};
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += stride) {
    const double cj = c[j];
    const double v = cj > 0.0 ? l[j] : cj < 0.0 ? u[j] : fmin(u[j], fmax(l[j], 0.0));
    if (!isfinite(v)) acc[1] += 1.0;
    else acc[0] += cj * v;
    if (x) x[j] = v;
    if (z) z[j] = cj;
  }
  block_partials(acc, partial);
}

__global__ void k_row_zero_violation(int32_t m, const double* AL, const double* AU, double* y, double* partial) {
  double acc[kRedWidth] = {};
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += stride) {
    const double v = AL[i] > 0.0 ? AL[i] : AU[i] < 0.0 ? -AU[i] : 0.0;  // distance of 0 from [AL, AU]
    acc[0] += v * v;
    if (y) y[i] = 0.0;
  }
  block_partials(acc, partial);
}

}  // namespace

void solve_separable(const gp_device_lp& lp, double* x, double* y, double* z, gp_lp_solve_info* info) {
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(info, 0, sizeof(*info));
  Sums s;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  hipLaunchKernelGGL(k_separable, dim3(kRedBlocks), dim3(kBlock), 0, 0, n, lp.c, lp.l, lp.u, x, z,
                     s.slot(0, kRedBlocks));
  hipLaunchKernelGGL(k_row_zero_violation, dim3(kRedBlocks), dim3(kBlock), 0, 0, m, lp.AL, lp.AU, y,
                     s.slot(1, kRedBlocks));
  GP_LAUNCH_CHECK("separable LP");
  s.finish(0);
  s.finish(1);
  const double* h = s.read(2);
  info->primal_objective = info->dual_objective = h[0] + lp.obj_constant;
  info->primal_residual = std::sqrt(h[kRedWidth]);
  info->status = h[1] > 0.0 ? GP_LP_UNBOUNDED : info->primal_residual > 0.0 ? GP_LP_INFEASIBLE : GP_LP_OPTIMAL;
  info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace gp

extern "C" void gp_lp_solve_default_params(gp_lp_solve_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->struct_size = static_cast<int32_t>(sizeof(*p));
  p->tolerance = 1e-6;
  p->max_iterations = 1000000;
  p->time_limit = 1000.0;
  p->check_every = 64;
  p->ruiz_iterations = 10;
  p->pock_chambolle = 1;
  p->bound_cost_scaling = 1;
  p->power_iterations = 5000;
  p->power_tolerance = 1e-4;
  p->restart_sufficient = 0.2;
  p->restart_necessary = 0.8;
  p->restart_artificial = 0.36;
  p->residual_balance = -1.0;
  p->residual_balance_band = 100.0;
}

extern "C" int gp_lp_solve(const gp_device_lp* lp, const gp_lp_solve_params* params, double* x, double* y, double* z,
                           gp_lp_solve_info* info) {
  return gp::guarded([&] {
    if (!lp || !params || !info) throw std::invalid_argument("gp_lp_solve: null argument");
    if (params->struct_size != static_cast<int32_t>(sizeof(gp_lp_solve_params)))
      throw std::invalid_argument("gp_lp_solve: params not initialised by gp_lp_solve_default_params");
    if (lp->A.rows < 0 || lp->A.cols < 0 || lp->A.rows != lp->AT.cols || lp->A.cols != lp->AT.rows ||
        lp->A.nnz != lp->AT.nnz)
      throw std::invalid_argument("gp_lp_solve: A and AT do not match");
    if (lp->A.rows == 0 || lp->A.cols == 0 || lp->A.nnz == 0) {
      gp::solve_separable(*lp, x, y, z, info);
      return;
    }
    gp::LpSolver solver(*lp, *params);
    solver.run(x, y, z, info);
  });
}
