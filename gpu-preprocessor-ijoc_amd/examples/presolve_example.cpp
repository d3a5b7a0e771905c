// presolve_example.cpp -- a caller of the drop-in C++ API
// (include/gpu_preprocessor/gpu_preprocessor.hpp: preprocess / postprocess /
// free_summary) on a small product-mix LP that presolve solves completely:
//
//   min  x0 - x1 + 0.5 x2 + x3
//   s.t. x0 + x1 + x2 = 6        (capacity, equality)
//             x1      <= 3       (market limit of product 1)
//                 2 x3 = 3       (fixed blend)
//        0 <= x0 <= 5, 0 <= x1 <= 10, x2 = 1, x3 free
//
// Optimum: x = (2, 3, 1, 1.5), objective 1.  The presolve removes every row
// and column, so the reduced solution is empty and postprocess() rebuilds the
// whole primal-dual point from the recovery tape.  The example checks the
// Karush-Kuhn-Tucker conditions of that point (convention c - A^T y - z = 0)
// and the objective, then shows that a run that changes nothing borrows the
// input LP.  Exit status 0 means every check passed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpu_preprocessor/gpu_preprocessor.hpp"

namespace gpp = gpu_preprocessor;

namespace {

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device copy of a host vector, released with the owner.
template <class T>
class DeviceBuffer {
 public:
  explicit DeviceBuffer(const std::vector<T>& host) : size_(host.size()) {
    if (host.empty()) return;
    hip_ok(hipMalloc(&ptr_, sizeof(T) * host.size()), "hipMalloc");
    hip_ok(hipMemcpy(ptr_, host.data(), sizeof(T) * host.size(), hipMemcpyHostToDevice), "hipMemcpy");
  }
  ~DeviceBuffer() {
    if (ptr_) (void)hipFree(ptr_);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  T* get() const { return ptr_; }

 private:
  T* ptr_ = nullptr;
  size_t size_ = 0;
};

// Row-major dense matrix -> CSR of the matrix (rows x cols) with the
// explicit nonzeros, in column order within each row.
struct HostCsr {
  std::vector<gpp::NnzOffset> ptr{0};
  std::vector<std::int32_t> idx;
  std::vector<double> val;
};

HostCsr to_csr(const std::vector<std::vector<double>>& dense, bool transpose) {
  const size_t rows = transpose ? dense[0].size() : dense.size();
  const size_t cols = transpose ? dense.size() : dense[0].size();
  HostCsr out;
  for (size_t i = 0; i < rows; ++i) {
    for (size_t j = 0; j < cols; ++j) {
      const double a = transpose ? dense[j][i] : dense[i][j];
      if (a == 0.0) continue;
      out.idx.push_back(static_cast<std::int32_t>(j));
      out.val.push_back(a);
    }
    out.ptr.push_back(static_cast<gpp::NnzOffset>(out.idx.size()));
  }
  return out;
}

void require(bool ok, const std::string& what) {
  if (!ok) throw std::runtime_error("check failed: " + what);
}

}  // namespace

int main() {
  try {
    hip_ok(hipSetDevice(0), "hipSetDevice");
    const double inf = INFINITY;
    const std::vector<std::vector<double>> A = {{1, 1, 1, 0}, {0, 1, 0, 0}, {0, 0, 0, 2}};
    const std::vector<double> c = {1, -1, 0.5, 1}, l = {0, 0, 1, -inf}, u = {5, 10, 1, inf};
    const std::vector<double> AL = {6, -inf, 3}, AU = {6, 3, 3};
    const int m = 3, n = 4;

    const HostCsr a = to_csr(A, false), at = to_csr(A, true);
    DeviceBuffer<gpp::NnzOffset> a_ptr(a.ptr), at_ptr(at.ptr);
    DeviceBuffer<std::int32_t> a_idx(a.idx), at_idx(at.idx);
    DeviceBuffer<double> a_val(a.val), at_val(at.val), d_c(c), d_l(l), d_u(u), d_al(AL), d_au(AU);
    gpp::DeviceLp lp;
    lp.A = {m, n, static_cast<gpp::NnzOffset>(a.idx.size()), a_ptr.get(), a_idx.get(), a_val.get()};
    lp.AT = {n, m, static_cast<gpp::NnzOffset>(at.idx.size()), at_ptr.get(), at_idx.get(), at_val.get()};
    lp.c = d_c.get();
    lp.l = d_l.get();
    lp.u = d_u.get();
    lp.AL = d_al.get();
    lp.AU = d_au.get();

    // Reference defaults, folding and Startup rules off, and the dual data of
    // the cleanup rules on the tape so that postprocess recovers y and z too.
    gpp::Parameters params;
    params.enable_folding = false;
    params.enable_structure_specific_rules = false;
    params.record_postsolve_tape_cpu = true;

    gpp::Summary summary = gpp::preprocess(lp, params);
    const auto& ps = summary.presolve;
    std::printf("presolve: %d x %d -> %d x %d (nnz %d) in %d iterations, %.3f ms (core %.3f ms)\n", ps.original_rows,
                ps.original_cols, ps.reduced_rows, ps.reduced_cols, static_cast<int>(ps.reduced_nnz), ps.iterations,
                1e3 * summary.preprocessing_seconds, 1e3 * ps.elapsed_seconds);
    require(!ps.has_infeasible && !ps.has_unbounded, "the LP is feasible and bounded");
    require(ps.reduced_rows == 0 && ps.reduced_cols == 0, "presolve removes every row and column");

    // The reduced LP is empty: its solution is three empty device vectors.
    const gpp::Solution sol = gpp::postprocess(nullptr, nullptr, nullptr, summary, lp);
    require(sol.x.size() == 4 && sol.y.size() == 3 && sol.z.size() == 4, "original dimensions");

    const double tol = 1e-9;
    double obj = 0.0;
    for (int j = 0; j < n; ++j) obj += c[j] * sol.x[j];
    for (int j = 0; j < n; ++j) {
      // bounds, reduced-cost sign at an active bound, zero reduced cost when inactive
      require(sol.x[j] >= l[j] - tol && sol.x[j] <= u[j] + tol, "bound of x" + std::to_string(j));
      const bool at_l = std::isfinite(l[j]) && sol.x[j] <= l[j] + tol;
      const bool at_u = std::isfinite(u[j]) && sol.x[j] >= u[j] - tol;
      if (!at_l) require(sol.z[j] <= tol, "z" + std::to_string(j) + " <= 0 off the lower bound");
      if (!at_u) require(sol.z[j] >= -tol, "z" + std::to_string(j) + " >= 0 off the upper bound");
      double aty = 0.0;
      for (int i = 0; i < m; ++i) aty += A[i][j] * sol.y[i];
      require(std::fabs(c[j] - aty - sol.z[j]) <= tol, "stationarity of column " + std::to_string(j));
    }
    for (int i = 0; i < m; ++i) {
      double act = 0.0;
      for (int j = 0; j < n; ++j) act += A[i][j] * sol.x[j];
      require(act >= AL[i] - tol && act <= AU[i] + tol, "row " + std::to_string(i) + " feasibility");
      const bool at_l = std::isfinite(AL[i]) && act <= AL[i] + tol;
      const bool at_u = std::isfinite(AU[i]) && act >= AU[i] - tol;
      if (!at_l) require(sol.y[i] <= tol, "y" + std::to_string(i) + " <= 0 off the lower side");
      if (!at_u) require(sol.y[i] >= -tol, "y" + std::to_string(i) + " >= 0 off the upper side");
    }
    require(std::fabs(obj - 1.0) <= tol, "optimal objective 1");
    std::printf("x = (%g, %g, %g, %g), y = (%g, %g, %g), objective %g: KKT conditions hold\n", sol.x[0], sol.x[1],
                sol.x[2], sol.x[3], sol.y[0], sol.y[1], sol.y[2], obj);
    gpp::free_summary(summary);
    gpp::free_summary(summary);  // idempotent

    // A run that changes nothing borrows the caller's LP instead of copying it.
    params.max_iters = 0;
    gpp::Summary identity = gpp::preprocess(lp, params);
    require(!identity.presolve.owns_reduced_lp && identity.presolve.reduced_lp.A.rowPtr == lp.A.rowPtr,
            "an unchanged LP is borrowed");
    gpp::free_summary(identity);
    std::printf("unchanged run borrows the input LP; summaries released\n");
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "presolve_example: %s\n", e.what());
    return 1;
  }
}
