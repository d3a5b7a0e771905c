// presolve_example.cpp -- the reference's toy drop-in example
// (examples/presolve_example.cpp) against this engine's C++ front end:
// min x0 + x1, x0 >= 1, x1 >= 1, 0 <= x0, x1 <= 10.  Device memory is HIP;
// the calls and types are the reference's gpu_preprocessor names.
#include <hip/hip_runtime.h>

#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpu_preprocessor/gpu_preprocessor.hpp"

namespace gpp = gpu_preprocessor;

static void check(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

template <class T>
struct DeviceVector {
  T* data = nullptr;
  explicit DeviceVector(const std::vector<T>& v) {
    if (v.empty()) return;
    check(hipMalloc(&data, v.size() * sizeof(T)));
    check(hipMemcpy(data, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  ~DeviceVector() {
    if (data) (void)hipFree(data);
  }
  DeviceVector(const DeviceVector&) = delete;
  DeviceVector& operator=(const DeviceVector&) = delete;
};

struct PipelineOwner {
  gpp::Summary value;
  ~PipelineOwner() { gpp::free_summary(value); }
};

int main() {
  try {
    gpp::Parameters params;
    check(hipSetDevice(0));
    DeviceVector<gpp::NnzOffset> rp({0, 1, 2});
    DeviceVector<std::int32_t> ci({0, 1});
    DeviceVector<double> av({1, 1}), c({1, 1}), al({1, 1}), au({INFINITY, INFINITY}), l({0, 0}), u({10, 10});
    gpp::DeviceLp original;
    original.A = {2, 2, 2, rp.data, ci.data, av.data};
    original.AT = original.A;  // the identity is its own transpose
    original.c = c.data;
    original.AL = al.data;
    original.AU = au.data;
    original.l = l.data;
    original.u = u.data;
    PipelineOwner pipeline;
    pipeline.value = gpp::preprocess(original, params);
    auto& p = pipeline.value.presolve;
    if (p.has_infeasible || p.has_unbounded) throw std::runtime_error("unexpected infeasible/unbounded example");
    std::cout << "original 2 x 2 -> reduced " << p.reduced_rows << " x " << p.reduced_cols << "\n";
    if (p.reduced_rows != p.reduced_cols || p.reduced_cols > 2)
      throw std::runtime_error("unexpected toy reduction");
    DeviceVector<double> xr(std::vector<double>(p.reduced_cols, 1.0));
    DeviceVector<double> yr(std::vector<double>(p.reduced_rows, p.reduced_cols == 1 ? 2.0 : 1.0));
    DeviceVector<double> zr(std::vector<double>(p.reduced_cols, 0.0));
    const auto sol = gpp::postprocess(xr.data, yr.data, zr.data, pipeline.value, original);
    if (sol.x.size() != 2 || std::abs(sol.x[0] - 1) > 1e-8 || std::abs(sol.x[1] - 1) > 1e-8)
      throw std::runtime_error("original primal recovery failed");
    std::cout << "original primal: " << sol.x[0] << ", " << sol.x[1] << "; objective: " << sol.x[0] + sol.x[1]
              << "\n";
    gpp::free_summary(pipeline.value);
    gpp::free_summary(pipeline.value);  // idempotent
    return 0;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
}
