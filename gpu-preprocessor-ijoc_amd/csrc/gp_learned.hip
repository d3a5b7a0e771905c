// gp_learned.hip -- the learned pass scheduler: per-round feature vector, the
// MLP 32-64-64-7 (FP32, ReLU, FP32 accumulation) with its inference, training
// (forward, masked BCE loss, backward) and Adam kernels, and the gate that the
// presolve scheduler consults before each round.
//
// The reference has no learned scheduler; the network decides which rule
// groups of the reference's phase machine (gpu_presolve.cu:2466-2592) run in a
// round.  Every dot product is summed in increasing input order, starting
// from 0 and adding the bias last, with no FMA contraction, so inference is
// bit-reproducible and equals a plain sequential CPU evaluation.
#include "gp_common.hpp"

namespace gp {

namespace {

constexpr int kIn = GP_MLP_IN, kH1 = GP_MLP_H1, kH2 = GP_MLP_H2, kOut = GP_MLP_OUT;
constexpr int kW1 = 0, kB1 = kW1 + kH1 * kIn, kW2 = kB1 + kH1, kB2 = kW2 + kH2 * kH1, kW3 = kB2 + kH2,
              kB3 = kW3 + kOut * kH2;
static_assert(kB3 + kOut == GP_MLP_PARAMS, "parameter layout");

__device__ inline float relu(float z) { return z > 0.0f ? z : 0.0f; }

// One dense layer, one thread per (sample, unit).  Z keeps the pre-activation
// (the backward pass needs its sign), H the activation (ReLU or identity).
template <bool Relu>
__global__ void k_dense_fwd(const float* __restrict__ W, const float* __restrict__ b, const float* __restrict__ X,
                            int32_t batch, int32_t in, int32_t out, float* Z, float* H) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(batch) * out) return;
  const int64_t s = idx / out;
  const int32_t o = static_cast<int32_t>(idx % out);
  const float* w = W + static_cast<int64_t>(o) * in;
  const float* x = X + s * in;
  float acc = 0.0f;
  for (int32_t i = 0; i < in; ++i) acc += w[i] * x[i];
  const float z = acc + b[o];
  Z[idx] = z;
  H[idx] = Relu ? relu(z) : z;
}

// Masked binary cross-entropy with logits.  dO = (sigmoid(o) - y) m / count;
// the per-element loss m (max(o,0) - o y + log1p(exp(-|o|))) goes to L.
__global__ void k_bce(const float* __restrict__ O, const float* __restrict__ Y, const float* __restrict__ M,
                      int64_t total, float inv_count, float* __restrict__ dO, float* __restrict__ L) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const float o = O[idx], y = Y[idx], m = M[idx];
  const float sig = 1.0f / (1.0f + expf(-o));
  dO[idx] = (sig - y) * m * inv_count;
  L[idx] = m * ((o > 0.0f ? o : 0.0f) - o * y + log1pf(expf(-fabsf(o))));
}

// Sum of L[0..total) in a fixed order: 256 lanes take strided partial sums,
// then a fixed pairwise tree; the result is scaled by inv_count.
__global__ void k_loss_sum(const float* __restrict__ L, int64_t total, float inv_count, float* __restrict__ loss) {
  __shared__ float part[256];
  float acc = 0.0f;
  for (int64_t i = threadIdx.x; i < total; i += 256) acc += L[i];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = part[0] * inv_count;
}

// Weight and bias gradients of one layer: gW[o][i] = sum_s dZ[s][o] X[s][i],
// gb[o] = sum_s dZ[s][o], samples in increasing order, one thread per value.
__global__ void k_dense_grad_w(const float* __restrict__ dZ, const float* __restrict__ X, int32_t batch,
                               int32_t in, int32_t out, float* __restrict__ gW, float* __restrict__ gb) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t nw = static_cast<int64_t>(out) * in;
  if (idx < nw) {
    const int32_t o = static_cast<int32_t>(idx / in), i = static_cast<int32_t>(idx % in);
    float acc = 0.0f;
    for (int32_t s = 0; s < batch; ++s) acc += dZ[static_cast<int64_t>(s) * out + o] * X[static_cast<int64_t>(s) * in + i];
    gW[idx] = acc;
  } else if (idx < nw + out) {
    const int32_t o = static_cast<int32_t>(idx - nw);
    float acc = 0.0f;
    for (int32_t s = 0; s < batch; ++s) acc += dZ[static_cast<int64_t>(s) * out + o];
    gb[o] = acc;
  }
}

// Gradient through one layer into the previous pre-activation:
// dZp[s][i] = (Zp[s][i] > 0) * sum_o dZ[s][o] W[o][i].
__global__ void k_dense_grad_x(const float* __restrict__ dZ, const float* __restrict__ W,
                               const float* __restrict__ Zp, int32_t batch, int32_t in, int32_t out,
                               float* __restrict__ dZp) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(batch) * in) return;
  const int64_t s = idx / in;
  const int32_t i = static_cast<int32_t>(idx % in);
  float acc = 0.0f;
  for (int32_t o = 0; o < out; ++o) acc += dZ[s * out + o] * W[static_cast<int64_t>(o) * in + i];
  dZp[idx] = Zp[idx] > 0.0f ? acc : 0.0f;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int32_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Single-sample inference for the scheduler: one 64-lane wavefront, hidden
// activations in LDS.  Same summation order as k_dense_fwd.
__global__ void __launch_bounds__(64) k_mlp_infer(const float* __restrict__ P, const float* __restrict__ x,
                                                  float* __restrict__ logits) {
  __shared__ float xs[kIn], h1[kH1], h2[kH2];
  const int t = threadIdx.x;
  if (t < kIn) xs[t] = x[t];
  __syncthreads();
  {
    const float* w = P + kW1 + t * kIn;
    float acc = 0.0f;
    for (int i = 0; i < kIn; ++i) acc += w[i] * xs[i];
    h1[t] = relu(acc + P[kB1 + t]);
  }
  __syncthreads();
  {
    const float* w = P + kW2 + t * kH1;
    float acc = 0.0f;
    for (int i = 0; i < kH1; ++i) acc += w[i] * h1[i];
    h2[t] = relu(acc + P[kB2 + t]);
  }
  __syncthreads();
  if (t < kOut) {
    const float* w = P + kW3 + t * kH2;
    float acc = 0.0f;
    for (int i = 0; i < kH2; ++i) acc += w[i] * h2[i];
    logits[t] = acc + P[kB3 + t];
  }
}
static_assert(kH1 == 64 && kH2 == 64 && kIn <= 64 && kOut <= 64, "k_mlp_infer assumes one 64-lane wave");

void check_batch(const void* a, const void* b, int32_t batch) {
  if (!a || !b) throw std::invalid_argument("null MLP buffer");
  if (batch < 1) throw std::invalid_argument("MLP batch must be >= 1");
}

// forward for a batch; activations in `act` (Z1 H1 Z2 H2 O, each [batch][units])
void mlp_forward(const float* P, const float* X, int32_t batch, float* act) {
  const int64_t B = batch;
  float* Z1 = act;
  float* H1 = Z1 + B * kH1;
  float* Z2 = H1 + B * kH1;
  float* H2 = Z2 + B * kH2;
  float* O = H2 + B * kH2;
  hipLaunchKernelGGL(k_dense_fwd<true>, dim3(grid_for(B * kH1)), dim3(kBlock), 0, 0, P + kW1, P + kB1, X, batch, kIn,
                     kH1, Z1, H1);
  hipLaunchKernelGGL(k_dense_fwd<true>, dim3(grid_for(B * kH2)), dim3(kBlock), 0, 0, P + kW2, P + kB2, H1, batch,
                     kH1, kH2, Z2, H2);
  hipLaunchKernelGGL(k_dense_fwd<false>, dim3(grid_for(B * kOut)), dim3(kBlock), 0, 0, P + kW3, P + kB3, H2, batch,
                     kH2, kOut, O, O);
  GP_LAUNCH_CHECK("mlp forward");
}

constexpr int64_t act_floats(int64_t B) { return B * (2 * kH1 + 2 * kH2 + kOut); }

}  // namespace

int rule_group(int32_t rule) {
  switch (rule) {
    case GP_RULE_FIXED_VARIABLES:
    case GP_RULE_EMPTY_COLUMNS:
    case GP_RULE_EMPTY_ROWS: return GP_GROUP_FIXED_EMPTY;
    case GP_RULE_SINGLETON_ROWS: return GP_GROUP_SINGLETON_ROWS;
    case GP_RULE_SINGLETON_COLUMNS: return GP_GROUP_SINGLETON_COLUMNS;
    case GP_RULE_BOUND_STRENGTHENING:
    case GP_RULE_REDUNDANT_ROWS: return GP_GROUP_BOUNDS_REDUNDANT;
    case GP_RULE_DUAL_FIX: return GP_GROUP_DUAL_FIX;
    case GP_RULE_DOUBLETON_EQUATIONS: return GP_GROUP_DOUBLETON;
    case GP_RULE_PARALLEL_ROWS:
    case GP_RULE_PARALLEL_COLUMNS: return GP_GROUP_PARALLEL;
    default: return -1;
  }
}

namespace {
inline float ratio(double a, double b) { return static_cast<float>(a / (b > 1.0 ? b : 1.0)); }
inline float clamp01(double v) { return static_cast<float>(v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v)); }
}  // namespace

void schedule_features(const gp_round_trace* hist, int32_t done, int32_t phase, int64_t nnz_now, int64_t nnz0,
                       int32_t m0, int32_t n0, int32_t bound_only, int32_t bound_limit, float* f) {
  for (int k = 0; k < kIn; ++k) f[k] = 0.0f;
  if (phase >= 0 && phase < 4) f[phase] = 1.0f;
  const gp_round_trace* p1 = done >= 1 ? &hist[done - 1] : nullptr;
  const gp_round_trace* p2 = done >= 2 ? &hist[done - 2] : nullptr;
  for (int r = 0; r < GP_RULE_COUNT; ++r) {
    if (p1) f[4 + r] = static_cast<float>((p1->fired_mask >> r) & 1);
    if (p2) f[15 + r] = static_cast<float>((p2->fired_mask >> r) & 1);
  }
  f[26] = ratio(static_cast<double>(nnz_now), static_cast<double>(nnz0));
  f[27] = ratio(static_cast<double>(p1 ? p1->rows_live : m0), static_cast<double>(m0));
  f[28] = ratio(static_cast<double>(p1 ? p1->cols_live : n0), static_cast<double>(n0));
  if (p1) {
    const double before = static_cast<double>(p1->nnz_before);
    f[29] = clamp01(10.0 * static_cast<double>(p1->nnz_before - p1->nnz_after) / (before > 1.0 ? before : 1.0));
  }
  f[30] = clamp01(static_cast<double>(bound_only) / static_cast<double>(bound_limit > 1 ? bound_limit : 1));
  f[31] = clamp01(static_cast<double>(done) / 64.0);
}

// ---- the scheduler's gate ---------------------------------------------------------------
void LearnedGate::init(const float* blob) {
  if (!blob) throw std::invalid_argument("learned_schedule=1 needs scheduler_weights");
  params.resize(GP_MLP_PARAMS);
  io.resize(kIn + kOut);
  h2d(params.data(), blob, GP_MLP_PARAMS);
  for (int g = 0; g < kOut; ++g) threshold[g] = blob[GP_MLP_PARAMS + g];
  ready = true;
}

uint32_t LearnedGate::decide(const float* features, float* logits_out) {
  float logits[kOut];
  h2d(io.data(), features, kIn);
  hipLaunchKernelGGL(k_mlp_infer, dim3(1), dim3(64), 0, 0, params.data(), io.data(), io.data() + kIn);
  GP_LAUNCH_CHECK("mlp infer");
  d2h(logits, io.data() + kIn, kOut);
  uint32_t run = 0;
  for (int g = 0; g < kOut; ++g) {
    if (logits[g] >= threshold[g]) run |= 1u << g;
    if (logits_out) logits_out[g] = logits[g];
  }
  return run;
}

}  // namespace gp

using namespace gp;

extern "C" {

int gp_rule_group(int32_t rule) { return rule_group(rule); }

void gp_schedule_features(const gp_round_trace* history, int32_t rounds_done, int32_t phase, int64_t nnz_now,
                          int64_t nnz0, int32_t m0, int32_t n0, int32_t bound_only_rounds, int32_t bound_only_limit,
                          float* features) {
  if (!features || (rounds_done > 0 && !history)) return;
  schedule_features(history, rounds_done, phase, nnz_now, nnz0, m0, n0, bound_only_rounds, bound_only_limit,
                    features);
}

int gp_mlp_forward(const float* params, const float* x, int32_t batch, float* logits) {
  return guarded([&] {
    check_batch(params, x, batch);
    if (!logits) throw std::invalid_argument("null logits");
    DVec<float> act(static_cast<size_t>(act_floats(batch)));
    mlp_forward(params, x, batch, act.data());
    const int64_t B = batch;
    d2d(logits, act.data() + B * (2 * kH1 + 2 * kH2), static_cast<size_t>(B * kOut));
    GP_CHECK(hipDeviceSynchronize());
  });
}

int gp_mlp_loss_grad(const float* params, const float* x, const float* y, const float* mask, int32_t batch,
                     float* loss, float* grad) {
  return guarded([&] {
    check_batch(params, x, batch);
    if (!y || !mask || !loss || !grad) throw std::invalid_argument("null MLP buffer");
    const int64_t B = batch;
    DVec<float> act(static_cast<size_t>(act_floats(B)));
    mlp_forward(params, x, batch, act.data());
    float* Z1 = act.data();
    float* H1 = Z1 + B * kH1;
    float* Z2 = H1 + B * kH1;
    float* H2 = Z2 + B * kH2;
    float* O = H2 + B * kH2;
    // the mean runs over the mask's ones: count them on the host (masks are 0/1)
    std::vector<float> hm(static_cast<size_t>(B * kOut));
    d2h(hm.data(), mask, hm.size());
    double count = 0.0;
    for (float v : hm) count += v;
    const float inv = static_cast<float>(1.0 / (count > 0.0 ? count : 1.0));
    DVec<float> dO(static_cast<size_t>(B * kOut)), L(static_cast<size_t>(B * kOut));
    DVec<float> dZ2(static_cast<size_t>(B * kH2)), dZ1(static_cast<size_t>(B * kH1));
    hipLaunchKernelGGL(k_bce, dim3(grid_for(B * kOut)), dim3(kBlock), 0, 0, O, y, mask, B * kOut, inv, dO.data(),
                       L.data());
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, 0, L.data(), B * kOut, inv, loss);
    // layer 3
    hipLaunchKernelGGL(k_dense_grad_w, dim3(grid_for(kOut * kH2 + kOut)), dim3(kBlock), 0, 0, dO.data(), H2, batch,
                       kH2, kOut, grad + kW3, grad + kB3);
    hipLaunchKernelGGL(k_dense_grad_x, dim3(grid_for(B * kH2)), dim3(kBlock), 0, 0, dO.data(), params + kW3, Z2,
                       batch, kH2, kOut, dZ2.data());
    // layer 2
    hipLaunchKernelGGL(k_dense_grad_w, dim3(grid_for(kH2 * kH1 + kH2)), dim3(kBlock), 0, 0, dZ2.data(), H1, batch,
                       kH1, kH2, grad + kW2, grad + kB2);
    hipLaunchKernelGGL(k_dense_grad_x, dim3(grid_for(B * kH1)), dim3(kBlock), 0, 0, dZ2.data(), params + kW2, Z1,
                       batch, kH1, kH2, dZ1.data());
    // layer 1
    hipLaunchKernelGGL(k_dense_grad_w, dim3(grid_for(kH1 * kIn + kH1)), dim3(kBlock), 0, 0, dZ1.data(), x, batch,
                       kIn, kH1, grad + kW1, grad + kB1);
    GP_LAUNCH_CHECK("mlp backward");
    GP_CHECK(hipDeviceSynchronize());
  });
}

int gp_mlp_adam_step(float* params, const float* grad, float* m, float* v, int32_t t, float lr, float beta1,
                     float beta2, float eps) {
  return guarded([&] {
    if (!params || !grad || !m || !v) throw std::invalid_argument("null Adam buffer");
    if (t < 1) throw std::invalid_argument("Adam step must be >= 1");
    const float c1 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta1), t));
    const float c2 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta2), t));
    hipLaunchKernelGGL(k_adam, dim3(grid_for(GP_MLP_PARAMS)), dim3(kBlock), 0, 0, params, grad, m, v, GP_MLP_PARAMS,
                       lr, beta1, beta2, eps, c1, c2);
    GP_LAUNCH_CHECK("adam");
    GP_CHECK(hipDeviceSynchronize());
  });
}

}  // extern "C"
