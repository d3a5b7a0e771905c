/*
 * oracle_mlp.c -- CPU restatement of the learned pass scheduler: the per-round
 * feature vector, the MLP 32-64-64-7 (FP32, ReLU) forward pass, the masked
 * binary cross-entropy loss with its gradient, and the Adam step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker of gp_learned.hip).  The reference
 * has no learned scheduler; what is restated here is the definition of the
 * network the engine trains and consults: every dot product is a sequential
 * FP32 sum in increasing input order starting from 0, with the bias added
 * last, so the forward pass matches the HIP kernels bit for bit (the file is
 * built with -ffp-contract=off).  The loss uses expf/log1pf, whose last bits
 * may differ between libm and the device library.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

enum { IN = GP_MLP_IN, H1 = GP_MLP_H1, H2 = GP_MLP_H2, OUT = GP_MLP_OUT };
enum { OW1 = 0, OB1 = OW1 + H1 * IN, OW2 = OB1 + H1, OB2 = OW2 + H2 * H1, OW3 = OB2 + H2, OB3 = OW3 + OUT * H2 };

int or_rule_group(int32_t rule) {
  static const int group[GP_RULE_COUNT] = {0, 0, 0, 1, 2, 3, 3, 4, 5, 6, 6};
  return rule >= 0 && rule < GP_RULE_COUNT ? group[rule] : -1;
}

static float ratio(double a, double b) { return (float)(a / (b > 1.0 ? b : 1.0)); }
static float clamp01(double v) { return (float)(v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v)); }

void or_schedule_features(const gp_round_trace* hist, int32_t done, int32_t phase, int64_t nnz_now,
                          int64_t nnz0, int32_t m0, int32_t n0, int32_t bound_only, int32_t limit,
                          float* f) {
  memset(f, 0, sizeof(float) * IN);
  if (phase >= 0 && phase < 4) f[phase] = 1.0f;
  const gp_round_trace* p1 = done >= 1 ? &hist[done - 1] : NULL;
  const gp_round_trace* p2 = done >= 2 ? &hist[done - 2] : NULL;
  for (int r = 0; r < GP_RULE_COUNT; ++r) {
    if (p1) f[4 + r] = (float)((p1->fired_mask >> r) & 1);
    if (p2) f[15 + r] = (float)((p2->fired_mask >> r) & 1);
  }
  f[26] = ratio((double)nnz_now, (double)nnz0);
  f[27] = ratio((double)(p1 ? p1->rows_live : m0), (double)m0);
  f[28] = ratio((double)(p1 ? p1->cols_live : n0), (double)n0);
  if (p1) {
    const double before = (double)p1->nnz_before;
    f[29] = clamp01(10.0 * (double)(p1->nnz_before - p1->nnz_after) / (before > 1.0 ? before : 1.0));
  }
  f[30] = clamp01((double)bound_only / (double)(limit > 1 ? limit : 1));
  f[31] = clamp01((double)done / 64.0);
}

static float relu(float z) { return z > 0.0f ? z : 0.0f; }

static void dense(const float* W, const float* b, const float* x, int in, int out, int act, float* z, float* h) {
  for (int o = 0; o < out; ++o) {
    float acc = 0.0f;
    for (int i = 0; i < in; ++i) acc += W[o * in + i] * x[i];
    const float v = acc + b[o];
    if (z) z[o] = v;
    h[o] = act ? relu(v) : v;
  }
}

void or_mlp_forward(const float* P, const float* x, int32_t batch, float* logits) {
  float h1[H1], h2[H2];
  for (int32_t s = 0; s < batch; ++s) {
    dense(P + OW1, P + OB1, x + (int64_t)s * IN, IN, H1, 1, NULL, h1);
    dense(P + OW2, P + OB2, h1, H1, H2, 1, NULL, h2);
    dense(P + OW3, P + OB3, h2, H2, OUT, 0, NULL, logits + (int64_t)s * OUT);
  }
}

/* loss = sum over masked entries of BCE-with-logits / count; grad by the
 * chain rule with the same per-value summation orders as the HIP kernels
 * (weights: over samples in increasing order; inputs: over units). */
void or_mlp_loss_grad(const float* P, const float* x, const float* y, const float* mask, int32_t batch,
                      double* loss, float* grad) {
  const int64_t B = batch;
  float* z1 = (float*)malloc(sizeof(float) * (size_t)(B * H1));
  float* a1 = (float*)malloc(sizeof(float) * (size_t)(B * H1));
  float* z2 = (float*)malloc(sizeof(float) * (size_t)(B * H2));
  float* a2 = (float*)malloc(sizeof(float) * (size_t)(B * H2));
  float* o = (float*)malloc(sizeof(float) * (size_t)(B * OUT));
  float* d3 = (float*)malloc(sizeof(float) * (size_t)(B * OUT));
  float* d2 = (float*)malloc(sizeof(float) * (size_t)(B * H2));
  float* d1 = (float*)malloc(sizeof(float) * (size_t)(B * H1));
  for (int64_t s = 0; s < B; ++s) {
    dense(P + OW1, P + OB1, x + s * IN, IN, H1, 1, z1 + s * H1, a1 + s * H1);
    dense(P + OW2, P + OB2, a1 + s * H1, H1, H2, 1, z2 + s * H2, a2 + s * H2);
    dense(P + OW3, P + OB3, a2 + s * H2, H2, OUT, 0, NULL, o + s * OUT);
  }
  double count = 0.0;
  for (int64_t k = 0; k < B * OUT; ++k) count += mask[k];
  const float inv = (float)(1.0 / (count > 0.0 ? count : 1.0));
  double total = 0.0;
  for (int64_t k = 0; k < B * OUT; ++k) {
    const float ok = o[k], yk = y[k], mk = mask[k];
    const float sig = 1.0f / (1.0f + expf(-ok));
    d3[k] = (sig - yk) * mk * inv;
    total += (double)(mk * ((ok > 0.0f ? ok : 0.0f) - ok * yk + log1pf(expf(-fabsf(ok)))));
  }
  *loss = total / (count > 0.0 ? count : 1.0);
  /* layer by layer: weight grads, then the gradient into the layer below */
  const struct {
    int in, out, ow, ob;
    const float* xin;
    const float* dz;
  } layers[3] = {{H2, OUT, OW3, OB3, a2, d3}, {H1, H2, OW2, OB2, a1, d2}, {IN, H1, OW1, OB1, x, d1}};
  for (int L = 0; L < 3; ++L) {
    const int in = layers[L].in, out = layers[L].out;
    const float* dz = layers[L].dz;
    const float* xin = layers[L].xin;
    for (int oi = 0; oi < out; ++oi) {
      for (int i = 0; i < in; ++i) {
        float acc = 0.0f;
        for (int64_t s = 0; s < B; ++s) acc += dz[s * out + oi] * xin[s * in + i];
        grad[layers[L].ow + oi * in + i] = acc;
      }
      float acc = 0.0f;
      for (int64_t s = 0; s < B; ++s) acc += dz[s * out + oi];
      grad[layers[L].ob + oi] = acc;
    }
    if (L == 2) break;
    float* dprev = L == 0 ? d2 : d1;
    const float* zprev = L == 0 ? z2 : z1;
    for (int64_t s = 0; s < B; ++s)
      for (int i = 0; i < in; ++i) {
        float acc = 0.0f;
        for (int oi = 0; oi < out; ++oi) acc += dz[s * out + oi] * P[layers[L].ow + oi * in + i];
        dprev[s * in + i] = zprev[s * in + i] > 0.0f ? acc : 0.0f;
      }
  }
  free(z1); free(a1); free(z2); free(a2); free(o); free(d3); free(d2); free(d1);
}

void or_mlp_adam_step(float* w, const float* g, float* m, float* v, int32_t t, float lr, float b1, float b2,
                      float eps) {
  const float c1 = (float)(1.0 - pow((double)b1, t)), c2 = (float)(1.0 - pow((double)b2, t));
  for (int i = 0; i < GP_MLP_PARAMS; ++i) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}
