// ORACLE -- test infrastructure only.  CPU restatement of the learned seed's
// network (6-64-64-2, ReLU): BF16 rounding at the points the device rounds
// (inputs, weights, hidden activations after bias and ReLU, output deltas and
// hidden deltas), products of BF16 values summed in double then rounded to FP32
// (the device sums them in FP32 on the MFMA units, in its own order).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

enum { MLP_IN = 6, MLP_H = 64, MLP_OUT = 2, MLP_K1 = 16, MLP_N3 = 16 };
enum { OFF_W1 = 0, OFF_B1 = 1024, OFF_W2 = 1088, OFF_B2 = 5184, OFF_W3 = 5248, OFF_B3 = 6272, MLP_PARAMS = 6288 };

// FP32 -> BF16 -> FP32, round to nearest even
static inline float bfr(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return f;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

static inline void mlp_features(double Tg, double p, double yinf, double Tp, double dp, double Re, float* x) {
  x[0] = (float)(Tg / 1000.0 - 1.0);
  x[1] = (float)std::log(p / 1.0e5);
  x[2] = (float)(3.0 * yinf);
  x[3] = (float)((Tp - 315.0) / 35.0);
  x[4] = (float)std::log(dp / 4.0e-5);
  x[5] = (float)std::log(1.0 + Re);
}

struct MlpCache {
  float x[MLP_K1], h1[MLP_H], h2[MLP_H], out[MLP_OUT];
};

// forward pass of one sample (P: FP32 master parameters)
static inline void mlp_forward(const float* P, const float* xin, MlpCache& c) {
  for (int k = 0; k < MLP_K1; ++k) c.x[k] = k < MLP_IN ? bfr(xin[k]) : 0.0f;
  for (int n = 0; n < MLP_H; ++n) {
    double z = 0.0;
    for (int k = 0; k < MLP_K1; ++k) z += (double)c.x[k] * (double)bfr(P[OFF_W1 + n * 16 + k]);
    float zf = (float)z + P[OFF_B1 + n];
    c.h1[n] = bfr(zf > 0.0f ? zf : 0.0f);
  }
  for (int n = 0; n < MLP_H; ++n) {
    double z = 0.0;
    for (int k = 0; k < MLP_H; ++k) z += (double)c.h1[k] * (double)bfr(P[OFF_W2 + n * 64 + k]);
    float zf = (float)z + P[OFF_B2 + n];
    c.h2[n] = bfr(zf > 0.0f ? zf : 0.0f);
  }
  for (int n = 0; n < MLP_OUT; ++n) {
    double z = 0.0;
    for (int k = 0; k < MLP_H; ++k) z += (double)c.h2[k] * (double)bfr(P[OFF_W3 + n * 64 + k]);
    c.out[n] = (float)z + P[OFF_B3 + n];
  }
}

// mean squared error over n samples and 2 outputs, and its gradient (double sums)
static inline float mlp_grad(const float* P, int n, const float* X, const float* Y, float* grad) {
  std::vector<double> g(MLP_PARAMS, 0.0);
  double loss = 0.0;
  const float scale = 1.0f / (float)(2 * n);
  MlpCache c;
  for (int s = 0; s < n; ++s) {
    mlp_forward(P, X + (size_t)s * MLP_IN, c);
    float dO[MLP_OUT];
    for (int o = 0; o < MLP_OUT; ++o) {
      float e = c.out[o] - Y[(size_t)s * MLP_OUT + o];
      loss += (double)e * (double)e;
      dO[o] = bfr(2.0f * e * scale);
      g[OFF_B3 + o] += dO[o];
      for (int k = 0; k < MLP_H; ++k) g[OFF_W3 + o * 64 + k] += (double)dO[o] * (double)c.h2[k];
    }
    float dZ2[MLP_H];
    for (int k = 0; k < MLP_H; ++k) {
      double d = 0.0;
      for (int o = 0; o < MLP_OUT; ++o) d += (double)dO[o] * (double)bfr(P[OFF_W3 + o * 64 + k]);
      dZ2[k] = c.h2[k] > 0.0f ? bfr((float)d) : 0.0f;
      g[OFF_B2 + k] += dZ2[k];
      for (int m = 0; m < MLP_H; ++m) g[OFF_W2 + k * 64 + m] += (double)dZ2[k] * (double)c.h1[m];
    }
    for (int k = 0; k < MLP_H; ++k) {
      double d = 0.0;
      for (int m = 0; m < MLP_H; ++m) d += (double)dZ2[m] * (double)bfr(P[OFF_W2 + m * 64 + k]);
      float dZ1 = c.h1[k] > 0.0f ? bfr((float)d) : 0.0f;
      g[OFF_B1 + k] += dZ1;
      for (int m = 0; m < MLP_K1; ++m) g[OFF_W1 + k * 16 + m] += (double)dZ1 * (double)c.x[m];
    }
  }
  for (int i = 0; i < MLP_PARAMS; ++i) grad[i] = (float)g[i];
  return (float)(loss / (2.0 * n));
}

}  // namespace orc
