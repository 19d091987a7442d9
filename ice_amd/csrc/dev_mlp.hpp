// The learned seed of the Langmuir-Knudsen interface solve: a 6-64-64-2
// multilayer perceptron (ReLU hidden layers) in BF16 on the MFMA units of
// gfx950, FP32 accumulation.
//
// Parameters, one FP32 master vector of kMlpParams floats:
//   W1t[64][16]  b1[64]  W2t[64][64]  b2[64]  W3t[16][64]  b3[16]
// Wt[n][k] is the weight from input k to output n.  Inputs are padded from 6 to
// 16 and outputs from 2 to 16 with zeros; those rows and columns stay zero
// (their gradients are zero).  The device works with BF16 copies of the
// weights, rounded to nearest even, in two layouts: Wt[n][k] (forward) and
// Wn[k][n] (backward products).
//
// A wave computes 64 samples at once: four 16-sample tiles, each layer a sum of
// v_mfma_f32_16x16x16_bf16 products (lane l holds A[l%16][4(l/16)+j],
// B[4(l/16)+j][l%16] and D[4(l/16)+r][l%16]).  Hidden activations are rounded
// to BF16 after the bias and the ReLU, in LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ice {

constexpr int kMlpIn = 6, kMlpH = 64, kMlpOut = 2, kMlpK1 = 16, kMlpN3 = 16;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpK1;       // 1024
constexpr int kOffW2 = kOffB1 + kMlpH;                // 1088
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;        // 5184
constexpr int kOffW3 = kOffB2 + kMlpH;                // 5248
constexpr int kOffB3 = kOffW3 + kMlpN3 * kMlpH;       // 6272
constexpr int kMlpParams = kOffB3 + kMlpN3;           // 6288

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// The BF16 weights the kernels read (device memory), and the FP32 biases.
struct MlpDev {
  const uint16_t* w1t;  // [64][16]
  const uint16_t* w2t;  // [64][64]
  const uint16_t* w3t;  // [16][64]
  const uint16_t* w2n;  // [64][64]  (k, n)
  const uint16_t* w3n;  // [64][16]  (k, n)
  const float* b1;      // [64]
  const float* b2;      // [64]
  const float* b3;      // [16]
};

// Four consecutive BF16 values (8-byte aligned) as an MFMA operand
__device__ __forceinline__ bf16x4_t load4(const uint16_t* p) { return *(const bf16x4_t*)p; }
// Four values at stride s
__device__ __forceinline__ bf16x4_t load4s(const uint16_t* p, int s) {
  bf16x4_t v;
  v[0] = (short)p[0];
  v[1] = (short)p[s];
  v[2] = (short)p[2 * s];
  v[3] = (short)p[3 * s];
  return v;
}

// One 16x16 output tile of D = A B over K (a multiple of 16) for the wave:
// A row i at a + i*lda with its k values contiguous, B column j at b + j*ldb
// with its k values contiguous (both operands "k-contiguous").
__device__ __forceinline__ f32x4_t wave_tile(const uint16_t* a, int lda, const uint16_t* b, int ldb, int K,
                                             f32x4_t acc) {
  const int l = threadIdx.x & 63, i = l & 15, kg = (l >> 4) * 4;
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t av = load4(a + i * lda + k0 + kg);
    bf16x4_t bv = load4(b + i * ldb + k0 + kg);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

// Bias + ReLU of a D tile, rounded to BF16, stored row-major at h (row stride
// ldh) and, when ht is given, transposed at ht (row stride ldt).
__device__ __forceinline__ void store_relu(f32x4_t d, const float* bias, uint16_t* h, int ldh, uint16_t* ht, int ldt) {
  const int l = threadIdx.x & 63, col = l & 15, r0 = (l >> 4) * 4;
  const float bb = bias[col];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float z = d[r] + bb;
    uint16_t v = f2bf(z > 0.0f ? z : 0.0f);
    h[(r0 + r) * ldh + col] = v;
    if (ht) ht[col * ldt + r0 + r] = v;
  }
}

// Forward pass of the wave's 64 samples.  x: LDS [64][16] BF16 inputs; h1, h2:
// LDS [64][64] scratch (h1 may alias h2 for inference: each 16-row tile of h2
// is written after the tile's reads of h1); out: LDS [64][2] FP32.  With h1t,
// h2t the transposed activations are kept too (training).
__device__ inline void mlp_forward_wave(const MlpDev& w, const uint16_t* x, uint16_t* h1, uint16_t* h2, float* out,
                                        uint16_t* h1t = nullptr, uint16_t* h2t = nullptr) {
  const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
  for (int t = 0; t < 4; ++t)
    for (int n = 0; n < 4; ++n) {
      f32x4_t d = wave_tile(x + t * 16 * kMlpK1, kMlpK1, w.w1t + n * 16 * kMlpK1, kMlpK1, kMlpK1, zero);
      store_relu(d, w.b1 + n * 16, h1 + t * 16 * kMlpH + n * 16, kMlpH, h1t ? h1t + n * 16 * 64 + t * 16 : nullptr,
                 64);
    }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  for (int t = 0; t < 4; ++t) {
    f32x4_t d[4];
    for (int n = 0; n < 4; ++n)
      d[n] = wave_tile(h1 + t * 16 * kMlpH, kMlpH, w.w2t + n * 16 * kMlpH, kMlpH, kMlpH, zero);
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    for (int n = 0; n < 4; ++n)
      store_relu(d[n], w.b2 + n * 16, h2 + t * 16 * kMlpH + n * 16, kMlpH, h2t ? h2t + n * 16 * 64 + t * 16 : nullptr,
                 64);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const int l = threadIdx.x & 63, col = l & 15, r0 = (l >> 4) * 4;
  for (int t = 0; t < 4; ++t) {
    f32x4_t d = wave_tile(h2 + t * 16 * kMlpH, kMlpH, w.w3t, kMlpH, kMlpH, zero);
    if (col < kMlpOut) {
#pragma unroll
      for (int r = 0; r < 4; ++r) out[(t * 16 + r0 + r) * kMlpOut + col] = d[r] + w.b3[col];
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
}

// The six features of a drop's interface solve, scaled to O(1):
// gas temperature, pressure, far-field vapour fraction, drop temperature, drop
// diameter and slip Reynolds number.
__host__ __device__ inline void mlp_features(double Tg, double p, double yinf, double Tp, double dp, double Re,
                                             float* x) {
  x[0] = (float)(Tg / 1000.0 - 1.0);
  x[1] = (float)log(p / 1.0e5);
  x[2] = (float)(3.0 * yinf);
  x[3] = (float)((Tp - 315.0) / 35.0);
  x[4] = (float)log(dp / 4.0e-5);
  x[5] = (float)log(1.0 + Re);
}

}  // namespace ice
