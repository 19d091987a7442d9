// The learned seed of the interface solve: sample generation by the plain
// solver, training on the GPU (BF16 MFMA forward and backward passes, FP32
// accumulation, Adam on FP32 master weights) and its C-ABI.  Part of the
// context's translation unit (included at the end of ice_gpu.hip).
#pragma once

namespace {

constexpr int kPartStride = kMlpParams + 1;  // gradient partial of one wave + its loss
constexpr int kWOff1t = 0, kWOff2t = 1024, kWOff3t = 5120, kWOff2n = 6144, kWOff3n = 10240, kWTotal = 11264;

MlpDev mlp_view(const uint16_t* w, const float* P) {
  MlpDev m;
  m.w1t = w + kWOff1t;
  m.w2t = w + kWOff2t;
  m.w3t = w + kWOff3t;
  m.w2n = w + kWOff2n;
  m.w3n = w + kWOff3n;
  m.b1 = P + kOffB1;
  m.b2 = P + kOffB2;
  m.b3 = P + kOffB3;
  return m;
}

// BF16 copies of the FP32 master weights, in both layouts
__global__ void k_mlp_pack(const float* __restrict__ P, uint16_t* __restrict__ w) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 1024) {  // W1t [64][16]
    w[kWOff1t + t] = f2bf(P[kOffW1 + t]);
  } else if (t < 5120) {  // W2t [64][64] and W2n [k][n]
    int e = t - 1024, n = e / 64, k = e % 64;
    uint16_t v = f2bf(P[kOffW2 + e]);
    w[kWOff2t + e] = v;
    w[kWOff2n + k * 64 + n] = v;
  } else if (t < 6144) {  // W3t [16][64] and W3n [64][16]
    int e = t - 5120, n = e / 64, k = e % 64;
    uint16_t v = f2bf(P[kOffW3 + e]);
    w[kWOff3t + e] = v;
    w[kWOff3n + k * 16 + n] = v;
  }
}

// LDS of one training wave (uint16 units)
constexpr int kTrX = 0, kTrXt = kTrX + 64 * 16, kTrH1 = kTrXt + 16 * 64, kTrH1t = kTrH1 + 64 * 64,
              kTrH2 = kTrH1t + 64 * 64, kTrH2t = kTrH2 + 64 * 64, kTrDO = kTrH2t + 64 * 64, kTrDOt = kTrDO + 64 * 16,
              kTrDZ2 = kTrDOt + 16 * 64, kTrDZ2t = kTrDZ2 + 64 * 64, kTrOut = kTrDZ2t + 64 * 64,
              kTrTotal = kTrOut + 64 * 2 * 2;

// Loss (mean squared error over n samples and the 2 outputs) and its gradient
// with respect to every parameter, for 64 samples per wave: one workgroup is
// one wave, looping over sample chunks blockIdx.x, +gridDim.x, ...  Deltas are
// rounded to BF16 for the products, like the activations; the bias gradients
// sum the rounded deltas.  Writes its partial sums to part[blockIdx.x].
__global__ __launch_bounds__(64) void k_mlp_grad(MlpDev w, const float* __restrict__ X, const float* __restrict__ Y,
                                                 int n, float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) uint16_t L[kTrTotal];
  uint16_t *x = L + kTrX, *xt = L + kTrXt, *h1 = L + kTrH1, *h1t = L + kTrH1t, *h2 = L + kTrH2, *h2t = L + kTrH2t;
  uint16_t *dO = L + kTrDO, *dOt = L + kTrDOt, *dZ2 = L + kTrDZ2, *dZ2t = L + kTrDZ2t;
  uint16_t* dZ1t = h2;  // h2 is free once dZ2 is formed
  float* out = (float*)(L + kTrOut);
  const int l = threadIdx.x, col = l & 15, r0 = (l >> 4) * 4;
  const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
  f32x4_t g1[4], g2[4][4], g3[4];
  for (int a = 0; a < 4; ++a) {
    g1[a] = zero;
    g3[a] = zero;
    for (int b = 0; b < 4; ++b) g2[a][b] = zero;
  }
  float db1 = 0.f, db2 = 0.f, db3 = 0.f, loss = 0.f;
  const float scale = 1.0f / (float)(2 * n);
  for (int k = 0; k < 16 * 64; k += 64) dOt[k + l] = 0;
  for (int k = 0; k < 64 * 16; k += 64) dO[k + l] = 0;
  for (long long c0 = (long long)blockIdx.x * 64; c0 < n; c0 += (long long)gridDim.x * 64) {
    const long long s = c0 + l;
    const bool valid = s < n;
    for (int k = 0; k < 16; ++k) {
      uint16_t v = (valid && k < kMlpIn) ? f2bf(X[s * kMlpIn + k]) : (uint16_t)0;
      x[l * 16 + k] = v;
      xt[k * 64 + l] = v;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    mlp_forward_wave(w, x, h1, h2, out, h1t, h2t);
    // output deltas
    for (int c = 0; c < kMlpOut; ++c) {
      float e = valid ? out[l * kMlpOut + c] - Y[s * kMlpOut + c] : 0.0f;
      loss += e * e;
      uint16_t d = f2bf(2.0f * e * scale);
      dO[l * 16 + c] = d;
      dOt[c * 64 + l] = d;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (l < kMlpOut)
      for (int q = 0; q < 64; ++q) db3 += bf2f(dOt[l * 64 + q]);
    // dW3[c][k] += sum_s dO[s][c] h2[s][k]
    for (int kt = 0; kt < 4; ++kt) g3[kt] = wave_tile(dOt, 64, h2t + kt * 16 * 64, 64, 64, g3[kt]);
    // dZ2 = (dO W3) * [h2 > 0]
    for (int t = 0; t < 4; ++t)
      for (int kt = 0; kt < 4; ++kt) {
        f32x4_t d = wave_tile(dO + t * 16 * 16, 16, w.w3n + kt * 16 * 16, 16, 16, zero);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int sr = t * 16 + r0 + r, u = kt * 16 + col;
          uint16_t v = bf2f(h2[sr * 64 + u]) > 0.0f ? f2bf(d[r]) : (uint16_t)0;
          dZ2[sr * 64 + u] = v;
          dZ2t[u * 64 + sr] = v;
        }
      }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    for (int q = 0; q < 64; ++q) db2 += bf2f(dZ2t[l * 64 + q]);
    // dW2[n][k] += sum_s dZ2[s][n] h1[s][k]
    for (int nt = 0; nt < 4; ++nt)
      for (int kt = 0; kt < 4; ++kt) g2[nt][kt] = wave_tile(dZ2t + nt * 16 * 64, 64, h1t + kt * 16 * 64, 64, 64, g2[nt][kt]);
    // dZ1 = (dZ2 W2) * [h1 > 0], kept transposed in the space of h2
    for (int t = 0; t < 4; ++t)
      for (int kt = 0; kt < 4; ++kt) {
        f32x4_t d = wave_tile(dZ2 + t * 16 * 64, 64, w.w2n + kt * 16 * 64, 64, 64, zero);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int sr = t * 16 + r0 + r, u = kt * 16 + col;
          dZ1t[u * 64 + sr] = bf2f(h1[sr * 64 + u]) > 0.0f ? f2bf(d[r]) : (uint16_t)0;
        }
      }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    for (int q = 0; q < 64; ++q) db1 += bf2f(dZ1t[l * 64 + q]);
    // dW1[n][k] += sum_s dZ1[s][n] x[s][k]
    for (int nt = 0; nt < 4; ++nt) g1[nt] = wave_tile(dZ1t + nt * 16 * 64, 64, xt, 64, 64, g1[nt]);
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
  }
  float* o = part + (long long)blockIdx.x * kPartStride;
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      o[kOffW1 + (nt * 16 + r0 + r) * 16 + col] = g1[nt][r];
      o[kOffW3 + (r0 + r) * 64 + nt * 16 + col] = g3[nt][r];
      for (int kt = 0; kt < 4; ++kt) o[kOffW2 + (nt * 16 + r0 + r) * 64 + kt * 16 + col] = g2[nt][kt][r];
    }
  o[kOffB1 + l] = db1;
  o[kOffB2 + l] = db2;
  if (l < kMlpN3) o[kOffB3 + l] = l < kMlpOut ? db3 : 0.0f;
  // the wave's loss: lanes summed in a fixed order through LDS
  out[l] = loss;
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  if (l == 0) {
    float t = 0.f;
    for (int q = 0; q < 64; ++q) t += out[q];
    o[kMlpParams] = t / (float)(2 * n);
  }
}

// Sum of the partials in workgroup order; then, with lr > 0, one Adam step on
// the master weights (step number t >= 1).  grad (optional) receives the sum.
__global__ void k_mlp_reduce_adam(const float* __restrict__ part, int nparts, float* __restrict__ P, float* m, float* v,
                                  float* grad, float* loss, float lr, int t) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > kMlpParams) return;
  float g = 0.f;
  for (int q = 0; q < nparts; ++q) g += part[(long long)q * kPartStride + i];
  if (i == kMlpParams) {
    if (loss) *loss = g;
    return;
  }
  if (grad) grad[i] = g;
  if (lr > 0.f) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float mi = b1 * m[i] + (1.f - b1) * g;
    float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    float mh = mi / (1.f - powf(b1, (float)t));
    float vh = vi / (1.f - powf(b2, (float)t));
    P[i] -= lr * mh / (sqrtf(vh) + eps);
  }
}

// Outputs of the network for n samples (one wave per workgroup)
__global__ __launch_bounds__(64) void k_mlp_infer(MlpDev w, const float* __restrict__ X, int n, float* __restrict__ Y) {
  __shared__ __attribute__((aligned(16))) uint16_t L[64 * 16 + 64 * 64 + 64 * 2 * 2];
  uint16_t* x = L;
  uint16_t* h = L + 64 * 16;
  float* out = (float*)(h + 64 * 64);
  const int l = threadIdx.x;
  for (long long c0 = (long long)blockIdx.x * 64; c0 < n; c0 += (long long)gridDim.x * 64) {
    const long long s = c0 + l;
    for (int k = 0; k < 16; ++k) x[l * 16 + k] = (s < n && k < kMlpIn) ? f2bf(X[s * kMlpIn + k]) : (uint16_t)0;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    mlp_forward_wave(w, x, h, h, out);
    if (s < n)
      for (int c = 0; c < kMlpOut; ++c) Y[s * kMlpOut + c] = out[l * kMlpOut + c];
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
  }
}

// Training samples from the plain solver: every interior cell of block b of the
// family whose drop runs the interface solve gives its six features and the
// solved interface state (rate over equilibrium rate, surface vapour molar
// fraction over its equilibrium value), appended at an atomic counter.
template <int M>
__global__ __launch_bounds__(kThreads) void k_lk_samples(FamArgs a, int b, float* X, float* Y, unsigned* count, int cap) {
  constexpr int NC = Closure<M>::NC;
  const BlkDev B = a.blk[b];
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.cells) return;
  int i, j, k;
  cell_of(B, t, i, j, k);
  long long g = gidx(B, i, j, k);
  double q[NC], gp[5], ep[kNepCell];
  load_cell<NC>(a.cur, a.N, g, q);
  for (int m = 0; m < 5; ++m) gp[m] = a.gas[(long long)m * a.N + g];
  cell_ep(a, g, q[NC - 2], ep);
  double p;
  if (!lk_needed(gp[0], gp[4], a.gR[g], q[NC - 2], a.mat.evap, a.mat.intf, ep, p)) return;
  double Rp, slip, Re;
  drop_state<M>(a, g, q, gp, Rp, slip, Re);
  GasSide s;
  s.evap = a.mat.evap;
  double Xs, Ys, BM;
  if (!evap_surface(s, gp[0], gp[4], a.gG[g], a.gR[g], a.gMu[g], a.gK[g], q[NC - 2], 2.0 * Rp, Re, ep, Xs, Ys, BM))
    return;
  gas_side_setup(s, ep);
  double md0, Qd;
  bool ovr;
  gas_side_rate_d(s, Ys, BM, md0, Qd, ovr);
  if (!(md0 < 0.0)) return;
  double LKd2 = lk_factor(s, ep), md = md0, mf;
  int ev = 0;
  if (lk_secant(s, Xs, LKd2, md0, md0, mf, Qd, ovr, ev)) md = mf;   // the fixed point to round-off
  else lk_picard_counted(s, Xs, LKd2, md, Qd, ovr, ev);
  double beta = -md * s.Pr / (2.0 * kPi * s.mug * s.dp);
  double XsN = dmax2(Xs - LKd2 * beta, 0.0);
  unsigned slot = atomicAdd(count, 1u);
  if ((int)slot >= cap) return;
  float f[kMlpIn];
  mlp_features(gp[4], p, ep[kYinf], q[NC - 2], 2.0 * Rp, Re, f);
  for (int m = 0; m < kMlpIn; ++m) X[(long long)slot * kMlpIn + m] = f[m];
  Y[(long long)slot * kMlpOut + 0] = (float)(md / md0);
  Y[(long long)slot * kMlpOut + 1] = (float)(XsN / Xs);
}

// splitmix64 stream for the initial weights
uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// He-uniform weights (fan-in 6, 64, 64), zero biases, zero padding
void mlp_init(std::vector<float>& P, uint64_t seed) {
  P.assign(kMlpParams, 0.0f);
  uint64_t st = seed;
  auto uni = [&](float lim) { return lim * (2.0f * (float)((splitmix(st) >> 11) * (1.0 / 9007199254740992.0)) - 1.0f); };
  for (int n = 0; n < 64; ++n)
    for (int k = 0; k < kMlpIn; ++k) P[kOffW1 + n * 16 + k] = uni(std::sqrt(6.0f / kMlpIn));
  for (int e = 0; e < 64 * 64; ++e) P[kOffW2 + e] = uni(std::sqrt(6.0f / 64.0f));
  for (int c = 0; c < kMlpOut; ++c)
    for (int k = 0; k < 64; ++k) P[kOffW3 + c * 64 + k] = uni(std::sqrt(6.0f / 64.0f));
}

int surr_alloc(ice_gpu_ctx* c, SurrDev& S) {
  if (S.P) return 0;
  if (dalloc(c, &S.P, kMlpParams + 16) || dalloc(c, &S.m, kMlpParams) || dalloc(c, &S.v, kMlpParams) ||
      dalloc(c, &S.w, kWTotal) || dalloc(c, &S.stats, 4) || dalloc(c, &S.loss, 1))
    return 1;
  S.nparts = 256;
  if (dalloc(c, &S.part, (size_t)S.nparts * kPartStride)) return 1;
  return 0;
}

hipStream_t stream_of(ice_gpu_ctx* c) { return c ? c->stream : nullptr; }

int surr_pack(ice_gpu_ctx* c, SurrDev& S) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(24), dim3(256), 0, stream_of(c), S.P, S.w);
  HIPCHK(hipGetLastError());
  return 0;
}

// loss and gradient (grad may be null) at the current weights; optional Adam step
int surr_grad(ice_gpu_ctx* c, SurrDev& S, const float* dX, const float* dY, int n, float* dgrad, float lr, int t) {
  int nwg = std::min(S.nparts, std::max(1, (n + 63) / 64));
  hipLaunchKernelGGL(k_mlp_grad, dim3(nwg), dim3(64), 0, stream_of(c), mlp_view(S.w, S.P), dX, dY, n, S.part);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((kMlpParams + 256) / 256), dim3(256), 0, stream_of(c), S.part, nwg, S.P, S.m,
                     S.v, dgrad, S.loss, lr, t);
  HIPCHK(hipGetLastError());
  if (lr > 0.f) return surr_pack(c, S);
  return 0;
}

int launch_samples(ice_gpu_ctx* c, int p, float* dX, float* dY, unsigned* dcount, int cap) {
  FamArgs a = fam_args(c, p);
  for (size_t b = 0; b < c->blk.size(); ++b) {
    dim3 grid(grid_of(c->blk[b].cells));
    switch (c->fams[p].closure) {
      case ICE_MK: hipLaunchKernelGGL(k_lk_samples<0>, grid, dim3(kThreads), 0, c->stream, a, (int)b, dX, dY, dcount, cap); break;
      case ICE_IG: hipLaunchKernelGGL(k_lk_samples<1>, grid, dim3(kThreads), 0, c->stream, a, (int)b, dX, dY, dcount, cap); break;
      default: hipLaunchKernelGGL(k_lk_samples<2>, grid, dim3(kThreads), 0, c->stream, a, (int)b, dX, dY, dcount, cap); break;
    }
    HIPCHK(hipGetLastError());
  }
  return 0;
}

// a temporary context for the stand-alone network calls
struct TmpBuf {
  std::vector<void*> p;
  ~TmpBuf() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T> int get(ice_gpu_ctx* c, T** out, size_t n) {
    if (dalloc(c, out, n)) return 1;
    p.push_back((void*)*out);
    return 0;
  }
};

}  // namespace

extern "C" {

// The stand-alone network calls take ctx = NULL (the default stream, no error text).
int ice_gpu_mlp_grad(ice_gpu_ctx* c, int n, const float* x, const float* y, const float* params, float* grad,
                     float* loss) {
  if (n < 1) return fail(c, "[ICE::mlp_grad] no samples");
  SurrDev S;
  TmpBuf T;
  float *dX, *dY, *dG;
  if (T.get(c, &S.P, kMlpParams + 16) || T.get(c, &S.w, kWTotal) || T.get(c, &S.loss, 1) ||
      T.get(c, &dX, (size_t)n * kMlpIn) || T.get(c, &dY, (size_t)n * kMlpOut) || T.get(c, &dG, kMlpParams))
    return 1;
  S.nparts = 256;
  if (T.get(c, &S.part, (size_t)S.nparts * kPartStride)) return 1;
  HIPCHK(hipMemcpy(S.P, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dX, x, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dY, y, (size_t)n * kMlpOut * sizeof(float), hipMemcpyHostToDevice));
  if (surr_pack(c, S) || surr_grad(c, S, dX, dY, n, dG, 0.f, 0)) return 1;
  HIPCHK(hipStreamSynchronize(stream_of(c)));
  HIPCHK(hipMemcpy(grad, dG, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(loss, S.loss, sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int ice_gpu_mlp_forward(ice_gpu_ctx* c, int n, const float* x, const float* params, float* out) {
  if (n < 1) return fail(c, "[ICE::mlp_forward] no samples");
  SurrDev S;
  TmpBuf T;
  float *dX, *dY;
  if (T.get(c, &S.P, kMlpParams + 16) || T.get(c, &S.w, kWTotal) || T.get(c, &dX, (size_t)n * kMlpIn) ||
      T.get(c, &dY, (size_t)n * kMlpOut))
    return 1;
  HIPCHK(hipMemcpy(S.P, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dX, x, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice));
  if (surr_pack(c, S)) return 1;
  int nwg = std::min(1024, (n + 63) / 64);
  hipLaunchKernelGGL(k_mlp_infer, dim3(nwg), dim3(64), 0, stream_of(c), mlp_view(S.w, S.P), dX, n, dY);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream_of(c)));
  HIPCHK(hipMemcpy(out, dY, (size_t)n * kMlpOut * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// Samples family p's current state with the plain solver and trains its network
// on them; warm: from its current weights (online retraining), else from fresh
// He-uniform weights.
static int surr_train_impl(ice_gpu_ctx* c, int p, int epochs, int batch, double lr, uint64_t seed, float* report,
                           bool warm) {
  if (p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::surrogate_train] no family %d", p);
  const ice_family_t& fd = c->fams[p - 1];
  if (fd.evap_select == 0 || fd.intf_select != 1)
    return fail(c, "[ICE::surrogate_train] family %d has no Langmuir-Knudsen interface solve", p);
  if (!c->cfg.coupled) return fail(c, "[ICE::surrogate_train] the run is not coupled to a gas");
  if (batch < 64) batch = 64;
  SurrDev& S = c->surr[p - 1];
  if (surr_alloc(c, S)) return 1;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  // samples of the current state, by the plain solver
  const int cap = 1 << 20;
  TmpBuf T;
  float *dX, *dY, *dOut;
  unsigned* dcount;
  if (T.get(c, &dX, (size_t)cap * kMlpIn) || T.get(c, &dY, (size_t)cap * kMlpOut) ||
      T.get(c, &dOut, (size_t)cap * kMlpOut) || T.get(c, &dcount, 1))
    return 1;
  const unsigned zero = 0u;
  HIPCHK(hipMemcpy(dcount, &zero, sizeof zero, hipMemcpyHostToDevice));
  if (launch_samples(c, p - 1, dX, dY, dcount, cap)) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));  // the context's stream does not block with the null stream
  unsigned ns = 0;
  HIPCHK(hipMemcpy(&ns, dcount, sizeof ns, hipMemcpyDeviceToHost));
  int n = (int)std::min<unsigned>(ns, (unsigned)cap);
  if (n < 64) return fail(c, "[ICE::surrogate_train] %d cells run the interface solve: too few samples", n);
  // the samples arrive in the order the atomics gave them: put them in a canonical
  // order (by their values), then shuffle once on the host (fixed seed), so that
  // mini-batches are spread over the domain and the training repeats bit for bit
  {
    std::vector<float> hx((size_t)n * kMlpIn), hy((size_t)n * kMlpOut);
    HIPCHK(hipMemcpy(hx.data(), dX, hx.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hy.data(), dY, hy.size() * sizeof(float), hipMemcpyDeviceToHost));
    {
      std::vector<int> order(n);
      for (int i = 0; i < n; ++i) order[i] = i;
      auto row = [&](int i, int m) { return m < kMlpIn ? hx[(size_t)i * kMlpIn + m] : hy[(size_t)i * kMlpOut + m - kMlpIn]; };
      std::sort(order.begin(), order.end(), [&](int a, int b) {
        for (int m = 0; m < kMlpIn + kMlpOut; ++m)
          if (row(a, m) != row(b, m)) return row(a, m) < row(b, m);
        return false;
      });
      std::vector<float> sx(hx.size()), sy(hy.size());
      for (int i = 0; i < n; ++i) {
        std::copy(&hx[(size_t)order[i] * kMlpIn], &hx[(size_t)order[i] * kMlpIn] + kMlpIn, &sx[(size_t)i * kMlpIn]);
        std::copy(&hy[(size_t)order[i] * kMlpOut], &hy[(size_t)order[i] * kMlpOut] + kMlpOut, &sy[(size_t)i * kMlpOut]);
      }
      hx.swap(sx);
      hy.swap(sy);
    }
    uint64_t st = seed ^ 0x5DEECE66Dull;
    for (int i = n - 1; i > 0; --i) {
      int j = (int)(splitmix(st) % (uint64_t)(i + 1));
      for (int m = 0; m < kMlpIn; ++m) std::swap(hx[(size_t)i * kMlpIn + m], hx[(size_t)j * kMlpIn + m]);
      for (int m = 0; m < kMlpOut; ++m) std::swap(hy[(size_t)i * kMlpOut + m], hy[(size_t)j * kMlpOut + m]);
    }
    HIPCHK(hipMemcpy(dX, hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dY, hy.data(), hy.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  std::vector<float> P;
  mlp_init(P, seed);
  std::vector<float> zeros(kMlpParams, 0.0f);
  if (!warm || !S.on) HIPCHK(hipMemcpy(S.P, P.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(S.m, zeros.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(S.v, zeros.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  if (surr_pack(c, S)) return 1;
  float loss0 = 0.f, loss1 = 0.f;
  batch = std::max(batch, (n + 255) / 256);  // at most 256 optimiser steps per pass over large sample sets
  if (surr_grad(c, S, dX, dY, n, nullptr, 0.f, 0)) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(&loss0, S.loss, sizeof(float), hipMemcpyDeviceToHost));
  int t = 0;
  for (int e = 0; e < epochs; ++e)
    for (int b0 = 0; b0 + batch <= n || (b0 == 0 && n < batch); b0 += batch) {
      int nb = std::min(batch, n - b0);
      float lr_t = (float)(lr * (e < epochs / 2 ? 1.0 : (e < 3 * epochs / 4 ? 0.3 : 0.1)));
      if (surr_grad(c, S, dX + (size_t)b0 * kMlpIn, dY + (size_t)b0 * kMlpOut, nb, nullptr, lr_t, ++t)) return 1;
    }
  if (surr_grad(c, S, dX, dY, n, nullptr, 0.f, 0)) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(&loss1, S.loss, sizeof(float), hipMemcpyDeviceToHost));
  // raw error of the network's rate ratio against the solver's
  int nwg = std::min(1024, (n + 63) / 64);
  hipLaunchKernelGGL(k_mlp_infer, dim3(nwg), dim3(64), 0, c->stream, mlp_view(S.w, S.P), dX, n, dOut);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::vector<float> hy((size_t)n * kMlpOut), ho((size_t)n * kMlpOut);
  HIPCHK(hipMemcpy(hy.data(), dY, hy.size() * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ho.data(), dOut, ho.size() * sizeof(float), hipMemcpyDeviceToHost));
  double sum = 0.0, mx = 0.0;
  for (int i = 0; i < n; ++i) {
    double d = std::fabs((double)ho[(size_t)i * kMlpOut] - (double)hy[(size_t)i * kMlpOut]);
    sum += d;
    mx = std::max(mx, d);
  }
  if (!S.on) batched_destroy(c);  // captured graphs hold the families' arguments (mlp_on)
  S.on = 1;
  S.batch = batch;
  S.lr = lr;
  S.seed = seed;
  if (report) {
    report[0] = (float)n;
    report[1] = loss0;
    report[2] = loss1;
    report[3] = (float)(sum / n);
    report[4] = (float)mx;
    report[5] = ms;
  }
  return 0;
}

int ice_gpu_surrogate_train(ice_gpu_ctx* c, int p, int epochs, int batch, double lr, uint64_t seed, float* report) {
  return surr_train_impl(c, p, epochs, batch, lr, seed, report, false);
}

int ice_gpu_surrogate_retrain(ice_gpu_ctx* c, int p, int epochs) {
  if (p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::surrogate_retrain] no family %d", p);
  if (epochs < 1) return fail(c, "[ICE::surrogate_retrain] epochs >= 1");
  SurrDev& S = c->surr[p - 1];
  if (!S.on) return fail(c, "[ICE::surrogate_retrain] family %d has no trained network", p);
  return surr_train_impl(c, p, epochs, S.batch, S.lr, S.seed, nullptr, true);
}

int ice_gpu_surrogate_retrain_every(ice_gpu_ctx* c, int every, int epochs) {
  if (every < 0 || epochs < 1) return fail(c, "[ICE::surrogate_retrain_every] every >= 0 and epochs >= 1");
  c->retrain_every = every;
  c->retrain_epochs = epochs;
  return 0;
}

int ice_gpu_surrogate_params(ice_gpu_ctx* c, int p, float* params) {
  if (p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::surrogate_params] no family %d", p);
  SurrDev& S = c->surr[p - 1];
  if (!S.on) return fail(c, "[ICE::surrogate_params] family %d has no trained network", p);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(params, S.P, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int ice_gpu_surrogate_set_params(ice_gpu_ctx* c, int p, const float* params) {
  if (p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::surrogate_set_params] no family %d", p);
  SurrDev& S = c->surr[p - 1];
  if (surr_alloc(c, S)) return 1;
  HIPCHK(hipMemcpy(S.P, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  if (surr_pack(c, S)) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!S.on) batched_destroy(c);  // captured graphs hold the families' arguments (mlp_on)
  S.on = 1;
  return 0;
}

int ice_gpu_set_lk_stats(ice_gpu_ctx* c, int on) {
  HIPCHK(hipStreamSynchronize(c->stream));
  c->lk_stats_on = on ? 1 : 0;
  batched_destroy(c);  // captured graphs hold the families' arguments
  for (auto& S : c->surr) {
    if (!S.stats && dalloc(c, &S.stats, 4)) return 1;
    const unsigned long long z[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpy(S.stats, z, sizeof z, hipMemcpyHostToDevice));
  }
  return 0;
}

int ice_gpu_lk_stats(ice_gpu_ctx* c, int p, int64_t* out) {
  if (p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::lk_stats] no family %d", p);
  SurrDev& S = c->surr[p - 1];
  unsigned long long h[4] = {0, 0, 0, 0};
  HIPCHK(hipStreamSynchronize(c->stream));
  if (S.stats) HIPCHK(hipMemcpy(h, S.stats, sizeof h, hipMemcpyDeviceToHost));
  for (int m = 0; m < 4; ++m) out[m] = (int64_t)h[m];
  return 0;
}

}  // extern "C"

int retrain_networks(ice_gpu_ctx* c) {
  for (int p = 1; p <= c->cfg.ngroups; ++p) {
    SurrDev& S = c->surr[p - 1];
    if (!S.on) continue;
    if (surr_train_impl(c, p, c->retrain_epochs, S.batch, S.lr, S.seed, nullptr, true)) return 1;
  }
  return 0;
}
