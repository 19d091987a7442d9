// The kernels of one family's stage, plain path (one launch per block for the
// per-cell kernels, one per family for the per-record kernels).
#pragma once
#include "kernels.hpp"

namespace ice {

constexpr int kThreads = 256;
// Waves per SIMD the stage kernels are compiled for, i.e. their register budget
// (512 VGPRs per SIMD lane shared by the resident waves).  The stage waits on
// its face passes' loads and barriers about half of its time, so two waves per
// SIMD are the floor; the face pass keeps no stencil in registers (the limited
// slopes come from LDS), so every closure, AG's 12 slots included, fits 256
// registers.  ICE_MKIG_WAVES / ICE_AG_WAVES override it for measurements.
#ifndef ICE_MKIG_WAVES
#define ICE_MKIG_WAVES 2
#endif
template <int M> struct StageWaves { static constexpr int value = ICE_MKIG_WAVES; };
#ifndef ICE_AG_WAVES
#define ICE_AG_WAVES 2
#endif
template <> struct StageWaves<2> { static constexpr int value = ICE_AG_WAVES; };

// ------------------------------------------------------------ learned seed
// LDS of the seed per workgroup: per wave 64x16 inputs, 64x64 hidden (BF16)
// and 64x2 outputs (FP32).
constexpr int kMlpWaveLds = 64 * 16 + 64 * 64 + 64 * 2 * 2;  // in uint16_t
constexpr int kMlpLds = kMlpWaveLds * (kThreads / 64);

// The network's forward weights (BF16: W1t, W2t, W3t) and biases (FP32: b1, b2,
// b3) as a workgroup keeps them in LDS
constexpr int kMlpWts = kMlpH * kMlpK1 + kMlpH * kMlpH + kMlpN3 * kMlpH;  // 6144
constexpr int kMlpBias = kMlpH + kMlpH + kMlpN3;                            // 144

// Copies the network's weights and biases into the workgroup's LDS (when load;
// every thread of the workgroup calls this, and it ends with a barrier) and
// returns the network pointing at that copy, so that every product of the
// forward pass reads LDS instead of memory.  The copy holds the same values, so
// the outputs are the same bits.
ICE_DEV MlpDev mlp_in_lds(const MlpDev& w, bool load, uint16_t* wts, float* bias) {
  constexpr int o2 = kMlpH * kMlpK1, o3 = o2 + kMlpH * kMlpH;
  if (load) {
    const uint32_t* s1 = (const uint32_t*)w.w1t;
    const uint32_t* s2 = (const uint32_t*)w.w2t;
    const uint32_t* s3 = (const uint32_t*)w.w3t;
    uint32_t* d = (uint32_t*)wts;
    for (int e = threadIdx.x; e < kMlpWts / 2; e += blockDim.x)
      d[e] = e < o2 / 2 ? s1[e] : (e < o3 / 2 ? s2[e - o2 / 2] : s3[e - o3 / 2]);
    for (int e = threadIdx.x; e < kMlpBias; e += blockDim.x)
      bias[e] = e < kMlpH ? w.b1[e] : (e < 2 * kMlpH ? w.b2[e - kMlpH] : w.b3[e - 2 * kMlpH]);
    __syncthreads();
  }
  MlpDev m = w;
  m.w1t = wts;
  m.w2t = wts + o2;
  m.w3t = wts + o3;
  m.b1 = bias;
  m.b2 = bias + kMlpH;
  m.b3 = bias + 2 * kMlpH;
  return m;
}

// Seed of the interface solve of the calling thread's cell.  mat.surrogate:
// 0 the plain solve (0), 2 the equilibrium rate (1), 1 the network's rate ratio
// for the cell, clamped to [1e-2, 1].  With 1, every thread of the workgroup
// calls this (the wave's 64 lanes feed one MLP pass on the MFMA units, skipped
// when no lane's cell needs the solve).  net: the network's weights (a.mlp, or
// its copy in LDS, mlp_in_lds).
template <int M, bool TAB = true>
ICE_DEV float stage_seed(const FamArgs& a, bool has_cell, long long g, uint16_t* lds, const MlpDev& net) {
  constexpr int NC = Closure<M>::NC;
  if (a.mat.surrogate == 2) return 1.0f;
  if (a.mat.surrogate != 1 || !a.mlp_on || a.mat.evap == 0 || a.mat.intf != 1) return 0.0f;
  bool need = false;
  float f[kMlpIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (has_cell) {
    double q[NC], gp[5];
    load_cell<NC>(a.cur, a.N, g, q);
#pragma unroll
    for (int m = 0; m < 5; ++m) gp[m] = a.gas[(long long)m * a.N + g];
    double ep[kNepCell];
    cell_ep<TAB>(a, g, q[NC - 2], ep);
    double p;
    need = lk_needed(gp[0], gp[4], a.gR[g], q[NC - 2], a.mat.evap, a.mat.intf, ep, p);
    if (need) {
      double Rp, slip, Re;
      drop_state<M, TAB>(a, g, q, gp, Rp, slip, Re);
      mlp_features(gp[4], p, ep[kYinf], q[NC - 2], 2.0 * Rp, Re, f);
    }
  }
  if (!__any(need)) return 0.0f;
  const int l = threadIdx.x & 63;
  uint16_t* x = lds + (threadIdx.x >> 6) * kMlpWaveLds;
  uint16_t* h = x + 64 * 16;
  float* out = (float*)(h + 64 * 64);
#pragma unroll
  for (int k = 0; k < 16; ++k) x[l * 16 + k] = k < kMlpIn ? f2bf(f[k]) : (uint16_t)0;
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  mlp_forward_wave(net, x, h, h, out);
  float r = out[l * kMlpOut];
  r = r < 1e-2f ? 1e-2f : (r > 1.0f ? 1.0f : r);
  return need ? r : 0.0f;
}

ICE_DEV void cell_of(const BlkDev& B, long long t, int& i, int& j, int& k) {
  i = (int)(t % B.ni) + 1;
  long long r = t / B.ni;
  j = (int)(r % B.nj) + 1;
  k = (int)(r / B.nj) + 1;
}

ICE_DEV long long dir_stride(const BlkDev& B, int d) {
  return d == 0 ? 1ll : (d == 1 ? (long long)B.NI : (long long)B.NI * (long long)B.NJ);
}

// block-wide minimum of a positive double, then one atomic per workgroup
ICE_DEV void block_min_atomic(double v, unsigned long long* target) {
  __shared__ double wmin[kThreads / 64];
  for (int o = 32; o > 0; o >>= 1) v = dmin2(v, __shfl_xor(v, o, 64));
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wmin[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmin[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = dmin2(m, wmin[q]);
    atomicMin(target, (unsigned long long)__double_as_longlong(m));
  }
}

// ------------------------------------------------------------------ dt
// compute_dt / compute_dt_ (Mod_dt.f90:11-112, 256-275)
template <int M>
ICE_DEV double dt_cell(const FamArgs& a, long long g) {
  constexpr int NC = Closure<M>::NC;
  double q[NC];
  load_cell<NC>(a.cur, a.N, g, q);
  double dt = 1e+5;
  double rate = 0.0;
  for (int d = 0; d < a.ndir; ++d) {
    double r0 = a.M[(long long)(d * 3 + 0) * a.N + g];
    double r1 = a.M[(long long)(d * 3 + 1) * a.N + g];
    double r2 = a.M[(long long)(d * 3 + 2) * a.N + g];
    double gn = norm3(r0, r1, r2);
    double vs[3] = {r0 / gn, r1 / gn, r2 / gn};
    double un = q[1] * vs[0] + q[2] * vs[1] + q[3] * vs[2];
    rate = rate + (fabs(un) + wavespeed<M>(q, vs)) * gn;
  }
  if (rate > 0.0) dt = dmin2(dt, 1.0 / rate);
  dt = dt * a.cfl;
  int iter = a.st->iter;
  if (iter <= a.cfl_ramp_iter) dt = dt * (double)iter / (double)a.cfl_ramp_iter;
  dt = dmin2(dt, a.dt_max);
  if (a.coupled && a.tau_factor > 0.0) {
    double tau = a.tau[g];
    if (q[0] > kRhoEmpty && !nonfinite(tau) && tau > 0.0) dt = dmin2(dt, a.tau_factor * tau);
  }
  if (!a.time_accurate) a.dt[g] = dt;
  return dt;
}

// ------------------------------------------------------------------ ghosts
enum { kCopy = 0, kMirror = 1, kInject = 2 };

ICE_DEV int ghost_rule(const DRec& r, double veln) {
  int rule = kCopy;
  if (r.type == 200) rule = kMirror;
  else if (r.type == 300 || r.type == 301) {
    if (veln * (double)(1 - 2 * (r.f % 2)) <= 0.0 || r.f == 3) rule = kMirror;
  } else if (r.type >= 401 && r.type <= 403) {
    if (!(veln * (double)(1 - 2 * (r.f % 2)) > 0.0)) {
      rule = kInject;
      if (r.massflux == 0.0) rule = kMirror;
    }
  }
  return rule;
}

// normal of the boundary face of a record (Lib_Ghost.f90:111-124)
ICE_DEV void bnd_normal(const FamArgs& a, const BlkDev& B, const DRec& r, double* nrm) {
  int i = r.i, j = r.j, k = r.k, d;
  if (r.f <= 2) { i -= r.f % 2; d = 0; }
  else if (r.f <= 4) { j -= r.f % 2; d = 1; }
  else { k -= r.f % 2; d = 2; }
  long long g = gidx(B, i, j, k);
  const double* fa = a.fa + (long long)(d * 4) * a.N;
  nrm[0] = fa[g];
  nrm[1] = fa[a.N + g];
  nrm[2] = fa[2 * a.N + g];
}

// conservative blend of chimera donors into ghost layer lay (Lib_Ghost.f90:439-466)
template <int M> ICE_DEV void chimera(const FamArgs& a, const DRec& r, int lay) {
  constexpr int NC = Closure<M>::NC;
  int c1 = lay == 1 ? 1 : r.nchi1 + 1;
  int c2 = lay == 1 ? r.nchi1 : r.nchi1 + r.nchi2;
  double cg[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) cg[v] = 0.0;
  for (int c = c1; c <= c2; ++c) {
    const int* d = a.donor + (long long)(r.donor0 + c - 1) * 4;
    const BlkDev Bd = a.blk[d[0] - 1];
    double q[NC], cs[NC];
    load_cell<NC>(a.prim_w, a.N, gidx(Bd, d[1], d[2], d[3]), q);
    prim2cons<M>(q, cs, a.mat);
    double w = a.weight[r.donor0 + c - 1];
#pragma unroll
    for (int v = 0; v < NC; ++v) cg[v] = cg[v] + w * cs[v];
  }
  const BlkDev B = a.blk[r.b - 1];
  const int* G = kGuide[r.f - 1];
  double q[NC];
  cons2prim<M>(cg, q, a.mat);
  store_cell<NC>(a.prim_w, a.N, gidx(B, r.i - lay * G[0], r.j - lay * G[1], r.k - lay * G[2]), q);
}

// first ghost layer (ghost_record, Lib_Ghost.f90:76-267)
template <int M>
__global__ void __launch_bounds__(kThreads) k_ghost1(FamArgs a) {
  constexpr int NC = Closure<M>::NC;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nlist) return;
  const DRec r = a.rec[a.list[t]];
  const BlkDev B = a.blk[r.b - 1];
  const long long N = a.N;
  const int* G = kGuide[r.f - 1];
  long long gg = gidx(B, r.i - G[0], r.j - G[1], r.k - G[2]);
  long long gm = gidx(B, r.i, r.j, r.k);
  double* P = a.prim_w;
  double in[NC], gh[NC];
  switch (r.type) {
    case 101:
    case 201: {
      const BlkDev Bs = a.blk[r.bs - 1];
      load_cell<NC>(P, N, gidx(Bs, r.is, r.js, r.ks), gh);
      store_cell<NC>(P, N, gg, gh);
      break;
    }
    case 200:
    case 300:
    case 301: {
      double nrm[3];
      bnd_normal(a, B, r, nrm);
      load_cell<NC>(P, N, gm, in);
#pragma unroll
      for (int v = 0; v < NC; ++v) gh[v] = in[v];
      double vel[3] = {in[1], in[2], in[3]};
      double veln = vel[0] * nrm[0] + vel[1] * nrm[1] + vel[2] * nrm[2];
      if (ghost_rule(r, veln) == kMirror) {
        for (int m = 0; m < 3; ++m) vel[m] = vel[m] - 2.0 * veln * nrm[m];
        if constexpr (M == 2) mirror6(&in[4], nrm, &gh[4]);
      }
      gh[1] = vel[0];
      gh[2] = vel[1];
      gh[3] = vel[2];
      store_cell<NC>(P, N, gg, gh);
      break;
    }
    case 0:
    case 400:
      load_cell<NC>(P, N, gm, gh);
      store_cell<NC>(P, N, gg, gh);
      break;
    case 401:
    case 402:
    case 403: {
      double nrm[3];
      bnd_normal(a, B, r, nrm);
      load_cell<NC>(P, N, gm, in);
      double vel[3] = {in[1], in[2], in[3]};
      double veln = vel[0] * nrm[0] + vel[1] * nrm[1] + vel[2] * nrm[2];
      int rule = ghost_rule(r, veln);
      a.inj[gg] = (rule == kInject);
      if (rule == kCopy) {
        store_cell<NC>(P, N, gg, in);
        break;
      }
      if (rule == kMirror) {
#pragma unroll
        for (int v = 0; v < NC; ++v) gh[v] = in[v];
        for (int m = 0; m < 3; ++m) vel[m] = vel[m] - 2.0 * veln * nrm[m];
        gh[1] = vel[0];
        gh[2] = vel[1];
        gh[3] = vel[2];
        if constexpr (M == 2) mirror6(&in[4], nrm, &gh[4]);
      } else {
#pragma unroll
        for (int v = 0; v < NC; ++v) gh[v] = 0.0;
        double sg = (double)(2 * (r.f % 2) - 1);
        double inward[3] = {sg * nrm[0], sg * nrm[1], sg * nrm[2]};
        double spd;
        if (r.type == 401) {
          double vg = a.gas[1 * N + gm] * inward[0] + a.gas[2 * N + gm] * inward[1] + a.gas[3 * N + gm] * inward[2];
          spd = r.velocity * vg;
        } else if (r.type == 402) {
          spd = r.velocity;
        } else {
          spd = r.velocity * norm3(a.gas[1 * N + gm], a.gas[2 * N + gm], a.gas[3 * N + gm]);
        }
        if (r.normal_dir) {
          gh[1] = spd * inward[0];
          gh[2] = spd * inward[1];
          gh[3] = spd * inward[2];
        } else {
          gh[1] = spd * cos(r.beta) * cos(r.alpha);
          gh[2] = spd * cos(r.beta) * sin(r.alpha);
          gh[3] = spd * sin(r.beta);
        }
        if (r.type == 401) {
          gh[0] = r.massflux * a.gas[gm] / r.velocity;
          double vn = fabs(gh[1] * nrm[0] + gh[2] * nrm[1] + gh[3] * nrm[2]);
          if (r.gfixed > 0.0 && vn > 0.0) gh[0] = gh[0] + r.massflux * r.gfixed / vn;
          gh[NC - 2] = r.temperature * a.gas[4 * N + gm];
        } else {
          double vn = gh[1] * nrm[0] + gh[2] * nrm[1] + gh[3] * nrm[2];
          gh[0] = r.massflux / fabs(vn);
          gh[NC - 2] = r.temperature;
        }
        gh[NC - 1] = gh[0] / mat_rho(a.mat, gh[NC - 2]) / (4.0 / 3.0 * kPi * pow(r.radius, 3.0));
        if constexpr (M == 1) gh[4] = 1e-6;
        if constexpr (M == 2) {
          gh[4] = 1e-6;
          gh[7] = 1e-6;
          gh[9] = 1e-6;
        }
      }
      store_cell<NC>(P, N, gg, gh);
      break;
    }
    case 102:
      chimera<M>(a, r, 1);
      break;
    default:
      break;
  }
}

// second ghost layer (second_ghost_record, Lib_Ghost.f90:312-387)
template <int M>
__global__ void __launch_bounds__(kThreads) k_ghost2(FamArgs a) {
  constexpr int NC = Closure<M>::NC;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nlist) return;
  const DRec r = a.rec[a.list[t]];
  if (r.type == 0) return;
  const BlkDev B = a.blk[r.b - 1];
  const long long N = a.N;
  const int* G = kGuide[r.f - 1];
  long long g1 = gidx(B, r.i - G[0], r.j - G[1], r.k - G[2]);
  long long g2 = gidx(B, r.i - 2 * G[0], r.j - 2 * G[1], r.k - 2 * G[2]);
  long long gm = gidx(B, r.i, r.j, r.k);
  long long gp = gidx(B, r.i + G[0], r.j + G[1], r.k + G[2]);
  double* P = a.prim_w;
  double q[NC];
  switch (r.type) {
    case 101:
    case 201: {
      const BlkDev Bs = a.blk[r.bs - 1];
      const int* Gs = kGuide[r.fs - 1];
      load_cell<NC>(P, N, gidx(Bs, r.is + Gs[0], r.js + Gs[1], r.ks + Gs[2]), q);
      store_cell<NC>(P, N, g2, q);
      break;
    }
    case 102:
      chimera<M>(a, r, 2);
      break;
    case 200:
    case 300:
    case 301:
    case 401:
    case 402:
    case 403: {
      double nrm[3];
      bnd_normal(a, B, r, nrm);
      double veln = P[1 * N + gm] * nrm[0] + P[2 * N + gm] * nrm[1] + P[3 * N + gm] * nrm[2];
      int rule = ghost_rule(r, veln);
      if (rule == kInject) {
        load_cell<NC>(P, N, g1, q);
        store_cell<NC>(P, N, g2, q);
      } else {
        load_cell<NC>(P, N, gp, q);
        double o[NC];
#pragma unroll
        for (int v = 0; v < NC; ++v) o[v] = q[v];
        if (rule == kMirror) {
          double d = q[1] * nrm[0] + q[2] * nrm[1] + q[3] * nrm[2];
          for (int m = 0; m < 3; ++m) o[1 + m] = q[1 + m] - 2.0 * d * nrm[m];
          if constexpr (M == 2) mirror6(&q[4], nrm, &o[4]);
        }
        store_cell<NC>(P, N, g2, o);
      }
      break;
    }
    default: {
      double c1[NC], cm[NC];
      load_cell<NC>(P, N, g1, c1);
      load_cell<NC>(P, N, gm, cm);
      load_cell<NC>(P, N, gp, q);
#pragma unroll
      for (int v = 0; v < NC; ++v) q[v] = 3.0 * c1[v] - 3.0 * cm[v] + q[v];
      store_cell<NC>(P, N, g2, q);
      break;
    }
  }
}

// ------------------------------------------------------------------ beta
// tile_beta / SD_rho (Mod_Fluxes.f90:366-389)
template <int M>
__global__ void __launch_bounds__(kThreads) k_beta(FamArgs a, int b) {
  const BlkDev B = a.blk[b];
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.cells) return;
  int i, j, k;
  cell_of(B, t, i, j, k);
  long long g = gidx(B, i, j, k);
  long long sj = B.NI, sk = (long long)B.NI * B.NJ;
  const double* r = a.cur;
  double rc = r[g];
  // a flat direction's neighbours are the cell itself (see BlkDev::flat)
  double ri0 = (B.flat & 1) ? rc : r[g - 1], ri1 = (B.flat & 1) ? rc : r[g + 1];
  double rj0 = (B.flat & 2) ? rc : r[g - sj], rj1 = (B.flat & 2) ? rc : r[g + sj];
  double rk0 = (B.flat & 4) ? rc : r[g - sk], rk1 = (B.flat & 4) ? rc : r[g + sk];
  a.beta[g] = sd_weight(rc, ri0, ri1, rj0, rj1, rk0, rk1);
}

// ------------------------------------------------------------------ boundary fluxes
// record_flux (Mod_BC_Fluxes.f90:180-272).  The four stencil cells lie on one
// grid line: (g2, g1, m, m+1) on an odd face, (m-1, m, g1, g2) on an even one,
// so the reconstruction reads them from memory with the direction's stride.
template <int M, int S>
ICE_DEV bool record_flux(const FamArgs& a, const DRec& r, double beta_val, double* flux) {
  constexpr int NC = Closure<M>::NC;
  const BlkDev B = a.blk[r.b - 1];
  const long long N = a.N;
  int dir = (r.f + 1) / 2 - 1;
  long long s = dir == 0 ? 1ll : (dir == 1 ? (long long)B.NI : (long long)B.NI * (long long)B.NJ);
  bool odd = r.f % 2 == 1;
  long long gm = gidx(B, r.i, r.j, r.k);
  long long g1 = odd ? gm - s : gm + s;
  long long gp = odd ? gm + s : gm - s;
  double dl_g1 = a.dl[(long long)dir * N + g1];
  double dl_m = a.dl[(long long)dir * N + gm];
  double dl_4 = a.dl[(long long)dir * N + gp];
  bool inj_g = a.inj[g1], inj_p = a.inj[gp];
  const double* fa = a.fa + (long long)(dir * 4) * N;
  long long gf = odd ? g1 : gm;
  double nrm[3] = {fa[gf], fa[N + gf], fa[2 * N + gf]};
  double area = fa[3 * N + gf];
  double dl0, dl1, dl2, dll, dlr;
  if (odd) {
    dl0 = dl_g1; dl1 = 0.5 * (dl_g1 + dl_m); dl2 = 0.5 * (dl_m + dl_4);
    dll = 0.5 * dl_g1; dlr = 0.5 * dl_m;
    if (inj_g) { dl1 = 0.5 * dl_m; dll = 0.0; }
    if (inj_p) dl2 = 0.5 * dl_m;
  } else {
    dl0 = 0.5 * (dl_4 + dl_m); dl1 = 0.5 * (dl_m + dl_g1); dl2 = dl_g1;
    dll = 0.5 * dl_m; dlr = 0.5 * dl_g1;
    if (inj_g) { dl1 = 0.5 * dl_m; dlr = 0.0; }
    if (inj_p) dl0 = 0.5 * dl_m;
  }
  double ql[NC], qr[NC];
  bool ok = reconstruct_mem<M>(a.limiter, a.cur + (odd ? gm - 2 * s : gm - s), s, N, dl0, dl1, dl2, dll, dlr, beta_val,
                               ql, qr);
  riemann_acc<M, S, 2>(ql, qr, nrm, area, flux, a.mat);
  if (!odd) {
#pragma unroll
    for (int v = 0; v < NC; ++v) flux[v] = -flux[v];
  }
  return ok;
}

// owner_flux (Mod_BC_Fluxes.f90:151-170): one thread per owner record
template <int M, int S, bool SD>
__global__ void __launch_bounds__(kThreads) k_conn(FamArgs a) {
  constexpr int NC = Closure<M>::NC;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nlist) return;
  const DRec r = a.rec[a.list[t]];
  double bf = 1.0;
  if (SD) {
    const DRec q = a.rec[r.partner - 1];
    double bq = a.beta[gidx(a.blk[q.b - 1], q.i, q.j, q.k)];
    double bn = a.beta[gidx(a.blk[r.b - 1], r.i, r.j, r.k)];
    bf = dmin2(bn, bq);
  }
  double fl[NC];
  if (!record_flux<M, S>(a, r, bf, fl)) atomicOr(&a.st->bad_recon, 1);
#pragma unroll
  for (int v = 0; v < NC; ++v) a.connF[(long long)(r.cf - 1) * NC + v] = fl[v];
}

// bound_cell (Mod_BC_Fluxes.f90:105-144): one thread per boundary cell, its
// records in face order
template <int M, int S, bool SD>
__global__ void __launch_bounds__(kThreads) k_bcell(FamArgs a) {
  constexpr int NC = Closure<M>::NC;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nbcell) return;
  double acc[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) acc[v] = 0.0;
  bool ok = true;
  for (int rr = a.bfirst[c]; rr < a.bfirst[c + 1]; ++rr) {
    const DRec r = a.rec[a.brec[rr]];
    if (r.type == 200 && M == 0) continue;
    double fl[NC];
    if (r.type == 101 || r.type == 201) {
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        double x = a.connF[(long long)(r.cf - 1) * NC + v];
        fl[v] = r.owner ? x : -x;
      }
    } else {
      double bv = SD ? a.beta[gidx(a.blk[r.b - 1], r.i, r.j, r.k)] : 1.0;
      ok = record_flux<M, S>(a, r, bv, fl) && ok;
    }
#pragma unroll
    for (int v = 0; v < NC; ++v) acc[v] = acc[v] + fl[v];
  }
  if (!ok) atomicOr(&a.st->bad_recon, 1);
#pragma unroll
  for (int v = 0; v < NC; ++v) a.bsum[(long long)c * NC + v] = acc[v];
}

// The source handed to the coupler: the last stage's (keep_source 1), or the
// step's, each stage's weighted as the update weighs its residual (keep_source 2:
// add_stage_source / RK_accumulate, Mod_Explicit.f90:257-281, Lib_RK.f90:33-44);
// the relaxation time of the last stage either way
template <int NC> ICE_DEV void keep_stage_source(const FamArgs& a, long long g, const double* src, double tau) {
  if (a.last_stage) a.tau[g] = tau;
  if (a.keep_source == 2) {
    const double w = rk_weight(a.nrk, a.srk);
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      double* p = a.src + (long long)v * a.N + g;
      *p = a.srk == 1 ? src[v] : (*p + src[v]) * w;
    }
  } else if (a.last_stage && a.keep_source) {
    store_cell<NC>(a.src, a.N, g, src);
  }
}

// add_births (Lib_Residual.f90:164-210): the rates a coupler set, over the cell's dt
template <int NC> ICE_DEV void add_imports(const FamArgs& a, long long g, double dt, double* r) {
  if (!a.imports) return;
#pragma unroll
  for (int v = 0; v < NC; ++v) r[v] = r[v] + a.imports[(long long)v * a.N + g] * dt;
}

// add_birth (Mod_Exchange.f90:283-300): packets pk = [mass, momentum (3), energy,
// number] added at once to the conservative state of their cells in the current
// buffer; cell c (padded index cell[c]) takes packets first[c] .. first[c+1]-1 in
// their order, all by one thread
template <int M>
__global__ void __launch_bounds__(kThreads) k_add_births(FamArgs a, const long long* __restrict__ cell,
                                                         const int* __restrict__ first, const double* __restrict__ pk,
                                                         int ncell) {
  constexpr int NC = Closure<M>::NC;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const long long g = cell[c];
  double q[NC], u[NC];
  load_cell<NC>(a.cur, a.N, g, q);
  for (int n = first[c]; n < first[c + 1]; ++n) {
    const double* k = pk + (long long)n * 6;
    prim2cons<M>(q, u, a.mat);
    u[0] = u[0] + k[0];
    u[1] = u[1] + k[1];
    u[2] = u[2] + k[2];
    u[3] = u[3] + k[3];
    u[NC - 2] = u[NC - 2] + k[4];
    u[NC - 1] = u[NC - 1] + k[5];
    cons2prim<M>(u, q, a.mat);
  }
  store_cell<NC>(a.prim_w, a.N, g, q);
}

// ------------------------------------------------------------------ update
// state_update_ (Lib_Newstate.f90:152-173) of one cell
template <int M, bool TAB = true>
ICE_DEV void update_cell(const FamArgs& a, const BlkDev& B, int i, int j, int k, long long g, const double* R) {
  constexpr int NC = Closure<M>::NC;
  const auto& mt = mat_of<TAB>(a);
  double q[NC], co[NC], cu[NC];
  load_cell<NC>(a.old, a.N, g, q);
  prim2cons<M>(q, co, mt);
  load_cell<NC>(a.cur, a.N, g, q);
  prim2cons<M>(q, cu, mt);
  double w = rk_weight(a.nrk, a.srk);
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    double rr = (cu[v] - co[v] + R[v]) * w;
    cu[v] = co[v] + rr;
  }
  if (cu[0] < 0.0) atomicAdd(&a.st->n_floored, 1ull);
  cons2prim<M>(cu, q, mt);
  store_cell<NC>(a.nxt, a.N, g, q);
  if (!checkprim<M>(q)) {
    if (atomicCAS(&a.st->bad_state, 0, 1) == 0) {
      int bidx = (int)(&B - a.blk);
      a.st->bad_b = bidx + 1;
      a.st->bad_i = i;
      a.st->bad_j = j;
      a.st->bad_k = k;
      a.st->bad_p = a.fam + 1;
    }
  }
}

// The end of one cell's stage: the source (coupled, seeded by seed), then the
// residual from the flux sum, which r_in(r) hands over only after the source so
// that it is not held through the source (Lib_Residual.f90:147-156), the imports
// (add_births) and the frozen slots, then the RK update, or the residual for
// smoothing (WRITE_R).
template <int M, bool TAB, int CP, bool WRITE_R, bool LK, class RIn>
ICE_DEV void finish_cell_with(const FamArgs& a, const BlkDev& Bref, int ci, int cj, int ck, long long g, RIn r_in,
                              float seed) {
  constexpr int NC = Closure<M>::NC;
  double src[NC];
  if constexpr (CP != 0) {
    double q[NC], tau;
    load_cell<NC>(a.cur, a.N, g, q);
    cell_source<M, TAB, CP == 2, LK>(a, g, q, src, tau, seed);
    for (int f = 0; f < a.nfrozen; ++f) src[a.frozen[f] - 1] = 0.0;
    keep_stage_source<NC>(a, g, src, tau);
  } else {
#pragma unroll
    for (int v = 0; v < NC; ++v) src[v] = 0.0;
  }
  double r[NC];
  r_in(r);
  double vol = a.vol[g];
  double dt = a.time_accurate ? a.st->dtglobal : a.dt[g];
#pragma unroll
  for (int v = 0; v < NC; ++v) r[v] = (r[v] / vol + src[v]) * dt;
  add_imports<NC>(a, g, dt, r);
  for (int f = 0; f < a.nfrozen; ++f) r[a.frozen[f] - 1] = 0.0;
  if constexpr (WRITE_R) {
    store_cell<NC>(a.R, a.N, g, r);
  } else {
    update_cell<M, TAB>(a, Bref, ci, cj, ck, g, r);
  }
}

// finish_cell_with the flux sum r already in hand
template <int M, bool TAB, int CP, bool WRITE_R, bool LK = true>
ICE_DEV void finish_cell(const FamArgs& a, const BlkDev& Bref, int ci, int cj, int ck, long long g, const double* r,
                         float seed) {
  constexpr int NC = Closure<M>::NC;
  finish_cell_with<M, TAB, CP, WRITE_R, LK>(
      a, Bref, ci, cj, ck, g,
      [&](double* o) {
#pragma unroll
        for (int v = 0; v < NC; ++v) o[v] = r[v];
      },
      seed);
}

// Whether the cell's source runs the Langmuir-Knudsen interface solve (lk_needed)
template <int M, bool TAB>
ICE_DEV bool lk_cell_needed(const FamArgs& a, long long g) {
  constexpr int NC = Closure<M>::NC;
  const double Tp = a.cur[(long long)(NC - 2) * a.N + g];
  double ep[kNepCell];
  cell_ep<TAB>(a, g, Tp, ep);
  double p;
  return lk_needed(a.gas[g], a.gas[4 * a.N + g], a.gR[g], Tp, a.mat.evap, a.mat.intf, ep, p);
}

// Appends the wave's cells that need the interface solve to the family's queue (one
// atomic per wave), with their block and flux sum; false for the others.  The
// queue holds every interior cell of the family, so no cell can find it full;
// one that did would be reported (lk_overflow) as an error of the step.
template <int NC>
ICE_DEV bool defer_cell(const FamArgs& a, bool need, long long g, int b, const double* r) {
  const unsigned long long mask = __ballot(need);
  if (mask == 0ull) return false;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned base = 0u;
  if (lane == leader) base = atomicAdd(a.lkq_n, (unsigned)__popcll(mask));
  base = (unsigned)__shfl((int)base, leader, 64);
  if (!need) return false;
  const unsigned slot = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
  if (slot >= (unsigned)a.lkq_cap) {
    atomicOr(&a.st->lk_overflow, 1);
    return false;
  }
  a.lkq_g[slot] = g;
  a.lkq_b[slot] = b;
#pragma unroll
  for (int v = 0; v < NC; ++v) a.lkq_r[(long long)v * a.lkq_cap + slot] = r[v];
  return true;
}

// The deferred cells of a stage, in full waves: their source with the interface
// solve (seeded as on the tile), residual and update, in any order (each cell's
// result is its own).  A grid-stride loop over the queue; every wave of a workgroup
// runs the same trips.
template <int M, bool WRITE_R>
__global__ void __launch_bounds__(kThreads, StageWaves<M>::value) k_lk_finish(FamArgs a) {
  constexpr int NC = Closure<M>::NC;
  __shared__ __attribute__((aligned(16))) uint16_t mlds[kMlpLds];
  __shared__ __attribute__((aligned(16))) uint16_t mwts[kMlpWts];
  __shared__ float mbias[kMlpBias];
  // the workgroup's waves run many network passes: its weights are read from LDS
  const MlpDev net =
      mlp_in_lds(a.mlp, a.mat.surrogate == 1 && a.mlp_on && a.mat.evap != 0 && a.mat.intf == 1, mwts, mbias);
  const unsigned n0 = *a.lkq_n, n = n0 < (unsigned)a.lkq_cap ? n0 : (unsigned)a.lkq_cap;
  for (unsigned base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const unsigned idx = base + threadIdx.x;
    const bool has = idx < n;
    const long long g = has ? a.lkq_g[idx] : 0ll;
    const float seed = stage_seed<M, false>(a, has, g, mlds, net);
    if (has) {
      const BlkDev& B = a.blk[a.lkq_b[idx]];
      const long long loc = g - B.off;
      const int i = (int)(loc % B.NI) - 1, j = (int)((loc / B.NI) % B.NJ) - 1, k = (int)(loc / ((long long)B.NI * B.NJ)) - 1;
      // the queued flux sum is read after the solve, which then has its registers
      finish_cell_with<M, false, 2, WRITE_R, true>(
          a, B, i, j, k, g,
          [&](double* o) {
#pragma unroll
            for (int v = 0; v < NC; ++v) o[v] = a.lkq_r[(long long)v * a.lkq_cap + idx];
          },
          seed);
    }
  }
}

// The cell's stage: boundary sum, interior faces in the reference's order
// (Mod_Fluxes.f90:471-508), source, residual (Lib_Residual.f90:147-156) and
// either the RK update (WRITE_R = false) or the residual for smoothing.
template <int M, int S, bool SD, int CP, bool WRITE_R>
__global__ void __launch_bounds__(kThreads, StageWaves<M>::value) k_stage(FamArgs a, int b) {
  constexpr bool COUPLED = CP != 0;
  constexpr int NC = Closure<M>::NC;
  const BlkDev& B = a.blk[b];
  const BlkDev Bv = B;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i = 1, j = 1, k = 1;
  if (t < Bv.cells) cell_of(Bv, t, i, j, k);
  long long g = gidx(Bv, i, j, k);
  float seed = 0.0f;
  if constexpr (CP == 2) {
    __shared__ __attribute__((aligned(16))) uint16_t mlds[kMlpLds];
    seed = stage_seed<M>(a, t < Bv.cells, g, mlds, a.mlp);
  }
  if (t >= Bv.cells) return;
  double r[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) r[v] = 0.0;
  bool surface = i == 1 || i == Bv.ni || j == 1 || j == Bv.nj || k == 1 || k == Bv.nk;
  if (surface) {
    int c = a.bidx[g];
    if (c >= 0) {
#pragma unroll
      for (int v = 0; v < NC; ++v) r[v] = a.bsum[(long long)c * NC + v];
    }
  }
  bool ok = true;
  const int idx[3] = {i, j, k};
  const int nd[3] = {Bv.ni, Bv.nj, Bv.nk};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int n = nd[d], m = idx[d];
    if (n < 2) continue;
    long long s = dir_stride(Bv, d);
    if (m % 2 == 1) {
      if (m < n) ok = face_acc<M, S, SD, 1>(a, g, s, d, m, n, r) && ok;
      if (m > 1) ok = face_acc<M, S, SD, 0>(a, g - s, s, d, m - 1, n, r) && ok;
    } else {
      ok = face_acc<M, S, SD, 0>(a, g - s, s, d, m - 1, n, r) && ok;
      if (m < n) ok = face_acc<M, S, SD, 1>(a, g, s, d, m, n, r) && ok;
    }
  }
  if (!ok) atomicOr(&a.st->bad_recon, 1);
  double src[NC];
  if constexpr (COUPLED) {
    double q[NC], tau;
    load_cell<NC>(a.cur, a.N, g, q);
    cell_source<M, true, CP == 2>(a, g, q, src, tau, seed);
    for (int f = 0; f < a.nfrozen; ++f) src[a.frozen[f] - 1] = 0.0;
    keep_stage_source<NC>(a, g, src, tau);
  } else {
#pragma unroll
    for (int v = 0; v < NC; ++v) src[v] = 0.0;
  }
  double vol = a.vol[g];
  double dt = a.time_accurate ? a.st->dtglobal : a.dt[g];
#pragma unroll
  for (int v = 0; v < NC; ++v) r[v] = (r[v] / vol + src[v]) * dt;
  add_imports<NC>(a, g, dt, r);
  for (int f = 0; f < a.nfrozen; ++f) r[a.frozen[f] - 1] = 0.0;
  if constexpr (WRITE_R) {
    store_cell<NC>(a.R, a.N, g, r);
  } else {
    update_cell<M>(a, B, i, j, k, g, r);
  }
}

// ------------------------------------------------------------------ IRS
// residual_smoothing_ (Lib_IRS.f90:34-86), direction d, over every block of the
// family in one launch (workgroups wg[b] .. wg[b+1]-1 take block b).  The
// reference's two Jacobi sweeps are fused: the first sweep's values at m-1 and
// m+1 are formed in registers by the same expression the first sweep applies
// there, so the second sweep reads exactly the doubles it would read from T.
// Ghosts are the zero-gradient copies of the residual the direction started
// from (src).  The result goes to dst, a buffer other than src.
// the block of workgroup w of an all-blocks launch (workgroups wg[b] .. wg[b+1]-1 take block b)
ICE_DEV int block_of_workgroup(const int* __restrict__ wg, int nb, int w) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (wg[mid] <= w) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The value direction d leaves at cell g (index m of n along d), for every slot:
// the first sweep's values at m-1 and m+1 (ghost copies at the line's ends, where
// the value is the cell's own), then the second sweep.  The five cells g-2s ..
// g+2s are all inside the block's padded box (two ghost layers), so their loads are
// issued unconditionally and the ends select.
template <int NC>
ICE_DEV void irs_line(const FamArgs& a, const double* __restrict__ src, long long g, long long s, int m, int n,
                      double* out) {
  const double beta = a.irs_beta;
  const double den = 1.0 + 2.0 * beta;
  const SharedDiv byd = div_prepare(den);  // every quotient of the smoothing divides by den
  double x[NC][5];
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const double* R = src + (long long)v * a.N + g;
#pragma unroll
    for (int q = 0; q < 5; ++q) x[v][q] = R[(q - 2) * s];
  }
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const double r0 = x[v][2], rm = x[v][1], rp = x[v][3];
    const double rmm = m - 1 > 1 ? x[v][0] : rm;
    const double rpp = m + 1 < n ? x[v][4] : rp;
    const double tm = m > 1 ? div_shared(rm + beta * (rmm + r0), byd) : r0;
    const double tp = m < n ? div_shared(rp + beta * (r0 + rpp), byd) : r0;
    out[v] = div_shared(r0 + beta * (tm + tp), byd);
  }
}

template <int NC>
__global__ void __launch_bounds__(kThreads) k_irs(FamArgs a, int d, const double* __restrict__ src,
                                                  double* __restrict__ dst, const int* __restrict__ wg, int nb) {
  const int w = (int)blockIdx.x;
  const int b = block_of_workgroup(wg, nb, w);
  const BlkDev B = a.blk[b];
  long long t = (long long)(w - wg[b]) * blockDim.x + threadIdx.x;
  if (t >= B.cells) return;
  int i, j, k;
  cell_of(B, t, i, j, k);
  long long g = gidx(B, i, j, k);
  int m = d == 0 ? i : (d == 1 ? j : k);
  int n = d == 0 ? B.ni : (d == 1 ? B.nj : B.nk);
  double r[NC];
  irs_line<NC>(a, src, g, dir_stride(B, d), m, n, r);
#pragma unroll
  for (int v = 0; v < NC; ++v) dst[(long long)v * a.N + g] = r[v];
}

// The value the smoothing along i leaves at cell g (index i of ni along i) for one
// slot Rv: irs_line's expressions for that slot, bit for bit.  g-2 .. g+2 are all
// inside the block's padded box, so the loads are unconditional and the ends select.
ICE_DEV double irs_i_value(const double* __restrict__ Rv, long long g, int i, int ni, double beta,
                            const SharedDiv& byd) {
  const double x0 = Rv[g - 2], x1 = Rv[g - 1], x2 = Rv[g], x3 = Rv[g + 1], x4 = Rv[g + 2];
  const double rmm = i - 1 > 1 ? x0 : x1;
  const double rpp = i + 1 < ni ? x4 : x3;
  const double tm = i > 1 ? div_shared(x1 + beta * (rmm + x2), byd) : x2;
  const double tp = i < ni ? div_shared(x3 + beta * (x2 + rpp), byd) : x2;
  return div_shared(x2 + beta * (tm + tp), byd);
}

// Direction d = 1 (j) or 2 (k) of the smoothing, marching: one thread per line and
// slot (blockIdx.y), lanes along i, down the line.  Each residual is read once and
// each first-sweep value T(j) formed once, by irs_line's expressions with its end
// copies (R(0) = R(1), R(n+1) = R(n); T(0) = R(1), T(n+1) = R(n)), so the result is
// irs_line's bit for bit.  FOLD_I: src holds the residual before the smoothing
// along i, and each value the march reads is that smoothing's at the cell
// (irs_i_value), formed from the line's five neighbours along i, so the
// direction-i pass is never written and read back.  Workgroups wl[b] .. wl[b+1]-1
// take block b's lines.
template <int NC, bool FOLD_I = false>
__global__ void __launch_bounds__(kThreads) k_irs_march(FamArgs a, int d, const double* __restrict__ src,
                                                        double* __restrict__ dst, const int* __restrict__ wl, int nb) {
  const int w = (int)blockIdx.x;
  const int b = block_of_workgroup(wl, nb, w);
  const BlkDev B = a.blk[b];
  const int no = d == 1 ? B.nk : B.nj;
  const long long t = (long long)(w - wl[b]) * blockDim.x + threadIdx.x;
  if (t >= (long long)B.ni * no) return;
  const int i = (int)(t % B.ni) + 1, o = (int)(t / B.ni) + 1;
  const int n = d == 1 ? B.nj : B.nk;
  const long long s = dir_stride(B, d);
  long long g = d == 1 ? gidx(B, i, 1, o) : gidx(B, i, o, 1);
  const int v = (int)blockIdx.y;
  const double* __restrict__ R = src + (long long)v * a.N;
  double* __restrict__ D = dst + (long long)v * a.N;
  const double beta = a.irs_beta;
  const double den = 1.0 + 2.0 * beta;
  const SharedDiv byd = div_prepare(den);  // every quotient of the smoothing divides by den
  const int ni = B.ni;
  auto rd = [&](long long gg) { return FOLD_I ? irs_i_value(R, gg, i, ni, beta, byd) : R[gg]; };
  double r0 = rd(g);                          // R(m)
  double rp = n > 1 ? rd(g + s) : r0;         // R(m+1)
  double tm = r0;                             // T(m-1)
  double tc = div_shared(r0 + beta * (r0 + rp), byd);  // T(m)
  int m = 1;
  // interior, eight cells at a time with their loads issued together: R(m+2) ..
  // R(m+9) are cells of the line
  for (; m + 9 <= n; m += 8) {
    double x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = rd(g + (2 + q) * s);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double tp = div_shared(rp + beta * (r0 + x[q]), byd);
      D[g] = div_shared(r0 + beta * (tm + tp), byd);
      tm = tc; tc = tp; r0 = rp; rp = x[q]; g += s;
    }
  }
  // the rest of the interior: R(m+2) is a cell of the line
  for (; m + 2 <= n; ++m) {
    const double rpp = rd(g + 2 * s);
    const double tp = div_shared(rp + beta * (r0 + rpp), byd);
    D[g] = div_shared(r0 + beta * (tm + tp), byd);
    tm = tc; tc = tp; r0 = rp; rp = rpp; g += s;
  }
  if (m + 1 == n) {  // m = n-1: R(n+1) = R(n)
    const double tp = div_shared(rp + beta * (r0 + rp), byd);
    D[g] = div_shared(r0 + beta * (tm + tp), byd);
    tm = tc; r0 = rp; g += s; ++m;
  }
  // m = n: T(n+1) = R(n)
  D[g] = div_shared(r0 + beta * (tm + r0), byd);
}

// The last direction of the smoothing (d = 1 or 2) fused with the RK update: one
// thread per line (lanes along i) and every slot, down the line.  Each cell's
// smoothed residual is formed by k_irs_march's expressions, bit for bit, and goes
// straight into update_cell, so it is never written and read back.  FOLD_I (d =
// 1): src holds the residual before the smoothing along i, formed here as in
// k_irs_march.  Workgroups wl[b] .. wl[b+1]-1 take block b's lines, as in
// k_irs_march.
template <int M, bool FOLD_I = false>
__global__ void __launch_bounds__(kThreads, 2) k_irs_march_update(FamArgs a, int d, const double* __restrict__ src,
                                                                  const int* __restrict__ wl, int nb) {
  constexpr int NC = Closure<M>::NC;
  const int w = (int)blockIdx.x;
  const int b = block_of_workgroup(wl, nb, w);
  const BlkDev& B = a.blk[b];
  const int ni = B.ni, no = d == 1 ? B.nk : B.nj;
  const long long t = (long long)(w - wl[b]) * blockDim.x + threadIdx.x;
  if (t >= (long long)ni * no) return;
  const int i = (int)(t % ni) + 1, o = (int)(t / ni) + 1;
  const int n = d == 1 ? B.nj : B.nk;
  const long long s = dir_stride(B, d);
  long long g = d == 1 ? gidx(B, i, 1, o) : gidx(B, i, o, 1);
  const long long N = a.N;
  const double beta = a.irs_beta;
  const double den = 1.0 + 2.0 * beta;
  // irs_i_value divides through byd; the march's own quotients are plain
  // divisions (the shared-divisor form, inlined for every slot of every case,
  // measured slower in this kernel)
  const SharedDiv byd = div_prepare(den);
  auto rd = [&](int v, long long gg) {
    return FOLD_I ? irs_i_value(src + (long long)v * N, gg, i, ni, beta, byd) : src[(long long)v * N + gg];
  };
  // per slot: R(m), R(m+1), T(m-1), T(m) of the cell m being finished
  double r0[NC], rp[NC], tm[NC], tc[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    r0[v] = rd(v, g);
    rp[v] = n > 1 ? rd(v, g + s) : r0[v];
    tm[v] = r0[v];
    tc[v] = (r0[v] + beta * (r0[v] + rp[v])) / den;
  }
  for (int m = 1; m <= n; ++m) {
    double out[NC];
    if (m + 2 <= n) {
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        const double rpp = rd(v, g + 2 * s);
        const double tp = (rp[v] + beta * (r0[v] + rpp)) / den;
        out[v] = (r0[v] + beta * (tm[v] + tp)) / den;
        tm[v] = tc[v];
        tc[v] = tp;
        r0[v] = rp[v];
        rp[v] = rpp;
      }
    } else if (m + 1 == n) {  // R(n+1) = R(n)
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        const double tp = (rp[v] + beta * (r0[v] + rp[v])) / den;
        out[v] = (r0[v] + beta * (tm[v] + tp)) / den;
        tm[v] = tc[v];
        r0[v] = rp[v];
      }
    } else {  // m = n: T(n+1) = R(n)
#pragma unroll
      for (int v = 0; v < NC; ++v) out[v] = (r0[v] + beta * (tm[v] + r0[v])) / den;
    }
    const int j = d == 1 ? m : o, k = d == 1 ? o : m;
    update_cell<M>(a, B, i, j, k, g, out);
    g += s;
  }
}

// the dt pass of every block of the family in one launch (workgroups wg[b] ..
// wg[b+1]-1 take block b's cells, one per thread)
template <int M>
__global__ void __launch_bounds__(kThreads) k_dt_blocks(FamArgs a, const int* __restrict__ wg, int nb) {
  const int w = (int)blockIdx.x;
  const int b = block_of_workgroup(wg, nb, w);
  const BlkDev B = a.blk[b];
  long long t = (long long)(w - wg[b]) * blockDim.x + threadIdx.x;
  double dtc = 1.79769313486231571e308;
  if (t < B.cells) {
    int i, j, k;
    cell_of(B, t, i, j, k);
    dtc = dt_cell<M>(a, gidx(B, i, j, k));
  }
  block_min_atomic(dtc, &a.st->dtbits);
}

// the update of every block of the family from the smoothed residual res, in one
// launch (workgroups wg[b] .. wg[b+1]-1 take block b, as in k_irs)
template <int M>
__global__ void __launch_bounds__(kThreads) k_update(FamArgs a, const double* __restrict__ res,
                                                     const int* __restrict__ wg, int nb) {
  constexpr int NC = Closure<M>::NC;
  const int w = (int)blockIdx.x;
  const int b = block_of_workgroup(wg, nb, w);
  const BlkDev& B = a.blk[b];
  long long t = (long long)(w - wg[b]) * blockDim.x + threadIdx.x;
  if (t >= B.cells) return;
  int i, j, k;
  cell_of(B, t, i, j, k);
  long long g = gidx(B, i, j, k);
  double r[NC];
  load_cell<NC>(res, a.N, g, r);
  update_cell<M>(a, B, i, j, k, g, r);
}

}  // namespace ice

namespace ice {

// Tile of a block handled by one workgroup: TX x TY x TZ cells; the faces of
// the tile in one direction are computed once each, by one thread, into LDS.
struct Tile {
  int tx, ty, tz;     // cells per tile
  int ntx, nty, ntz;  // tiles per block
};

// slots of the face fluxes of one direction in LDS, structure of arrays
constexpr int kMaxTileFaces = kThreads;
// cells of a direction pass's extended box: the tile widened by one cell on each
// side along the direction (every tile ice_gpu_set_tiles accepts fits)
#ifndef ICE_MAX_EXT_CELLS
#define ICE_MAX_EXT_CELLS 352
#endif
constexpr int kMaxExtCells = ICE_MAX_EXT_CELLS;
// LDS of the stage tile, in doubles: the face fluxes of one direction (NC slots
// of kMaxTileFaces faces), the limited slopes of the extended box (NC slots of
// kMaxExtCells cells), the cells' half lengths and their first-order flags (bytes)
template <int NC> constexpr int tile_lds_doubles() {
  return NC * kMaxTileFaces + (NC + 1) * kMaxExtCells + kMaxExtCells / 8;
}

// The limited slope of every slot of one cell along direction d (padded index g,
// index m of n along d, neighbours g -/+ s).  The MUSCL reconstruction
// (Lib_Reconstruction.f90:17-78) of the face above the cell (its left state) calls
// the limiter with (slope to the upper neighbour, slope to the lower neighbour,
// cx, cy) over the mean spacings and the cell's half length, and that of the face
// below it (its right state) with the same five values, so each cell's value is
// formed once, with the same operations, and both faces read it.  The ghost
// lengths an inflow face zeroes (Mod_Fluxes.f90:613-618) are those of the
// neighbour past the first or the last cell.  first: the cell or a neighbour is
// empty, and the cell's side of either face is first order.
template <int M, int LIM>
ICE_DEV void cell_slopes(const FamArgs& a, long long g, long long s, int d, int m, int n, double* slope, int e,
                         double& half, bool& first) {
  constexpr int NC = Closure<M>::NC;
  int lim = a.limiter;
  if constexpr (LIM >= 0) lim = LIM;
  const long long N = a.N;
  // every load of the cell (its three cells' slots, lengths and inflow flags) is
  // issued before the first-order test, so they make one round trip to memory
  // instead of two; the empty asm keeps the compiler from moving them below the
  // test's branch
  const double* p = a.cur;
  double prev[NC], local[NC], next[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const double* pv = p + (long long)v * N;
    prev[v] = pv[g - s];
    local[v] = pv[g];
    next[v] = pv[g + s];
  }
  const double* len = a.dl + (long long)d * N;
  double len0 = len[g - s], len1 = len[g], len2 = len[g + s];
  const bool inj0 = a.inj[g - s], inj2 = a.inj[g + s];
  if (m == 1 && inj0) len0 = 0.0;
  if (m == n && inj2) len2 = 0.0;
  const double dl0 = 0.5 * (len0 + len1), dl1 = 0.5 * (len1 + len2);
  half = 0.5 * len1;
#pragma unroll
  for (int v = 0; v < NC; ++v) asm volatile("" ::"v"(prev[v]), "v"(local[v]), "v"(next[v]));
  asm volatile("" ::"v"(dl0), "v"(dl1));
  first = prev[0] <= kRhoEmpty || local[0] <= kRhoEmpty || next[0] <= kRhoEmpty;
  if (first) return;
  double cx, cy;
  if (half > 0.0) { cx = dl1 / half; cy = dl0 / half; } else { cx = 2.0; cy = 2.0; }
  const SharedDiv by0 = div_prepare(dl0), by1 = div_prepare(dl1);
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const double s_up = div_shared(next[v] - local[v], by1);
    const double s_dn = div_shared(local[v] - prev[v], by0);
    slope[v * kMaxExtCells + e] = limiter(lim, s_up, s_dn, cx, cy);
  }
}

// One interior face of a tile from its two cells' limited slopes (cell_slopes):
// the left and right states with the reference's halving of the limiter factor
// until both are realizable, or first order below 1e-18 (Lib_Reconstruction.f90:
// 53-76), then the Riemann flux times the area (Mod_Fluxes.f90:655-685), into q.
// g: the lower cell, s: the stride along d; elo/ehi: the cells' extended-box index.
template <int M, int S, bool SD, bool TAB>
ICE_DEV bool face_from_slopes(const FamArgs& a, long long g, long long s, int d, const double* slope,
                              const double* half, const uint8_t* first, int elo, int ehi, double* q) {
  constexpr int NC = Closure<M>::NC;
  const long long N = a.N;
  // the face's geometry and both cells' values are requested first, in one round
  // trip to memory, ahead of the LDS reads and the branches that use them
  const double* fa = a.fa + (long long)(d * 4) * N;
  const double nrm[3] = {fa[g], fa[N + g], fa[2 * N + g]};
  const double area = fa[3 * N + g];
  double lo[NC], hi[NC];
  load_cell<NC>(a.cur, N, g, lo);
  load_cell<NC>(a.cur, N, g + s, hi);
  const bool first_l = first[elo], first_r = first[ehi];
  const double dll = half[elo], dlr = half[ehi];
  const double bf = SD ? dmin2(a.beta[g], a.beta[g + s]) : 1.0;
  double ql[NC], qr[NC];
  bool ok = true;
  if (first_l && first_r) {
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      ql[v] = lo[v];
      qr[v] = hi[v];
    }
  } else {
    double limval = bf;
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      ql[v] = first_l ? lo[v] : lo[v] + slope[v * kMaxExtCells + elo] * limval * dll;
      qr[v] = first_r ? hi[v] : hi[v] - slope[v * kMaxExtCells + ehi] * limval * dlr;
    }
    bool ok_l = first_l || checkprim<M>(ql), ok_r = first_r || checkprim<M>(qr);
    if (__builtin_expect(!(ok_l && ok_r), 0)) {
      // the reference's retries: every pass forms both sides from the halved factor
      for (;;) {
        limval = 0.5 * limval;
        if (limval < 1.e-18) {
#pragma unroll
          for (int v = 0; v < NC; ++v) {
            ql[v] = lo[v];
            qr[v] = hi[v];
          }
          ok = checkprim<M>(ql) && checkprim<M>(qr);
          break;
        }
#pragma unroll
        for (int v = 0; v < NC; ++v) {
          if (!first_l) ql[v] = lo[v] + slope[v * kMaxExtCells + elo] * limval * dll;
          if (!first_r) qr[v] = hi[v] - slope[v * kMaxExtCells + ehi] * limval * dlr;
        }
        if (!first_l) ok_l = checkprim<M>(ql);
        if (!first_r) ok_r = checkprim<M>(qr);
        if (ok_l && ok_r) break;
      }
    }
  }
  riemann_acc<M, S, 2>(ql, qr, nrm, area, q, mat_of<TAB>(a));
  return ok;
}

// The stage with the interior faces shared through LDS (Mod_Fluxes.f90:421-521:
// each face once, each cell its two faces per direction in the odd-then-even
// order), then the cell's source, residual and update as in k_stage.
// One tile of the stage: the body of k_stage_tiled, also run by the persistent
// kernel for the tiles it takes from its work queue.  lds is the workgroup's
// buffer of tile_lds_doubles<NC>() doubles.  Each direction pass: the limited
// slopes of the extended box's cells, then each face once, then each cell sums
// its two faces; two barriers a pass.
// TAB = false: the launch serves only families without tabulated properties
template <int M, int S, bool SD, int CP, bool WRITE_R, int LIM, bool TAB = false>
ICE_DEV void stage_tile_body(const FamArgs& a, int b, const Tile& T, int tile, double* lds) {
  constexpr int NC = Closure<M>::NC;
  double* fp = lds;
  double* slope = lds + NC * kMaxTileFaces;
  double* half = slope + NC * kMaxExtCells;
  uint8_t* first = (uint8_t*)(half + kMaxExtCells);
  const BlkDev& Bref = a.blk[b];
  const BlkDev B = Bref;
  int ox = (tile % T.ntx) * T.tx + 1;
  int oy = ((tile / T.ntx) % T.nty) * T.ty + 1;
  int oz = (tile / (T.ntx * T.nty)) * T.tz + 1;
  int t = threadIdx.x;
  int ncell = T.tx * T.ty * T.tz;
  int ci = ox + t % T.tx, cj = oy + (t / T.tx) % T.ty, ck = oz + t / (T.tx * T.ty);
  bool has_cell = t < ncell && ci <= B.ni && cj <= B.nj && ck <= B.nk;
  long long g = has_cell ? gidx(B, ci, cj, ck) : 0;
  double r[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) r[v] = 0.0;
  if (has_cell && (ci == 1 || ci == B.ni || cj == 1 || cj == B.nj || ck == 1 || ck == B.nk)) {
    int c = a.bidx[g];
    if (c >= 0) {
#pragma unroll
      for (int v = 0; v < NC; ++v) r[v] = a.bsum[(long long)c * NC + v];
    }
  }
  bool ok = true;
  const int nd[3] = {B.ni, B.nj, B.nk};
  for (int d = 0; d < 3; ++d) {
    int n = nd[d];
    if (n < 2) continue;
    const long long stride = dir_stride(B, d);
    // the extended box: offsets -1 .. t_d along d, the tile's extent across it
    const int ex = T.tx + 2 * (d == 0), ey = T.ty + 2 * (d == 1), ez = T.tz + 2 * (d == 2);
    const int ne = ex * ey * ez;
    const int bx = ox - (d == 0), by = oy - (d == 1), bz = oz - (d == 2);
    for (int e = t; e < ne; e += blockDim.x) {
      const int li = bx + e % ex, lj = by + (e / ex) % ey, lk = bz + e / (ex * ey);
      const int m = d == 0 ? li : (d == 1 ? lj : lk);
      if (m >= 1 && m <= n && li <= B.ni && lj <= B.nj && lk <= B.nk) {
        double h;
        bool f1;
        cell_slopes<M, LIM>(a, gidx(B, li, lj, lk), stride, d, m, n, slope, e, h, f1);
        half[e] = h;
        first[e] = f1;
      }
    }
    __syncthreads();
    // faces of the tile along d: lower-cell offset -1 .. t_d-1 along d; the face's
    // lower cell has the same coordinates in the extended box, its upper the next
    int fx = T.tx + (d == 0), fy = T.ty + (d == 1), fz = T.tz + (d == 2);
    int nf = fx * fy * fz;
    const int estep = d == 0 ? 1 : (d == 1 ? ex : ex * ey);
    for (int f = t; f < nf; f += blockDim.x) {
      const int fi = f % fx, fj = (f / fx) % fy, fk = f / (fx * fy);
      int li = bx + fi, lj = by + fj, lk = bz + fk;
      int m = d == 0 ? li : (d == 1 ? lj : lk);
      bool inside = m >= 1 && m <= n - 1 && li >= 1 && lj >= 1 && lk >= 1 && li <= B.ni && lj <= B.nj && lk <= B.nk;
      if (inside) {
        const int elo = fi + ex * (fj + ey * fk);
        double q[NC];
        ok = face_from_slopes<M, S, SD, TAB>(a, gidx(B, li, lj, lk), stride, d, slope, half, first, elo,
                                             elo + estep, q) && ok;
#pragma unroll
        for (int v = 0; v < NC; ++v) fp[v * kMaxTileFaces + f] = q[v];
      }
    }
    __syncthreads();
    if (has_cell) {
      int ti = ci - ox, tj = cj - oy, tk = ck - oz;
      // face index of the cell's lower face in this direction's face grid
      int flo = ti + fx * (tj + fy * tk);
      int fup = flo + (d == 0 ? 1 : (d == 1 ? fx : fx * fy));
      int m = d == 0 ? ci : (d == 1 ? cj : ck);
      if (m % 2 == 1) {
        if (m < n) {
#pragma unroll
          for (int v = 0; v < NC; ++v) r[v] = r[v] - fp[v * kMaxTileFaces + fup];
        }
        if (m > 1) {
#pragma unroll
          for (int v = 0; v < NC; ++v) r[v] = r[v] + fp[v * kMaxTileFaces + flo];
        }
      } else {
#pragma unroll
        for (int v = 0; v < NC; ++v) r[v] = r[v] + fp[v * kMaxTileFaces + flo];
        if (m < n) {
#pragma unroll
          for (int v = 0; v < NC; ++v) r[v] = r[v] - fp[v * kMaxTileFaces + fup];
        }
      }
    }
  }
  if (!ok) atomicOr(&a.st->bad_recon, 1);
  // a cell whose source runs the Langmuir-Knudsen solve finishes in k_lk_finish,
  // where such cells fill whole waves instead of holding up the rest of theirs;
  // the solve is not compiled into the tile (a family with the solve always has
  // the queue)
  bool deferred = false;
  if constexpr (CP == 2) {
    if (a.lkq_g) {
      const bool need = has_cell && lk_cell_needed<M, TAB>(a, g);
      deferred = defer_cell<NC>(a, need, g, b, r);
    }
  }
  if (!has_cell || deferred) return;
  finish_cell<M, TAB, CP, WRITE_R, false>(a, Bref, ci, cj, ck, g, r, 0.0f);
}

// Workgroup w of a launch runs on XCD w % 8 (workgroups are dealt round-robin
// over the 8 XCDs): give each XCD one contiguous run of the n tiles, so that
// tiles sharing halo cells share that XCD's L2.
ICE_DEV int xcd_tile(int w, int n) {
#ifdef ICE_NO_XCD_MAP
  return w;
#endif
  constexpr int kXcd = 8;
  int q = n / kXcd, r = n % kXcd, x = w % kXcd, j = w / kXcd;
  return x * q + (x < r ? x : r) + j;
}

template <int M, int S, bool SD, int CP, bool WRITE_R, int LIM>
__global__ void __launch_bounds__(kThreads, StageWaves<M>::value) k_stage_tiled(FamArgs a, int b, Tile T) {
  constexpr int NC = Closure<M>::NC;
  __shared__ double lds[tile_lds_doubles<NC>()];
  stage_tile_body<M, S, SD, CP, WRITE_R, LIM>(a, b, T, xcd_tile(blockIdx.x, gridDim.x), lds);
}

// ------------------------------------------------------------------ batched path
// A unit of work of the persistent kernels: one tile of one block of one family.
struct WorkItem {
  int fam, block, tile, pad;
  Tile T;
};

// Work queues, one per XCD: workgroup w serves queue w % nq (the hardware places
// consecutive workgroups on consecutive XCDs, so a queue's tiles, which are
// contiguous in the grid, share that XCD's L2), taking items in order until the
// queue is empty.  Every wave reaches the exit: a workgroup leaves at its first
// ticket past the end of its queue.
struct WorkQueues {
  const WorkItem* items;
  const int* qstart;               // nq + 1 offsets into items
  unsigned long long* ticket;      // nq counters, zeroed before each launch
  int nq;
};

ICE_DEV int next_item(const WorkQueues& Q, int q, int* shared_slot) {
  if (threadIdx.x == 0) {
    unsigned long long t = atomicAdd(&Q.ticket[q], 1ull);
    long long it = (long long)Q.qstart[q] + (long long)t;
    *shared_slot = it < Q.qstart[q + 1] ? (int)it : -1;
  }
  __syncthreads();
  // the item is the workgroup's: saying so keeps its fields, and the family's
  // arguments they index, in scalar registers
  int it = __builtin_amdgcn_readfirstlane(*shared_slot);
  __syncthreads();
  return it;
}

template <int M, int S, bool SD, int CP, bool WRITE_R, int LIM>
__global__ void __launch_bounds__(kThreads, StageWaves<M>::value) k_stage_persist(const FamArgs* __restrict__ fams,
                                                                                   WorkQueues Q) {
  constexpr int NC = Closure<M>::NC;
  __shared__ double lds[tile_lds_doubles<NC>()];
  __shared__ int slot;
  int q = blockIdx.x % Q.nq;
  for (;;) {
    int it = next_item(Q, q, &slot);
    if (it < 0) break;
    const WorkItem w = Q.items[it];
    stage_tile_body<M, S, SD, CP, WRITE_R, LIM>(fams[w.fam], w.block, w.T, w.tile, lds);
  }
}

// The batched stage with one workgroup per work item instead of persistent
// workgroups: the hardware dispatcher balances the items, dealt to the XCDs in
// contiguous runs (an item's tile neighbours and its sibling families share
// that XCD's L2).  Same tiles, same bits as k_stage_persist.
template <int M, int S, bool SD, int CP, bool WRITE_R, int LIM>
__global__ void __launch_bounds__(kThreads, StageWaves<M>::value) k_stage_items(const FamArgs* __restrict__ fams,
                                                                                 const WorkItem* __restrict__ items) {
  constexpr int NC = Closure<M>::NC;
  __shared__ double lds[tile_lds_doubles<NC>()];
  const int it = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x));
  const WorkItem w = items[it];
  stage_tile_body<M, S, SD, CP, WRITE_R, LIM>(fams[w.fam], w.block, w.T, w.tile, lds);
}

// the sensor weights over the same queues, one cell per thread
template <int M>
__global__ void __launch_bounds__(kThreads) k_beta_persist(const FamArgs* fams, WorkQueues Q) {
  __shared__ int slot;
  int q = blockIdx.x % Q.nq;
  for (;;) {
    int it = next_item(Q, q, &slot);
    if (it < 0) break;
    const WorkItem w = Q.items[it];
    const FamArgs& a = fams[w.fam];
    const BlkDev B = a.blk[w.block];
    const Tile& T = w.T;
    int t = threadIdx.x;
    int ox = (w.tile % T.ntx) * T.tx + 1, oy = ((w.tile / T.ntx) % T.nty) * T.ty + 1,
        oz = (w.tile / (T.ntx * T.nty)) * T.tz + 1;
    int ci = ox + t % T.tx, cj = oy + (t / T.tx) % T.ty, ck = oz + t / (T.tx * T.ty);
    if (t < T.tx * T.ty * T.tz && ci <= B.ni && cj <= B.nj && ck <= B.nk) {
      long long g = gidx(B, ci, cj, ck), sj = B.NI, sk = (long long)B.NI * B.NJ;
      const double* r = a.cur;
      double rc = r[g];
      double ri0 = (B.flat & 1) ? rc : r[g - 1], ri1 = (B.flat & 1) ? rc : r[g + 1];
      double rj0 = (B.flat & 2) ? rc : r[g - sj], rj1 = (B.flat & 2) ? rc : r[g + sj];
      double rk0 = (B.flat & 4) ? rc : r[g - sk], rk1 = (B.flat & 4) ? rc : r[g + sk];
      a.beta[g] = sd_weight(rc, ri0, ri1, rj0, rj1, rk0, rk1);
    }
  }
}


namespace {  // one copy per translation unit
__global__ void k_zero_tickets(unsigned long long* t, int n) {
  if (threadIdx.x < n) t[threadIdx.x] = 0ull;
}
__global__ void k_zero_count(unsigned* n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *n = 0u;
}
}  // namespace

}  // namespace ice
