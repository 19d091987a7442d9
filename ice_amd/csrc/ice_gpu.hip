// C-ABI of the MI355X explicit residual-and-update path (include/ice_gpu.h):
// the context, its set-up, and the step driver that replaces the body of
// Explicit_Step (src/lib/numerics/time/explicit/Mod_Explicit.f90:11-245).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/ice_gpu.h"
#include "batched.hpp"
#include "launch.hpp"

using namespace ice;

namespace {

struct FamDev {
  double* buf[3] = {nullptr, nullptr, nullptr};  // three state buffers, roles rotated per stage
  int cur = 0, old = 0;
  double *beta = nullptr, *dt = nullptr, *tau = nullptr, *src = nullptr, *R = nullptr, *T = nullptr;
  double *bsum = nullptr, *connF = nullptr;
  uint8_t* inj = nullptr;
  int* bidx = nullptr;
  int *grp = nullptr, *own = nullptr, *bfirst = nullptr, *brec = nullptr;
  long long* bcell_g = nullptr;
  int ngrp = 0, nown = 0, nbcell = 0, nconn = 0;
  int nc = 0;
};

// host view of a record and its pairing (ICE_bc_type)
struct HRec {
  int v[ICE_REC_NI];
  double d[ICE_REC_ND];
  int partner = 0, owner = 0, cf = 0;
};

}  // namespace

struct ice_gpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  ice_config_t cfg;
  std::vector<ice_family_t> fams;
  std::vector<BlkDev> blk;
  BlkDev* d_blk = nullptr;
  long long N = 0;
  double *vol = nullptr, *dl = nullptr, *M = nullptr, *fa = nullptr;
  double *gas = nullptr, *gR = nullptr, *gG = nullptr, *gK = nullptr, *gMu = nullptr, *yinf = nullptr;
  std::vector<FamDev> fam;
  std::vector<HRec> rec;
  DRec* d_rec = nullptr;
  int* d_donor = nullptr;
  double* d_weight = nullptr;
  std::vector<std::vector<int>> bfirst_h, brec_h;  // per family, 1-based (collect_boundary_cells)
  DevStatus* d_st = nullptr;
  int iter = 0;
  int keep_source = 0;
  int kernel_timing = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double step_ms = 0.0, kernel_ms = 0.0;
  long long kernel_launches = 0;
  std::vector<hipEvent_t> kev;
  BatchedPlan* batched = nullptr;
  int tile[3] = {-1, -1, -1};  // cells per stage tile; -1 default shape, 0 per-cell kernel
  std::string err;
};

namespace {

int fail(ice_gpu_ctx* c, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return 1;
}

#define HIPCHK(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return fail(c, "%s: %s", #call, hipGetErrorString(e_));         \
  } while (0)

template <typename T> int dalloc(ice_gpu_ctx* c, T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(c, "hipMalloc of %zu bytes: %s", n * sizeof(T), hipGetErrorString(e));
  return 0;
}

int nc_of(int closure) { return closure == ICE_MK ? 6 : (closure == ICE_IG ? 7 : 12); }

long long padded_of(const BlkDev& B, int i, int j, int k) {
  return B.off + (long long)(i + 1) + (long long)B.NI * ((long long)(j + 1) + (long long)B.NJ * (long long)(k + 1));
}

// A fill kernel for NaN / zero initialisation, so no memset node is ever needed
__global__ void k_fill(double* p, long long n, double v) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}
__global__ void k_fill_i32(int* p, long long n, int v) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}
__global__ void k_fill_u8(uint8_t* p, long long n, uint8_t v) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

unsigned grid_of(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

int fill(ice_gpu_ctx* c, double* p, long long n, double v) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill, dim3(grid_of(n)), dim3(kThreads), 0, c->stream, p, n, v);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- set-up
// Pair_Connections (IO_BC.f90:360-510): partner, owner and face number of every
// connection record, family by family, faces numbered by their owners in table order.
int pair(ice_gpu_ctx* c) {
  int ng = c->cfg.ngroups, nrec = (int)c->rec.size(), nb = (int)c->blk.size();
  auto tangent = [](int f, int i, int j, int k, int& a, int& b) {
    if (f <= 2) { a = j; b = k; }
    else if (f <= 4) { a = i; b = k; }
    else { a = i; b = j; }
  };
  for (int p = 1; p <= ng; ++p) {
    std::vector<std::vector<int>> face(nb * 6);
    for (int n = p; n <= nrec; n += ng) {
      const int* v = c->rec[n - 1].v;
      if (v[ICE_R_TYPE] != 101 && v[ICE_R_TYPE] != 201) continue;
      const BlkDev& B = c->blk[v[ICE_R_B] - 1];
      int n1, n2, a, b;
      tangent(v[ICE_R_F], B.ni, B.nj, B.nk, n1, n2);
      auto& m = face[(v[ICE_R_B] - 1) * 6 + v[ICE_R_F] - 1];
      if (m.empty()) m.assign((size_t)n1 * n2, 0);
      tangent(v[ICE_R_F], v[ICE_R_I], v[ICE_R_J], v[ICE_R_K], a, b);
      m[(size_t)(a - 1) + (size_t)n1 * (b - 1)] = n;
    }
    for (int n = p; n <= nrec; n += ng) {
      HRec& R = c->rec[n - 1];
      const int* v = R.v;
      if (v[ICE_R_TYPE] != 101 && v[ICE_R_TYPE] != 201) continue;
      int bs = v[ICE_R_BS], fs = v[ICE_R_FS], q = 0;
      if (bs >= 1 && bs <= nb && fs >= 1 && fs <= 6 && !face[(bs - 1) * 6 + fs - 1].empty()) {
        const BlkDev& B = c->blk[bs - 1];
        int n1, n2, a, b;
        tangent(fs, B.ni, B.nj, B.nk, n1, n2);
        tangent(fs, v[ICE_R_IS], v[ICE_R_JS], v[ICE_R_KS], a, b);
        if (a >= 1 && b >= 1 && a <= n1 && b <= n2) q = face[(bs - 1) * 6 + fs - 1][(size_t)(a - 1) + (size_t)n1 * (b - 1)];
      }
      if (q) {
        const int* w = c->rec[q - 1].v;
        if (w[ICE_R_I] != v[ICE_R_IS] || w[ICE_R_J] != v[ICE_R_JS] || w[ICE_R_K] != v[ICE_R_KS]) q = 0;
      }
      if (q == 0 || q == n)
        return fail(c, "[ICE::Setup_BC] the connection record %d (block %d, cell %d %d %d, face %d) has no partner", n,
                    v[ICE_R_B], v[ICE_R_I], v[ICE_R_J], v[ICE_R_K], v[ICE_R_F]);
      const int* w = c->rec[q - 1].v;
      if (w[ICE_R_TYPE] != v[ICE_R_TYPE] || w[ICE_R_BS] != v[ICE_R_B] || w[ICE_R_FS] != v[ICE_R_F] ||
          w[ICE_R_IS] != v[ICE_R_I] || w[ICE_R_JS] != v[ICE_R_J] || w[ICE_R_KS] != v[ICE_R_K])
        return fail(c, "[ICE::Setup_BC] the connection records %d and %d do not point at each other", n, q);
      const BlkDev& Bs = c->blk[bs - 1];
      int across = (fs + 1) / 2 == 1 ? Bs.ni : ((fs + 1) / 2 == 2 ? Bs.nj : Bs.nk);
      if (across < ICE_GC) return fail(c, "[ICE::Setup_BC] record %d: block %d has fewer than gc cells across", n, bs);
      R.partner = q;
      int f = v[ICE_R_F], g = w[ICE_R_F];
      if (f % 2 == 0 && g % 2 == 1) R.owner = 1;
      else if (f % 2 == 1 && g % 2 == 0) R.owner = 0;
      else {
        int k1[5] = {v[ICE_R_B], f, v[ICE_R_K], v[ICE_R_J], v[ICE_R_I]};
        int k2[5] = {w[ICE_R_B], g, w[ICE_R_K], w[ICE_R_J], w[ICE_R_I]};
        int less = 0;
        for (int m = 0; m < 5; ++m)
          if (k1[m] != k2[m]) { less = k1[m] < k2[m]; break; }
        R.owner = less;
      }
    }
    int nf = 0;
    for (int n = p; n <= nrec; n += ng) {
      HRec& R = c->rec[n - 1];
      if ((R.v[ICE_R_TYPE] != 101 && R.v[ICE_R_TYPE] != 201) || !R.owner) continue;
      R.cf = ++nf;
      c->rec[R.partner - 1].cf = nf;
    }
    c->fam[p - 1].nconn = nf;
  }
  return 0;
}

// build_local_bc_index / collect_boundary_cells / collect_owner_records
// (Mod_GhostExchange.f90:103-360) for one process owning every block
int build_lists(ice_gpu_ctx* c) {
  int ng = c->cfg.ngroups, nrec = (int)c->rec.size();
  c->bfirst_h.assign(ng, {});
  c->brec_h.assign(ng, {});
  for (int p = 1; p <= ng; ++p) {
    FamDev& F = c->fam[p - 1];
    std::vector<int> grp, own, bfirst, brec;
    std::vector<long long> bg;
    for (int n = 1; n <= nrec; ++n) {
      const int* v = c->rec[n - 1].v;
      if (v[ICE_R_P] != p) continue;
      // a planar record of a flat direction fills ghosts nothing reads (BlkDev::flat)
      if (v[ICE_R_TYPE] == 0 && (c->blk[v[ICE_R_B] - 1].flat >> ((v[ICE_R_F] - 1) / 2)) & 1) continue;
      grp.push_back(n - 1);
    }
    for (int n0 : grp) {
      const HRec& R = c->rec[n0];
      if ((R.v[ICE_R_TYPE] == 101 || R.v[ICE_R_TYPE] == 201) && R.owner) own.push_back(n0);
    }
    // records of each cell of each block, cells in k, j, i order, records by face (stable)
    for (size_t b = 0; b < c->blk.size(); ++b) {
      const BlkDev& B = c->blk[b];
      std::vector<std::pair<long long, int>> at;  // (cell key, record)
      for (int n0 : grp) {
        const int* v = c->rec[n0].v;
        if (v[ICE_R_B] != (int)b + 1 || v[ICE_R_TYPE] == 0) continue;
        long long key = (long long)(v[ICE_R_I] - 1) + (long long)B.ni * ((v[ICE_R_J] - 1) + (long long)B.nj * (v[ICE_R_K] - 1));
        at.push_back({key, n0});
      }
      std::stable_sort(at.begin(), at.end(), [&](const std::pair<long long, int>& x, const std::pair<long long, int>& y) {
        if (x.first != y.first) return x.first < y.first;
        return c->rec[x.second].v[ICE_R_F] < c->rec[y.second].v[ICE_R_F];
      });
      for (size_t m = 0; m < at.size(); ++m) {
        if (m == 0 || at[m].first != at[m - 1].first) {
          bfirst.push_back((int)brec.size());
          const int* v = c->rec[at[m].second].v;
          bg.push_back(padded_of(B, v[ICE_R_I], v[ICE_R_J], v[ICE_R_K]));
        }
        brec.push_back(at[m].second);
      }
    }
    bfirst.push_back((int)brec.size());
    F.ngrp = (int)grp.size();
    F.nown = (int)own.size();
    F.nbcell = (int)bg.size();
    c->bfirst_h[p - 1] = bfirst;
    c->brec_h[p - 1] = brec;
    if (dalloc(c, &F.grp, grp.size()) || dalloc(c, &F.own, own.size()) || dalloc(c, &F.bfirst, bfirst.size()) ||
        dalloc(c, &F.brec, brec.size()) || dalloc(c, &F.bcell_g, bg.size()) ||
        dalloc(c, &F.bsum, (size_t)F.nbcell * F.nc) || dalloc(c, &F.connF, (size_t)std::max(F.nconn, 1) * F.nc))
      return 1;
    HIPCHK(hipMemcpy(F.grp, grp.data(), grp.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(F.own, own.data(), own.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(F.bfirst, bfirst.data(), bfirst.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(F.brec, brec.data(), brec.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(F.bcell_g, bg.data(), bg.size() * sizeof(long long), hipMemcpyHostToDevice));
    // boundary cell of each surface cell
    std::vector<int> bidx(c->N, -1);
    for (size_t m = 0; m < bg.size(); ++m) bidx[bg[m]] = (int)m;
    HIPCHK(hipMemcpy(F.bidx, bidx.data(), c->N * sizeof(int), hipMemcpyHostToDevice));
  }
  return 0;
}

FamArgs fam_args(ice_gpu_ctx* c, int p) {
  FamDev& F = c->fam[p];
  const ice_family_t& fd = c->fams[p];
  FamArgs a;
  std::memset(&a, 0, sizeof a);
  a.blk = c->d_blk;
  a.N = c->N;
  a.vol = c->vol; a.dl = c->dl; a.M = c->M; a.fa = c->fa;
  a.gas = c->gas; a.gR = c->gR; a.gG = c->gG; a.gK = c->gK; a.gMu = c->gMu; a.yinf = c->yinf;
  a.cur = F.buf[F.cur]; a.old = F.buf[F.old];
  a.prim_w = F.buf[F.cur];
  a.beta = F.beta; a.dt = F.dt; a.tau = F.tau; a.inj = F.inj; a.src = F.src; a.R = F.R; a.T = F.T;
  a.bsum = F.bsum; a.bidx = F.bidx; a.connF = F.connF;
  a.rec = c->d_rec; a.donor = c->d_donor; a.weight = c->d_weight;
  a.bfirst = F.bfirst; a.brec = F.brec; a.bcell_g = F.bcell_g; a.nbcell = F.nbcell;
  a.st = c->d_st;
  a.mat.rho = fd.rho_mat; a.mat.cs = fd.cs; a.mat.lv = fd.lv; a.mat.emiss = fd.emiss;
  for (int m = 0; m < 8; ++m) a.mat.ep[m] = fd.ep[m];
  a.mat.evap = fd.evap_select; a.mat.intf = fd.intf_select; a.mat.surrogate = fd.surrogate;
  a.limiter = c->cfg.limiter; a.nrk = c->cfg.nrk; a.time_accurate = c->cfg.time_accurate;
  a.coupled = c->cfg.coupled; a.drag_sel = c->cfg.drag_select; a.heat_sel = c->cfg.heat_select;
  a.blow_sel = c->cfg.blow_select;
  a.nfrozen = fd.nfrozen;
  for (int m = 0; m < 4; ++m) a.frozen[m] = fd.frozen[m];
  a.keep_source = c->keep_source; a.var_yinf = c->cfg.var_yinf; a.iter = c->iter;
  a.cfl_ramp_iter = c->cfg.cfl_ramp_iter; a.ndir = c->cfg.ndir; a.fam = p;
  a.cfl = c->cfg.cfl; a.dt_max = c->cfg.dt_max; a.tau_factor = c->cfg.tau_factor; a.irs_beta = c->cfg.irs_beta;
  return a;
}

// ---------------------------------------------------------------- dispatch
__global__ void k_step_begin(DevStatus* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) st->dtbits = (unsigned long long)__double_as_longlong(1e+5);
}

__global__ void k_dt_end(DevStatus* st, int time_accurate) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double dt = __longlong_as_double((long long)st->dtbits);
    st->dtglobal = dt;
    if (time_accurate) st->time = st->time + dt;
  }
}

void launch_dt(ice_gpu_ctx* c, int p, const FamArgs& a) {
  int cl = c->fams[p].closure, nb = (int)c->blk.size();
  if (cl == ICE_MK) launch_dt_mk(c->stream, c->blk.data(), nb, a);
  else if (cl == ICE_IG) launch_dt_ig(c->stream, c->blk.data(), nb, a);
  else launch_dt_ag(c->stream, c->blk.data(), nb, a);
}

// The stage tile of a block: the context's shape, else one whose face grids in
// every resolved direction fit the 256 threads of a workgroup.
Tile stage_tile(ice_gpu_ctx* c, const BlkDev& B) {
  Tile T;
  if (c->tile[0] == 0) {
    T.tx = T.ty = T.tz = 0;
    T.ntx = T.nty = T.ntz = 0;
    return T;
  }
  int tx, ty, tz;
  if (c->tile[0] > 0) {
    tx = c->tile[0]; ty = c->tile[1]; tz = c->tile[2];
  } else if (B.nj == 1 && B.nk == 1) {
    tx = 255; ty = 1; tz = 1;
  } else if (B.nk == 1) {
    tx = 15; ty = 16; tz = 1;
  } else {
    tx = 7; ty = 6; tz = 5;
  }
  tx = std::min(tx, B.ni); ty = std::min(ty, B.nj); tz = std::min(tz, B.nk);
  T.tx = tx; T.ty = ty; T.tz = tz;
  T.ntx = (B.ni + tx - 1) / tx; T.nty = (B.nj + ty - 1) / ty; T.ntz = (B.nk + tz - 1) / tz;
  return T;
}

void dispatch_stage(ice_gpu_ctx* c, FamArgs& a) {
  const FamDev& F = c->fam[a.fam];
  std::vector<Tile> tiles(c->blk.size());
  for (size_t b = 0; b < c->blk.size(); ++b) tiles[b] = stage_tile(c, c->blk[b]);
  StageLaunch L;
  L.stream = c->stream;
  L.blk = c->blk.data();
  L.tiles = tiles.data();
  L.nb = (int)c->blk.size();
  L.grp = F.grp; L.ngrp = F.ngrp; L.own = F.own; L.nown = F.nown; L.nbcell = F.nbcell;
  L.riemann = c->fams[a.fam].riemann; L.sd = c->cfg.shock_detector; L.coupled = c->cfg.coupled;
  L.irs = c->cfg.irs; L.limiter = c->cfg.limiter; L.ndir = c->cfg.ndir;
  L.timing = c->kernel_timing; L.kev = &c->kev;
  int cl = c->fams[a.fam].closure;
  if (cl == ICE_MK) launch_stage_mk(L, a);
  else if (cl == ICE_IG) launch_stage_ig(L, a);
  else launch_stage_ag(L, a);
}

// the roles of the three buffers at stage srk: read cur (and old), write nxt
int other_buffer(int x, int y) {
  for (int m = 0; m < 3; ++m)
    if (m != x && m != y) return m;
  return 0;
}

// One Explicit_Step, plain path: dt pass, then per family per stage the
// ghost, beta, boundary, flux/residual/update kernels.
int plain_step(ice_gpu_ctx* c) {
  int ng = c->cfg.ngroups;
  c->iter += 1;
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, c->stream, c->d_st);
  for (int p = 0; p < ng; ++p) {
    FamDev& F = c->fam[p];
    F.old = F.cur;  // state_copy: the step's start state stays in its buffer
    FamArgs a = fam_args(c, p);
    launch_dt(c, p, a);
  }
  hipLaunchKernelGGL(k_dt_end, dim3(1), dim3(64), 0, c->stream, c->d_st, c->cfg.time_accurate);
  for (int p = 0; p < ng; ++p) {
    FamDev& F = c->fam[p];
    for (int srk = 1; srk <= c->cfg.nrk; ++srk) {
      int nxt = other_buffer(F.cur, F.old);
      FamArgs a = fam_args(c, p);
      a.nxt = F.buf[nxt];
      a.srk = srk;
      a.last_stage = srk == c->cfg.nrk;
      dispatch_stage(c, a);
      F.cur = nxt;
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace

// ================================================================= C-ABI
extern "C" {

int ice_gpu_create(int device, const ice_config_t* cfg, const ice_family_t* fams, int nb, const int32_t* dims,
                   ice_gpu_ctx** out) {
  *out = nullptr;
  ice_gpu_ctx* c = new ice_gpu_ctx();
  c->device = device;
  c->cfg = *cfg;
  if (cfg->ngroups < 1 || nb < 1) {
    fail(c, "[ICE::create] ngroups and the block count must be positive");
    *out = c;
    return 1;
  }
  if (cfg->nrk < 1 || cfg->nrk > 3) {
    fail(c, "  [ERROR] unknown time-scheme: nrk must be 1, 2 or 3");
    *out = c;
    return 1;
  }
  if (cfg->irs && cfg->time_accurate) {
    fail(c, "[ICE::Assign_Setup] implicit residual smoothing is for steady runs (time-accurate = false)");
    *out = c;
    return 1;
  }
  c->fams.assign(fams, fams + cfg->ngroups);
  for (int p = 0; p < cfg->ngroups; ++p) {
    if (fams[p].riemann == ICE_SAUREL && fams[p].closure != ICE_MK) {
      fail(c, "[ICE::Check_Riemann_Solver] Saurel serves the MK closure only (family %d)", p + 1);
      *out = c;
      return 1;
    }
  }
  *out = c;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(c, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&c->ev0));
  HIPCHK(hipEventCreate(&c->ev1));
  long long off = 0;
  for (int b = 0; b < nb; ++b) {
    BlkDev B;
    std::memset(&B, 0, sizeof B);
    B.ni = dims[3 * b]; B.nj = dims[3 * b + 1]; B.nk = dims[3 * b + 2];
    if (B.ni < 1 || B.nj < 1 || B.nk < 1) return fail(c, "[ICE::create] block %d has an empty dimension", b + 1);
    B.NI = B.ni + 4; B.NJ = B.nj + 4; B.NK = B.nk + 4;
    B.off = off;
    B.cells = (long long)B.ni * B.nj * B.nk;
    off += (long long)B.NI * B.NJ * B.NK;
    c->blk.push_back(B);
  }
  c->N = off;
  long long N = c->N;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  if (dalloc(c, &c->d_blk, nb)) return 1;
  HIPCHK(hipMemcpy(c->d_blk, c->blk.data(), nb * sizeof(BlkDev), hipMemcpyHostToDevice));
  if (dalloc(c, &c->vol, N) || dalloc(c, &c->dl, 3 * N) || dalloc(c, &c->M, 9 * N) || dalloc(c, &c->fa, 12 * N))
    return 1;
  if (fill(c, c->vol, N, nan) || fill(c, c->dl, 3 * N, nan) || fill(c, c->M, 9 * N, nan) || fill(c, c->fa, 12 * N, nan))
    return 1;
  if (cfg->coupled) {
    if (dalloc(c, &c->gas, 5 * N) || dalloc(c, &c->gR, N) || dalloc(c, &c->gG, N) || dalloc(c, &c->gK, N) ||
        dalloc(c, &c->gMu, N))
      return 1;
    if (fill(c, c->gas, 5 * N, nan) || fill(c, c->gR, N, nan) || fill(c, c->gG, N, nan) || fill(c, c->gK, N, nan) ||
        fill(c, c->gMu, N, nan))
      return 1;
    if (cfg->var_yinf) {
      if (dalloc(c, &c->yinf, N) || fill(c, c->yinf, N, 0.0)) return 1;
    }
  }
  c->fam.resize(cfg->ngroups);
  for (int p = 0; p < cfg->ngroups; ++p) {
    FamDev& F = c->fam[p];
    F.nc = nc_of(fams[p].closure);
    for (int m = 0; m < 3; ++m) {
      if (dalloc(c, &F.buf[m], (size_t)F.nc * N) || fill(c, F.buf[m], (long long)F.nc * N, nan)) return 1;
    }
    if (dalloc(c, &F.beta, N) || fill(c, F.beta, N, 1.0)) return 1;
    if (dalloc(c, &F.tau, N) || fill(c, F.tau, N, std::numeric_limits<double>::infinity())) return 1;
    if (!cfg->time_accurate && (dalloc(c, &F.dt, N) || fill(c, F.dt, N, 0.0))) return 1;
    if (dalloc(c, &F.inj, N)) return 1;
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_of(N)), dim3(kThreads), 0, c->stream, F.inj, N, (uint8_t)0);
    if (dalloc(c, &F.bidx, N)) return 1;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_of(N)), dim3(kThreads), 0, c->stream, F.bidx, N, -1);
    if (cfg->irs) {
      if (dalloc(c, &F.R, (size_t)F.nc * N) || dalloc(c, &F.T, (size_t)F.nc * N)) return 1;
    }
  }
  if (dalloc(c, &c->d_st, 1)) return 1;
  DevStatus s0;
  std::memset(&s0, 0, sizeof s0);
  HIPCHK(hipMemcpy(c->d_st, &s0, sizeof s0, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int ice_gpu_set_geometry(ice_gpu_ctx* c, int b, const double* vol, const double* dl, const double* M, const double* fi,
                         const double* fj, const double* fk) {
  if (b < 1 || b > (int)c->blk.size()) return fail(c, "[ICE::set_geometry] no block %d", b);
  const BlkDev& B = c->blk[b - 1];
  long long npad = (long long)B.NI * B.NJ * B.NK, N = c->N;
  int ni = B.ni, nj = B.nj, nk = B.nk;
  // stage the block's padded slices on the host, then one copy per array
  std::vector<double> h(npad * 12, std::numeric_limits<double>::quiet_NaN());
  auto loc = [&](int i, int j, int k) { return (long long)(i + 1) + (long long)B.NI * ((j + 1) + (long long)B.NJ * (k + 1)); };
  auto put = [&](double* dev, int ncomp, const std::vector<double>& src) -> int {
    for (int m = 0; m < ncomp; ++m)
      HIPCHK(hipMemcpy(dev + (long long)m * N + B.off, src.data() + (long long)m * npad, npad * sizeof(double),
                       hipMemcpyHostToDevice));
    return 0;
  };
  for (int k = 1; k <= nk; ++k)
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) h[loc(i, j, k)] = vol[(i - 1) + (long long)ni * ((j - 1) + (long long)nj * (k - 1))];
  if (put(c->vol, 1, h)) return 1;
  std::fill(h.begin(), h.end(), std::numeric_limits<double>::quiet_NaN());
  for (int k = 0; k <= nk + 1; ++k)
    for (int j = 0; j <= nj + 1; ++j)
      for (int i = 0; i <= ni + 1; ++i)
        for (int m = 0; m < 3; ++m)
          h[m * npad + loc(i, j, k)] = dl[(i + (long long)(ni + 2) * (j + (long long)(nj + 2) * k)) * 3 + m];
  if (put(c->dl, 3, h)) return 1;
  std::fill(h.begin(), h.end(), std::numeric_limits<double>::quiet_NaN());
  for (int k = 1; k <= nk; ++k)
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) {
        long long s = (i - 1) + (long long)ni * ((j - 1) + (long long)nj * (k - 1));
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) h[(r * 3 + cc) * npad + loc(i, j, k)] = M[s * 9 + r + 3 * cc];
      }
  if (put(c->M, 9, h)) return 1;
  std::fill(h.begin(), h.end(), std::numeric_limits<double>::quiet_NaN());
  for (int k = 1; k <= nk; ++k)
    for (int j = 1; j <= nj; ++j)
      for (int i = 0; i <= ni; ++i)
        for (int m = 0; m < 4; ++m)
          h[(0 * 4 + m) * npad + loc(i, j, k)] = fi[(i + (long long)(ni + 1) * ((j - 1) + (long long)nj * (k - 1))) * 4 + m];
  for (int k = 1; k <= nk; ++k)
    for (int j = 0; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i)
        for (int m = 0; m < 4; ++m)
          h[(1 * 4 + m) * npad + loc(i, j, k)] = fj[((i - 1) + (long long)ni * (j + (long long)(nj + 1) * (k - 1))) * 4 + m];
  for (int k = 0; k <= nk; ++k)
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i)
        for (int m = 0; m < 4; ++m)
          h[(2 * 4 + m) * npad + loc(i, j, k)] = fk[((i - 1) + (long long)ni * ((j - 1) + (long long)nj * k)) * 4 + m];
  if (put(c->fa, 12, h)) return 1;
  return 0;
}

int ice_gpu_set_boundaries(ice_gpu_ctx* c, int nrec, const int32_t* reci, const double* recd, int ndonor,
                           const int32_t* donor, const double* weight) {
  int ng = c->cfg.ngroups;
  if (nrec % ng != 0) return fail(c, "[ICE::Setup_BC] bc array is not a multiple of ngroups");
  c->rec.assign(nrec, HRec());
  for (int n = 0; n < nrec; ++n) {
    std::memcpy(c->rec[n].v, reci + (size_t)n * ICE_REC_NI, sizeof(int) * ICE_REC_NI);
    std::memcpy(c->rec[n].d, recd + (size_t)n * ICE_REC_ND, sizeof(double) * ICE_REC_ND);
    const int* v = c->rec[n].v;
    int b = v[ICE_R_B], f = v[ICE_R_F], p = v[ICE_R_P], code = v[ICE_R_TYPE];
    if (b < 1 || b > (int)c->blk.size() || f < 1 || f > 6 || p < 1 || p > ng)
      return fail(c, "[ICE::Setup_BC] record %d names block %d, face %d, family %d", n + 1, b, f, p);
    const BlkDev& B = c->blk[b - 1];
    int i = v[ICE_R_I], j = v[ICE_R_J], k = v[ICE_R_K];
    int on = (f == 1 && i == 1) || (f == 2 && i == B.ni) || (f == 3 && j == 1) || (f == 4 && j == B.nj) ||
             (f == 5 && k == 1) || (f == 6 && k == B.nk);
    if (i < 1 || i > B.ni || j < 1 || j > B.nj || k < 1 || k > B.nk || !on)
      return fail(c, "[ICE::Setup_BC] record %d: cell %d %d %d is not on face %d of block %d", n + 1, i, j, k, f, b);
    if (code != 0 && code != 101 && code != 201 && code != 102 && code != 200 && code != 300 && code != 301 &&
        code != 400 && (code < 401 || code > 403))
      return fail(c, "[ICE::Setup_BC] unsupported ATLAS BC code %d at record %d", code, n + 1);
    if (code == 102) {
      int d0 = v[ICE_R_DONOR0], n1 = v[ICE_R_NCHI1], n2 = v[ICE_R_NCHI2];
      if (d0 < 0 || n1 < 0 || n2 < 0 || d0 + n1 + n2 > ndonor)
        return fail(c, "[ICE::Setup_BC] chimera record %d: donors out of the donor list", n + 1);
      for (int m = d0; m < d0 + n1 + n2; ++m) {
        const int* d = donor + 4 * (size_t)m;
        if (d[0] < 1 || d[0] > (int)c->blk.size()) return fail(c, "[ICE::Setup_BC] chimera donor in no block");
        const BlkDev& Bd = c->blk[d[0] - 1];
        if (d[1] < 1 || d[1] > Bd.ni || d[2] < 1 || d[2] > Bd.nj || d[3] < 1 || d[3] > Bd.nk)
          return fail(c, "[ICE::Setup_BC] chimera donor %d is not a cell of block %d", m + 1, d[0]);
      }
    }
  }
  if (pair(c)) return 1;
  std::vector<DRec> dr(nrec);
  for (int n = 0; n < nrec; ++n) {
    const HRec& H = c->rec[n];
    DRec& r = dr[n];
    std::memset(&r, 0, sizeof r);
    r.b = H.v[ICE_R_B]; r.i = H.v[ICE_R_I]; r.j = H.v[ICE_R_J]; r.k = H.v[ICE_R_K]; r.f = H.v[ICE_R_F];
    r.type = H.v[ICE_R_TYPE]; r.p = H.v[ICE_R_P]; r.bs = H.v[ICE_R_BS]; r.is = H.v[ICE_R_IS]; r.js = H.v[ICE_R_JS];
    r.ks = H.v[ICE_R_KS]; r.fs = H.v[ICE_R_FS]; r.normal_dir = H.v[ICE_R_NORMAL_DIR];
    r.nchi1 = H.v[ICE_R_NCHI1]; r.nchi2 = H.v[ICE_R_NCHI2]; r.donor0 = H.v[ICE_R_DONOR0];
    r.partner = H.partner; r.owner = H.owner; r.cf = H.cf;
    r.massflux = H.d[ICE_RD_MASSFLUX]; r.velocity = H.d[ICE_RD_VELOCITY]; r.temperature = H.d[ICE_RD_TEMPERATURE];
    r.radius = H.d[ICE_RD_RADIUS]; r.alpha = H.d[ICE_RD_ALPHA]; r.beta = H.d[ICE_RD_BETA]; r.gfixed = H.d[ICE_RD_GFIXED];
  }
  if (dalloc(c, &c->d_rec, nrec) || dalloc(c, &c->d_donor, (size_t)ndonor * 4) || dalloc(c, &c->d_weight, ndonor))
    return 1;
  HIPCHK(hipMemcpy(c->d_rec, dr.data(), nrec * sizeof(DRec), hipMemcpyHostToDevice));
  if (ndonor) {
    HIPCHK(hipMemcpy(c->d_donor, donor, (size_t)ndonor * 4 * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_weight, weight, (size_t)ndonor * sizeof(double), hipMemcpyHostToDevice));
  }
  // directions of one cell whose faces are all planar (0), in every family
  for (size_t b = 0; b < c->blk.size(); ++b) {
    BlkDev& B = c->blk[b];
    const int n[3] = {B.ni, B.nj, B.nk};
    B.flat = 0;
    for (int d = 0; d < 3; ++d) B.flat |= (n[d] == 1) << d;
  }
  for (const HRec& H : c->rec)
    if (H.v[ICE_R_TYPE] != 0) c->blk[H.v[ICE_R_B] - 1].flat &= ~(1 << ((H.v[ICE_R_F] - 1) / 2));
  HIPCHK(hipMemcpy(c->d_blk, c->blk.data(), c->blk.size() * sizeof(BlkDev), hipMemcpyHostToDevice));
  if (build_lists(c)) return 1;
  return 0;
}

int ice_gpu_get_bc_maps(ice_gpu_ctx* c, int32_t* partner, int32_t* owner, int32_t* cf, int p, int32_t* ncell,
                        int32_t* first, int32_t* recs, int cap) {
  for (size_t n = 0; n < c->rec.size(); ++n) {
    partner[n] = c->rec[n].partner;
    owner[n] = c->rec[n].owner;
    cf[n] = c->rec[n].cf;
  }
  if (p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::get_bc_maps] no family %d", p);
  const std::vector<int>& fh = c->bfirst_h[p - 1];
  const std::vector<int>& rh = c->brec_h[p - 1];
  *ncell = (int)fh.size() - 1;
  if ((int)fh.size() > cap || (int)rh.size() > cap) return fail(c, "[ICE::get_bc_maps] capacity %d too small", cap);
  for (size_t m = 0; m < fh.size(); ++m) first[m] = fh[m] + 1;
  for (size_t m = 0; m < rh.size(); ++m) recs[m] = rh[m] + 1;
  return 0;
}

static int state_copy_block(ice_gpu_ctx* c, int b, int p, const double* prim, double* out) {
  if (b < 1 || b > (int)c->blk.size() || p < 1 || p > c->cfg.ngroups)
    return fail(c, "[ICE::state] no block %d / family %d", b, p);
  const BlkDev& B = c->blk[b - 1];
  FamDev& F = c->fam[p - 1];
  long long npad = (long long)B.NI * B.NJ * B.NK;
  std::vector<double> h((size_t)npad * F.nc);
  if (prim) {
    // Fortran (nc, cells) -> SoA slices, into every buffer
    for (long long g = 0; g < npad; ++g)
      for (int v = 0; v < F.nc; ++v) h[(size_t)v * npad + g] = prim[g * F.nc + v];
    for (int m = 0; m < 3; ++m)
      for (int v = 0; v < F.nc; ++v)
        HIPCHK(hipMemcpy(F.buf[m] + (long long)v * c->N + B.off, h.data() + (size_t)v * npad, npad * sizeof(double),
                         hipMemcpyHostToDevice));
  } else {
    for (int v = 0; v < F.nc; ++v)
      HIPCHK(hipMemcpy(h.data() + (size_t)v * npad, F.buf[F.cur] + (long long)v * c->N + B.off, npad * sizeof(double),
                       hipMemcpyDeviceToHost));
    for (long long g = 0; g < npad; ++g)
      for (int v = 0; v < F.nc; ++v) out[g * F.nc + v] = h[(size_t)v * npad + g];
  }
  return 0;
}

int ice_gpu_put_state(ice_gpu_ctx* c, int b, int p, const double* prim) { return state_copy_block(c, b, p, prim, nullptr); }
int ice_gpu_get_state(ice_gpu_ctx* c, int b, int p, double* prim) {
  HIPCHK(hipStreamSynchronize(c->stream));
  return state_copy_block(c, b, p, nullptr, prim);
}

static int put_interior(ice_gpu_ctx* c, int b, double* dev, int ncomp, const double* src, bool comp_fastest) {
  const BlkDev& B = c->blk[b - 1];
  long long npad = (long long)B.NI * B.NJ * B.NK;
  std::vector<double> h((size_t)npad * ncomp, std::numeric_limits<double>::quiet_NaN());
  for (int k = 1; k <= B.nk; ++k)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i) {
        long long s = (i - 1) + (long long)B.ni * ((j - 1) + (long long)B.nj * (k - 1));
        long long g = (long long)(i + 1) + (long long)B.NI * ((j + 1) + (long long)B.NJ * (k + 1));
        for (int m = 0; m < ncomp; ++m)
          h[(size_t)m * npad + g] = comp_fastest ? src[s * ncomp + m] : src[(size_t)m * B.cells + s];
      }
  for (int m = 0; m < ncomp; ++m)
    HIPCHK(hipMemcpy(dev + (long long)m * c->N + B.off, h.data() + (size_t)m * npad, npad * sizeof(double),
                     hipMemcpyHostToDevice));
  return 0;
}

int ice_gpu_put_gas(ice_gpu_ctx* c, int b, const double* gprim, const double* R, const double* gam, const double* k,
                    const double* mu) {
  if (!c->cfg.coupled) return fail(c, "[ICE::put_gas] this context is not coupled to a gas");
  if (b < 1 || b > (int)c->blk.size()) return fail(c, "[ICE::put_gas] no block %d", b);
  if (put_interior(c, b, c->gas, 5, gprim, true) || put_interior(c, b, c->gR, 1, R, true) ||
      put_interior(c, b, c->gG, 1, gam, true) || put_interior(c, b, c->gK, 1, k, true) ||
      put_interior(c, b, c->gMu, 1, mu, true))
    return 1;
  return 0;
}

int ice_gpu_put_gas_yinf(ice_gpu_ctx* c, int b, const double* yinf) {
  if (!c->yinf) return fail(c, "[ICE::put_gas_yinf] the context was created without var_yinf");
  if (b < 1 || b > (int)c->blk.size()) return fail(c, "[ICE::put_gas_yinf] no block %d", b);
  return put_interior(c, b, c->yinf, 1, yinf, true);
}

int ice_gpu_put_tau(ice_gpu_ctx* c, int b, int p, const double* tau) {
  if (b < 1 || b > (int)c->blk.size() || p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::put_tau] no block/family");
  return put_interior(c, b, c->fam[p - 1].tau, 1, tau, true);
}

int ice_gpu_set_keep_source(ice_gpu_ctx* c, int on) {
  c->keep_source = on ? 1 : 0;
  for (auto& F : c->fam)
    if (on && !F.src) {
      if (dalloc(c, &F.src, (size_t)F.nc * c->N) || fill(c, F.src, (long long)F.nc * c->N, 0.0)) return 1;
    }
  return 0;
}

int ice_gpu_step(ice_gpu_ctx* c, int nsteps, ice_status_t* st) {
  if (c->rec.empty()) return fail(c, "[ICE::step] no boundary table: call ice_gpu_set_boundaries first");
  for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
  c->kev.clear();
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  for (int s = 0; s < nsteps; ++s) {
    int rc = (c->cfg.mode == ICE_MODE_BATCHED) ? batched_step(c) : plain_step(c);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->step_ms = ms;
  c->kernel_ms = 0.0;
  c->kernel_launches = 0;
  for (size_t m = 0; m + 1 < c->kev.size(); m += 2) {
    float km = 0.f;
    (void)hipEventElapsedTime(&km, c->kev[m], c->kev[m + 1]);
    c->kernel_ms += km;
    c->kernel_launches += 1;
  }
  DevStatus s;
  HIPCHK(hipMemcpy(&s, c->d_st, sizeof s, hipMemcpyDeviceToHost));
  if (st) {
    st->n_floored = (int64_t)s.n_floored;
    st->bad_state = s.bad_state;
    st->bad_block = s.bad_b; st->bad_i = s.bad_i; st->bad_j = s.bad_j; st->bad_k = s.bad_k; st->bad_p = s.bad_p;
    st->bad_recon = s.bad_recon;
    st->iter = c->iter;
    st->dtglobal = s.dtglobal;
    st->time = s.time;
  }
  if (s.bad_recon) return fail(c, " [ERROR] [ICE::compute_flux] unphysical state at first order on a face");
  if (s.bad_state)
    return fail(c,
                " [ERROR] [ICE::state_update] invalid state after the update (NaN or infinity): block %d cell %d %d %d "
                "phase %d",
                s.bad_b, s.bad_i, s.bad_j, s.bad_k, s.bad_p);
  return 0;
}

int ice_gpu_get_source(ice_gpu_ctx* c, int b, int p, double* source, double* tau) {
  if (b < 1 || b > (int)c->blk.size() || p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::get_source] no block/family");
  FamDev& F = c->fam[p - 1];
  if (!F.src) return fail(c, "[ICE::get_source] call ice_gpu_set_keep_source(ctx, 1) before the step");
  HIPCHK(hipStreamSynchronize(c->stream));
  const BlkDev& B = c->blk[b - 1];
  long long npad = (long long)B.NI * B.NJ * B.NK;
  std::vector<double> h((size_t)npad * (F.nc + 1));
  for (int v = 0; v < F.nc; ++v)
    HIPCHK(hipMemcpy(h.data() + (size_t)v * npad, F.src + (long long)v * c->N + B.off, npad * sizeof(double),
                     hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(h.data() + (size_t)F.nc * npad, F.tau + B.off, npad * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 1; k <= B.nk; ++k)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i) {
        long long s = (i - 1) + (long long)B.ni * ((j - 1) + (long long)B.nj * (k - 1));
        long long g = (long long)(i + 1) + (long long)B.NI * ((j + 1) + (long long)B.NJ * (k + 1));
        for (int v = 0; v < F.nc; ++v) source[s * F.nc + v] = h[(size_t)v * npad + g];
        tau[s] = h[(size_t)F.nc * npad + g];
      }
  return 0;
}

int ice_gpu_timing(ice_gpu_ctx* c, double* step_ms, double* kernel_ms, int64_t* launches) {
  if (step_ms) *step_ms = c->step_ms;
  if (kernel_ms) *kernel_ms = c->kernel_ms;
  if (launches) *launches = c->kernel_launches;
  return 0;
}

int ice_gpu_set_tiles(ice_gpu_ctx* c, int tx, int ty, int tz) {
  if (tx == 0) {
    c->tile[0] = c->tile[1] = c->tile[2] = 0;
    return 0;
  }
  if (tx < 0) {
    c->tile[0] = c->tile[1] = c->tile[2] = -1;
    return 0;
  }
  for (const BlkDev& B : c->blk) {
    int n[3] = {B.ni, B.nj, B.nk}, t[3] = {std::min(tx, B.ni), std::min(ty, B.nj), std::min(tz, B.nk)};
    bool bad = ty < 1 || tz < 1 || t[0] * t[1] * t[2] > kThreads;
    for (int d = 0; d < 3; ++d)
      if (n[d] > 1) bad = bad || t[0] * t[1] * t[2] / t[d] * (t[d] + 1) > kMaxTileFaces;
    if (bad)
      return fail(c, "[ICE::set_tiles] a %d x %d x %d tile has more cells or faces than the %d threads of a workgroup",
                  tx, ty, tz, kThreads);
  }
  c->tile[0] = tx; c->tile[1] = ty; c->tile[2] = tz;
  return 0;
}

int ice_gpu_set_kernel_timing(ice_gpu_ctx* c, int on) {
  c->kernel_timing = on ? 1 : 0;
  return 0;
}

int ice_gpu_last_error(ice_gpu_ctx* c, char* buf, int len) {
  if (!c || len <= 0) return 1;
  std::snprintf(buf, (size_t)len, "%s", c->err.c_str());
  return 0;
}

void ice_gpu_destroy(ice_gpu_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
  auto F = [](void* p) {
    if (p) (void)hipFree(p);
  };
  F(c->d_blk); F(c->vol); F(c->dl); F(c->M); F(c->fa); F(c->gas); F(c->gR); F(c->gG); F(c->gK); F(c->gMu);
  F(c->yinf); F(c->d_rec); F(c->d_donor); F(c->d_weight); F(c->d_st);
  for (auto& f : c->fam) {
    for (int m = 0; m < 3; ++m) F(f.buf[m]);
    F(f.beta); F(f.dt); F(f.tau); F(f.src); F(f.R); F(f.T); F(f.bsum); F(f.connF); F(f.inj); F(f.bidx);
    F(f.grp); F(f.own); F(f.bfirst); F(f.brec); F(f.bcell_g);
  }
  batched_destroy(c);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"

// Batched path: not yet built; the mode refuses to run rather than fall back.
int batched_step(ice_gpu_ctx* c) { return fail(c, "[ICE::step] the batched mode is not available in this build"); }
void batched_destroy(ice_gpu_ctx*) {}
