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
#include "inflow400.hpp"
#include "launch.hpp"
#include "multigrid.hpp"

using namespace ice;

namespace {

struct FamDev {
  double* buf[3] = {nullptr, nullptr, nullptr};  // three state buffers, roles rotated per stage
  int cur = 0, old = 0;
  double *beta = nullptr, *dt = nullptr, *tau = nullptr, *src = nullptr, *R = nullptr, *T = nullptr;
  double* imports = nullptr;  // rates in force a coupler set (add_births), [nc][N], or null
  // the queue of the cells whose interface solve a stage defers (Langmuir-Knudsen families)
  unsigned* lkq_n = nullptr;
  long long* lkq_g = nullptr;
  int* lkq_b = nullptr;
  double* lkq_r = nullptr;
  int lkq_cap = 0;
  double *bsum = nullptr, *connF = nullptr;
  uint8_t* inj = nullptr;
  int* bidx = nullptr;
  int *grp = nullptr, *own = nullptr, *bfirst = nullptr, *brec = nullptr;
  long long* bcell_g = nullptr;
  int ngrp = 0, nown = 0, nbcell = 0, nconn = 0;
  int nc = 0;
  double* prop = nullptr;   // the material's property table (kPropCols columns), or null
  int prop_tmin = 0, prop_tmax = 0, prop_flags = 0;
};

// The 400 records of a family and its boundaries' accumulators and totals
// (count_400_inflow); n = 0: the family has no 400 record
struct I400Fam {
  int* rec = nullptr;
  int n = 0;
  I400Chunk* chunks = nullptr;
  int nchunk = 0;
  I400Acc* acc = nullptr;
  I400Tot* tot = nullptr;
};

// The learned seed of a family: FP32 master weights, Adam moments, BF16
// copies, gradient partials, interface-solve counters
struct SurrDev {
  float *P = nullptr, *m = nullptr, *v = nullptr, *part = nullptr, *loss = nullptr;
  uint16_t* w = nullptr;
  unsigned long long* stats = nullptr;
  int nparts = 0, on = 0;
  int batch = 512;                 // training settings of the last training, for retraining
  double lr = 3e-3;
  uint64_t seed = 7;
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
  int* d_irs_wg = nullptr;   // first workgroup of each block in the all-blocks cell launches, then the total
  int irs_nwg = 0;
  int* d_irs_wl[2] = {nullptr, nullptr};  // the same over the lines along j and k (marching smoothing)
  int irs_nwl[2] = {0, 0};
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
  int* h_flags = nullptr;    // pinned mirror of the status's bad_state and bad_recon
  int lk_grid = 1024;        // workgroups of the deferred interface solves' kernel (4 per CU)
  int iter = 0;
  int keep_source = 0;
  int kernel_timing = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double step_ms = 0.0, kernel_ms = 0.0, lk_ms = 0.0;
  long long kernel_launches = 0, lk_launches = 0;
  std::vector<hipEvent_t> kev;
  std::vector<hipEvent_t> lkev;    // brackets of the deferred interface-solve launches (kernel timing)
  std::vector<FamArgs*> kev_args;  // FamArgs arrays of eager batched steps, freed with the events
  BatchedPlan* batched = nullptr;
  std::vector<SurrDev> surr;   // per family
  int lk_stats_on = 0;
  int retrain_every = 0, retrain_epochs = 20;  // online retraining of the trained networks
  int tile[3] = {-1, -1, -1};  // cells per stage tile; -1 default shape, 0 per-cell kernel
  int dispatch = 0;            // batched stage: 0 persistent queues, 1 one workgroup per item
  double res_threshold = 0.0;  // res-threshold of a steady run (0: never stop on the residual)
  double res_reference = 0.0;  // the first non-zero density residual of the run
  int res_stopped = 0;         // the last ice_gpu_step ended on the residual
  double *res_part = nullptr, *res_sums = nullptr;  // residual-norm partial sums
  std::vector<I400Fam> i400;   // per family
  ice_inflow400_line_t* i400_lines = nullptr;  // the line log, ICE_INFLOW400_LOG entries
  int* i400_n = nullptr;                        // lines logged since the last read
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

// the scratch array of the entry point that allocates nothing, on the context's device
int noalloc_setup(ice_gpu_ctx* c);

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
  c->i400.assign(ng, I400Fam());
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
    std::vector<std::vector<int>> of_blk(c->blk.size());  // the family's records of each block, in table order
    for (int n0 : grp) of_blk[c->rec[n0].v[ICE_R_B] - 1].push_back(n0);
    for (size_t b = 0; b < c->blk.size(); ++b) {
      const BlkDev& B = c->blk[b];
      std::vector<std::pair<long long, int>> at;  // (cell key, record)
      for (int n0 : of_blk[b]) {
        const int* v = c->rec[n0].v;
        if (v[ICE_R_TYPE] == 0) continue;
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
    // the family's 400 records in table order (list_records, Mod_Inflow400.f90:185-200)
    // (sorted by boundary, block then face, and cut into chunks: the counts and
    // maxima of a boundary do not depend on the order of its records)
    std::vector<int> r400;
    for (int n0 : grp)
      if (c->rec[n0].v[ICE_R_TYPE] == 400) r400.push_back(n0);
    auto bnd_of = [&](int n0) { return (c->rec[n0].v[ICE_R_B] - 1) * 6 + (c->rec[n0].v[ICE_R_F] - 1); };
    std::stable_sort(r400.begin(), r400.end(), [&](int x, int y) { return bnd_of(x) < bnd_of(y); });
    std::vector<I400Chunk> chunks;
    for (size_t s0 = 0; s0 < r400.size();) {
      size_t e0 = s0;
      while (e0 < r400.size() && bnd_of(r400[e0]) == bnd_of(r400[s0])) ++e0;
      for (size_t a = s0; a < e0; a += kI400Chunk)
        chunks.push_back(I400Chunk{bnd_of(r400[s0]), (int)a, (int)std::min(e0, a + kI400Chunk), 0});
      s0 = e0;
    }
    I400Fam& I = c->i400[p - 1];
    I.n = (int)r400.size();
    I.nchunk = (int)chunks.size();
    if (I.n) {
      size_t nbnd = 6 * c->blk.size();
      if (dalloc(c, &I.rec, r400.size()) || dalloc(c, &I.acc, nbnd) || dalloc(c, &I.tot, nbnd) ||
          dalloc(c, &I.chunks, chunks.size()))
        return 1;
      HIPCHK(hipMemcpy(I.rec, r400.data(), r400.size() * sizeof(int), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(I.chunks, chunks.data(), chunks.size() * sizeof(I400Chunk), hipMemcpyHostToDevice));
      HIPCHK(hipMemset(I.acc, 0, nbnd * sizeof(I400Acc)));
      HIPCHK(hipMemset(I.tot, 0, nbnd * sizeof(I400Tot)));
    }
  }
  if (!c->i400_n) {
    if (dalloc(c, &c->i400_lines, ICE_INFLOW400_LOG) || dalloc(c, &c->i400_n, 1)) return 1;
    HIPCHK(hipMemset(c->i400_n, 0, sizeof(int)));
  }
  return 0;
}

// count_400_inflow of family p's stage a (Mod_Explicit.f90:129): it reads the
// stage input's interior boundary cells only, so it may run anywhere in the stage
// before the input changes roles
void launch_i400(ice_gpu_ctx* c, int p, const FamArgs& a) {
  const I400Fam& I = c->i400[p];
  if (!I.n) return;
  hipLaunchKernelGGL(k_i400_count, dim3(I.nchunk), dim3(256), 0, c->stream, a, c->fams[p].closure, I.rec, I.chunks,
                     I.acc, I.tot);
  hipLaunchKernelGGL(k_i400_fold, dim3(1), dim3(256), 0, c->stream, (int)(6 * c->blk.size()), p + 1, I.acc, I.tot,
                     c->i400_lines, c->i400_n, c->d_st);
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
  a.imports = F.imports;
  a.lkq_n = F.lkq_n; a.lkq_g = F.lkq_g; a.lkq_b = F.lkq_b; a.lkq_r = F.lkq_r; a.lkq_cap = F.lkq_cap;
  a.bsum = F.bsum; a.bidx = F.bidx; a.connF = F.connF;
  a.rec = c->d_rec; a.donor = c->d_donor; a.weight = c->d_weight;
  a.bfirst = F.bfirst; a.brec = F.brec; a.bcell_g = F.bcell_g; a.nbcell = F.nbcell;
  a.st = c->d_st;
  a.mat.rho = fd.rho_mat; a.mat.cs = fd.cs; a.mat.lv = fd.lv; a.mat.emiss = fd.emiss;
  a.mat.tab = F.prop; a.mat.tmin = F.prop_tmin; a.mat.tmax = F.prop_tmax; a.mat.flags = F.prop_flags;
  for (int m = 0; m < 8; ++m) a.mat.ep[m] = fd.ep[m];
  a.mat.evap = fd.evap_select; a.mat.intf = fd.intf_select; a.mat.surrogate = fd.surrogate;
  a.limiter = c->cfg.limiter; a.nrk = c->cfg.nrk; a.time_accurate = c->cfg.time_accurate;
  a.coupled = c->cfg.coupled; a.drag_sel = c->cfg.drag_select; a.heat_sel = c->cfg.heat_select;
  a.blow_sel = c->cfg.blow_select;
  a.nfrozen = fd.nfrozen;
  static_assert(kNFrozenMax == ICE_NFROZEN_MAX, "frozen slots");
  for (int m = 0; m < ICE_NFROZEN_MAX; ++m) a.frozen[m] = fd.frozen[m];
  a.keep_source = c->keep_source; a.var_yinf = c->cfg.var_yinf; a.iter = c->iter;
  a.cfl_ramp_iter = c->cfg.cfl_ramp_iter; a.ndir = c->cfg.ndir; a.fam = p;
  a.cfl = c->cfg.cfl; a.dt_max = c->cfg.dt_max; a.tau_factor = c->cfg.tau_factor; a.irs_beta = c->cfg.irs_beta;
  const SurrDev& S = c->surr[p];
  a.mlp_on = S.on;
  if (S.on) {
    a.mlp.w1t = S.w; a.mlp.w2t = S.w + 1024; a.mlp.w3t = S.w + 5120; a.mlp.w2n = S.w + 6144; a.mlp.w3n = S.w + 10240;
    a.mlp.b1 = S.P + kOffB1; a.mlp.b2 = S.P + kOffB2; a.mlp.b3 = S.P + kOffB3;
  }
  a.lk_stats = c->lk_stats_on ? S.stats : nullptr;
  return a;
}

// ---------------------------------------------------------------- dispatch
__global__ void k_step_begin(DevStatus* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->dtbits = (unsigned long long)__double_as_longlong(1e+5);
    st->iter = st->iter + 1;
  }
}

__global__ void k_dt_end(DevStatus* st, int time_accurate) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double dt = __longlong_as_double((long long)st->dtbits);
    st->dtglobal = dt;
    if (time_accurate) st->time = st->time + dt;
  }
}

// residual_norm partial sums: workgroup w of block b strides over the block's
// interior cells and writes the five sums of squared changes of rho, u, v, w, T
// (state `q` against the step's start state `o`), reduced in a fixed tree order.
constexpr int kResWG = 64;  // workgroups per (block, family)
__global__ __launch_bounds__(256) void k_resnorm_part(const BlkDev* __restrict__ blk, const double* __restrict__ q,
                                                      const double* __restrict__ o, long long N, int it,
                                                      double* __restrict__ part, int ngroups, int p) {
  const int b = blockIdx.y;
  const BlkDev B = blk[b];
  const long long cells = (long long)B.ni * B.nj * B.nk;
  double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += (long long)gridDim.x * blockDim.x) {
    int i = (int)(t % B.ni) + 1;
    long long r = t / B.ni;
    int j = (int)(r % B.nj) + 1;
    int k = (int)(r / B.nj) + 1;
    long long g = gidx(B, i, j, k);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double d = fabs(q[(long long)m * N + g] - o[(long long)m * N + g]);
      s[m] += d * d;
    }
    double d = fabs(q[(long long)it * N + g] - o[(long long)it * N + g]);
    s[4] += d * d;
  }
  __shared__ double red[5][256];
#pragma unroll
  for (int m = 0; m < 5; ++m) red[m][threadIdx.x] = s[m];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int m = 0; m < 5; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 5) part[(((long long)b * ngroups + p) * kResWG + blockIdx.x) * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// the kResWG partials of every (block, family) summed in workgroup order
__global__ void k_resnorm_sum(const double* __restrict__ part, double* __restrict__ out, int npairs) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs * 5) return;
  int pair = t / 5, m = t % 5;
  double s = 0.0;
  for (int w = 0; w < kResWG; ++w) s += part[((long long)pair * kResWG + w) * 5 + m];
  out[t] = s;
}

void launch_dt(ice_gpu_ctx* c, int p, const FamArgs& a) {
  int cl = c->fams[p].closure, nb = (int)c->blk.size();
  if (cl == ICE_MK) launch_dt_mk(c->stream, c->d_irs_wg, c->irs_nwg, nb, a);
  else if (cl == ICE_IG) launch_dt_ig(c->stream, c->d_irs_wg, c->irs_nwg, nb, a);
  else launch_dt_ag(c->stream, c->d_irs_wg, c->irs_nwg, nb, a);
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
    tx = 254; ty = 1; tz = 1;
  } else if (B.nk == 1) {
    tx = 15; ty = 15; tz = 1;
  } else {
    tx = 6; ty = 6; tz = 5;
  }
  tx = std::min(tx, B.ni); ty = std::min(ty, B.nj); tz = std::min(tz, B.nk);
  T.tx = tx; T.ty = ty; T.tz = tz;
  T.ntx = (B.ni + tx - 1) / tx; T.nty = (B.nj + ty - 1) / ty; T.ntz = (B.nk + tz - 1) / tz;
  return T;
}

// Families with tabulated properties (flags set) run the per-cell stage kernel,
// which takes the table's branches; the tiled and batched kernels are compiled
// for constant properties only.
bool any_table(const ice_gpu_ctx* c) {
  for (const FamDev& F : c->fam)
    if (F.prop_flags) return true;
  return false;
}

void dispatch_stage(ice_gpu_ctx* c, FamArgs& a) {
  const FamDev& F = c->fam[a.fam];
  std::vector<Tile> tiles(c->blk.size());
  for (size_t b = 0; b < c->blk.size(); ++b) {
    tiles[b] = stage_tile(c, c->blk[b]);
    if (F.prop_flags) tiles[b].tx = 0;
  }
  StageLaunch L;
  L.stream = c->stream;
  L.blk = c->blk.data();
  L.tiles = tiles.data();
  L.nb = (int)c->blk.size();
  L.grp = F.grp; L.ngrp = F.ngrp; L.own = F.own; L.nown = F.nown; L.nbcell = F.nbcell;
  L.riemann = c->fams[a.fam].riemann; L.sd = c->cfg.shock_detector;
  L.coupled = c->cfg.coupled ? (c->fams[a.fam].evap_select != 0 ? 2 : 1) : 0;
  L.irs = c->cfg.irs; L.limiter = c->cfg.limiter; L.ndir = c->cfg.ndir;
  L.irs_wg = c->d_irs_wg; L.irs_nwg = c->irs_nwg;
  for (int d = 0; d < 2; ++d) { L.irs_wl[d] = c->d_irs_wl[d]; L.irs_nwl[d] = c->irs_nwl[d]; }
  L.timing = c->kernel_timing; L.kev = &c->kev; L.lkev = &c->lkev;
  L.lk_grid = c->lk_grid;
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
      launch_i400(c, p, a);
      dispatch_stage(c, a);
      F.cur = nxt;
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace

// online retraining of every trained network on the current state (surrogate.hpp)
int retrain_networks(ice_gpu_ctx* c);
namespace {
int residual_norm(ice_gpu_ctx* c, double* total, double* fam);
}

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
  if (cfg->ndir < 1 || cfg->ndir > 3) {
    fail(c, "[ICE::create] ndir (the mesh's resolved directions) must be 1, 2 or 3");
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
  if (noalloc_setup(c)) return 1;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  {
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c->lk_grid = std::max(1, 4 * prop.multiProcessorCount);
  }
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
  {
    std::vector<int> wg(nb + 1, 0);
    for (int b = 0; b < nb; ++b) wg[b + 1] = wg[b] + (int)grid_cells(c->blk[b].cells);
    c->irs_nwg = wg[nb];
    if (dalloc(c, &c->d_irs_wg, nb + 1)) return 1;
    HIPCHK(hipMemcpy(c->d_irs_wg, wg.data(), (nb + 1) * sizeof(int), hipMemcpyHostToDevice));
    for (int d = 1; d <= 2; ++d) {
      for (int b = 0; b < nb; ++b) {
        const BlkDev& B = c->blk[b];
        wg[b + 1] = wg[b] + (int)grid_cells((long long)B.ni * (d == 1 ? B.nk : B.nj));
      }
      c->irs_nwl[d - 1] = wg[nb];
      if (dalloc(c, &c->d_irs_wl[d - 1], nb + 1)) return 1;
      HIPCHK(hipMemcpy(c->d_irs_wl[d - 1], wg.data(), (nb + 1) * sizeof(int), hipMemcpyHostToDevice));
    }
  }
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
  c->surr.resize(cfg->ngroups);
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
    if (cfg->coupled && fams[p].evap_select != 0 && fams[p].intf_select == 1) {
      // room for every interior cell: the tile kernels leave each cell that needs the
      // interface solve to k_lk_finish (108 B per cell for AG, 60 for MK)
      long long cells = 0;
      for (const BlkDev& B : c->blk) cells += B.cells;
      if (cells > (1ll << 31) - 1)
        return fail(c, "[ICE::create] %lld cells: the interface-solve queue holds at most 2^31-1", cells);
      F.lkq_cap = (int)cells;
      if (dalloc(c, &F.lkq_n, 1) || dalloc(c, &F.lkq_g, F.lkq_cap) || dalloc(c, &F.lkq_b, F.lkq_cap) ||
          dalloc(c, &F.lkq_r, (size_t)F.nc * F.lkq_cap))
        return 1;
    }
  }
  if (dalloc(c, &c->d_st, 1)) return 1;
  HIPCHK(hipHostMalloc((void**)&c->h_flags, 2 * sizeof(int), hipHostMallocDefault));
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
    // one copy over the bus, then on the device into the other two buffers
    for (int v = 0; v < F.nc; ++v)
      HIPCHK(hipMemcpy(F.buf[0] + (long long)v * c->N + B.off, h.data() + (size_t)v * npad, npad * sizeof(double),
                       hipMemcpyHostToDevice));
    for (int m = 1; m < 3; ++m)
      for (int v = 0; v < F.nc; ++v)
        HIPCHK(hipMemcpy(F.buf[m] + (long long)v * c->N + B.off, F.buf[0] + (long long)v * c->N + B.off,
                         npad * sizeof(double), hipMemcpyDeviceToDevice));
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

int ice_gpu_get_tau(ice_gpu_ctx* c, int b, int p, double* tau) {
  if (b < 1 || b > (int)c->blk.size() || p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::get_tau] no block/family");
  HIPCHK(hipStreamSynchronize(c->stream));
  const BlkDev& B = c->blk[b - 1];
  long long npad = (long long)B.NI * B.NJ * B.NK;
  std::vector<double> h((size_t)npad);
  HIPCHK(hipMemcpy(h.data(), c->fam[p - 1].tau + B.off, npad * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 1; k <= B.nk; ++k)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i)
        tau[(i - 1) + (long long)B.ni * ((j - 1) + (long long)B.nj * (k - 1))] =
            h[(size_t)((i + 1) + (long long)B.NI * ((j + 1) + (long long)B.NJ * (k + 1)))];
  return 0;
}

int ice_gpu_set_keep_source(ice_gpu_ctx* c, int on) {
  c->keep_source = on == 2 ? 2 : (on ? 1 : 0);
  batched_destroy(c);  // captured graphs hold the families' arguments
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
  for (hipEvent_t e : c->lkev) (void)hipEventDestroy(e);
  c->lkev.clear();
  for (FamArgs* d : c->kev_args) (void)hipFree(d);
  c->kev_args.clear();
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  c->res_stopped = 0;
  const bool res_stop = c->res_threshold > 0.0 && !c->cfg.time_accurate;
  for (int s = 0; s < nsteps; ++s) {
    int rc = (c->cfg.mode == ICE_MODE_BATCHED && !any_table(c)) ? batched_step(c) : plain_step(c);
    if (rc) return rc;
    // the reference stops in the stage whose state failed (Lib_Newstate.f90:143-147,
    // Mod_Fluxes.f90:238-241): no step follows the one that set a fatal flag
    HIPCHK(hipMemcpyAsync(c->h_flags, &c->d_st->bad_state, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->h_flags[0] || c->h_flags[1]) break;
    if (c->retrain_every > 0 && c->iter % c->retrain_every == 0 && retrain_networks(c)) return 1;
    if (res_stop) {
      // Mod_Explicit.f90:196-225: a steady run ends when its density residual has
      // fallen by res-threshold from the first non-zero one
      std::vector<double> tot(5), fam(5 * c->cfg.ngroups);
      if (residual_norm(c, tot.data(), fam.data())) return 1;
      if (c->res_reference == 0.0) c->res_reference = tot[0];
      if (c->res_reference > 0.0 && tot[0] <= c->res_threshold * c->res_reference) {
        c->res_stopped = 1;
        break;
      }
    }
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
  c->lk_ms = 0.0;
  c->lk_launches = 0;
  for (size_t m = 0; m + 1 < c->lkev.size(); m += 2) {
    float km = 0.f;
    (void)hipEventElapsedTime(&km, c->lkev[m], c->lkev[m + 1]);
    c->lk_ms += km;
    c->lk_launches += 1;
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
  if (s.lk_overflow) return fail(c, " [ERROR] [ICE::compute_source] the queue of deferred interface solves overflowed");
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

int ice_gpu_residual_norm(ice_gpu_ctx* c, double* total, double* fam) {
  if (!c) return 1;
  if (!total || !fam) return fail(c, "[ICE::residual_norm] null output");
  return residual_norm(c, total, fam);
}

}  // extern "C"

namespace {

// residual_norm (Mod_Diagnostic.f90:22-69) of the last step into total[5] and
// fam[5 * ngroups], with the context's partial-sum buffers
int residual_norm(ice_gpu_ctx* c, double* total, double* fam) {
  const int nb = (int)c->blk.size(), ng = c->cfg.ngroups;
  if (!c->res_part) {
    if (dalloc(c, &c->res_part, (size_t)nb * ng * kResWG * 5) || dalloc(c, &c->res_sums, (size_t)nb * ng * 5))
      return 1;
  }
  double *part = c->res_part, *sums = c->res_sums;
  for (int p = 0; p < ng; ++p) {
    const FamDev& F = c->fam[p];
    hipLaunchKernelGGL(k_resnorm_part, dim3(kResWG, nb), dim3(256), 0, c->stream, c->d_blk, F.buf[F.cur],
                       F.buf[F.old], c->N, F.nc - 2, part, ng, p);
  }
  hipLaunchKernelGGL(k_resnorm_sum, dim3((nb * ng * 5 + 255) / 256), dim3(256), 0, c->stream, part, sums, nb * ng);
  std::vector<double> h((size_t)nb * ng * 5);
  hipError_t e = hipMemcpyAsync(h.data(), sums, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, "[ICE::residual_norm] %s", hipGetErrorString(e));
  // totals over blocks then families, family sums over blocks (Mod_Diagnostic.f90:106-127)
  for (int m = 0; m < 5; ++m) total[m] = 0.0;
  for (int m = 0; m < 5 * ng; ++m) fam[m] = 0.0;
  for (int b = 0; b < nb; ++b)
    for (int p = 0; p < ng; ++p)
      for (int m = 0; m < 5; ++m) {
        double x = h[((size_t)b * ng + p) * 5 + m];
        total[m] = total[m] + x;
        fam[p * 5 + m] = fam[p * 5 + m] + x;
      }
  for (int m = 0; m < 5; ++m) total[m] = std::sqrt(total[m]);
  for (int m = 0; m < 5 * ng; ++m) fam[m] = std::sqrt(fam[m]);
  return 0;
}

}  // namespace

extern "C" {

int ice_gpu_set_res_threshold(ice_gpu_ctx* c, double threshold) {
  if (!c) return 1;
  if (!(threshold >= 0.0)) return fail(c, "[ICE::set_res_threshold] res-threshold must be >= 0");
  c->res_threshold = threshold;
  c->res_reference = 0.0;
  c->res_stopped = 0;
  return 0;
}

int ice_gpu_res_stop(ice_gpu_ctx* c, int32_t* stopped, double* reference) {
  if (!c) return 1;
  if (stopped) *stopped = c->res_stopped;
  if (reference) *reference = c->res_reference;
  return 0;
}

int ice_gpu_timing(ice_gpu_ctx* c, double* step_ms, double* kernel_ms, int64_t* launches) {
  if (step_ms) *step_ms = c->step_ms;
  if (kernel_ms) *kernel_ms = c->kernel_ms;
  if (launches) *launches = c->kernel_launches;
  return 0;
}

int ice_gpu_lk_timing(ice_gpu_ctx* c, double* ms, int64_t* launches) {
  if (ms) *ms = c->lk_ms;
  if (launches) *launches = c->lk_launches;
  return 0;
}

int ice_gpu_set_tiles(ice_gpu_ctx* c, int tx, int ty, int tz) {
  if (tx == 0) {
    c->tile[0] = c->tile[1] = c->tile[2] = 0;
    batched_destroy(c);
    return 0;
  }
  if (tx < 0) {
    c->tile[0] = c->tile[1] = c->tile[2] = -1;
    batched_destroy(c);
    return 0;
  }
  for (const BlkDev& B : c->blk) {
    int n[3] = {B.ni, B.nj, B.nk}, t[3] = {std::min(tx, B.ni), std::min(ty, B.nj), std::min(tz, B.nk)};
    bool bad = ty < 1 || tz < 1 || t[0] * t[1] * t[2] > kThreads;
    for (int d = 0; d < 3; ++d)
      if (n[d] > 1)
        bad = bad || t[0] * t[1] * t[2] / t[d] * (t[d] + 1) > kMaxTileFaces ||
              t[0] * t[1] * t[2] / t[d] * (t[d] + 2) > kMaxExtCells;
    if (bad)
      return fail(c, "[ICE::set_tiles] a %d x %d x %d tile has more cells or faces than the %d threads of a workgroup, "
                  "or more than %d cells widened by one along a direction", tx, ty, tz, kThreads, kMaxExtCells);

  }
  c->tile[0] = tx; c->tile[1] = ty; c->tile[2] = tz;
  batched_destroy(c);  // the batched plan's work items are tiles
  return 0;
}

// ---------------------------------------------------------------- autotuner
// Candidate stage tiles (cells along i, j, k) for blocks resolving 3, 2 and 1
// directions; each must keep its cells and the faces of each direction within
// the 256 threads of a workgroup (checked by ice_gpu_set_tiles).
namespace {
// (most of them widened by one cell along any direction still fit the 256 threads,
// so each direction's limited slopes take one pass of the workgroup)
const int kTiles3[][3] = {{6, 6, 5}, {7, 5, 5}, {5, 6, 6}, {6, 5, 6}, {7, 6, 4}, {8, 5, 4}, {7, 7, 4}, {7, 6, 5}};
const int kTiles2[][3] = {{15, 15, 1}, {14, 16, 1}, {16, 14, 1}, {12, 18, 1}, {18, 12, 1}, {15, 16, 1}};
const int kTiles1[][3] = {{254, 1, 1}, {128, 1, 1}, {192, 1, 1}};

// a device copy of everything a step changes, to run trial steps and put it back
struct Snapshot {
  std::vector<double*> buf;  // per family: 3 state buffers, tau, dt
  std::vector<int> cur, old;
  DevStatus st;
  int iter = 0;
  std::vector<std::vector<I400Tot>> i400;  // per family: the 400 boundaries' totals
  std::vector<ice_inflow400_line_t> lines;
  int nlines = 0;
};

// the 400 counts, totals and line log (the accumulators are zero between steps)
int snapshot_i400(ice_gpu_ctx* c, Snapshot& S, bool take) {
  const size_t nbnd = 6 * c->blk.size();
  if (!c->i400_n) return 0;
  if (take) {
    S.i400.assign(c->i400.size(), {});
    for (size_t p = 0; p < c->i400.size(); ++p)
      if (c->i400[p].n) {
        S.i400[p].resize(nbnd);
        HIPCHK(hipMemcpy(S.i400[p].data(), c->i400[p].tot, nbnd * sizeof(I400Tot), hipMemcpyDeviceToHost));
      }
    HIPCHK(hipMemcpy(&S.nlines, c->i400_n, sizeof(int), hipMemcpyDeviceToHost));
    S.lines.resize(ICE_INFLOW400_LOG);
    HIPCHK(hipMemcpy(S.lines.data(), c->i400_lines, S.lines.size() * sizeof(ice_inflow400_line_t),
                     hipMemcpyDeviceToHost));
    return 0;
  }
  for (size_t p = 0; p < c->i400.size(); ++p)
    if (c->i400[p].n)
      HIPCHK(hipMemcpy(c->i400[p].tot, S.i400[p].data(), nbnd * sizeof(I400Tot), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->i400_n, &S.nlines, sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->i400_lines, S.lines.data(), S.lines.size() * sizeof(ice_inflow400_line_t),
                   hipMemcpyHostToDevice));
  return 0;
}

int snapshot(ice_gpu_ctx* c, Snapshot& S, bool take) {
  const size_t N = (size_t)c->N;
  if (take) {
    S.buf.clear();
    for (auto& F : c->fam) {
      for (int m = 0; m < 5; ++m) {
        double* src = m < 3 ? F.buf[m] : (m == 3 ? F.tau : F.dt);
        size_t n = (m < 3 ? (size_t)F.nc : 1) * N;
        double* d = nullptr;
        if (src) {
          if (dalloc(c, &d, n)) return 1;
          HIPCHK(hipMemcpyAsync(d, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        }
        S.buf.push_back(d);
      }
      S.cur.push_back(F.cur);
      S.old.push_back(F.old);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&S.st, c->d_st, sizeof S.st, hipMemcpyDeviceToHost));
    S.iter = c->iter;
    return snapshot_i400(c, S, true);
  }
  for (size_t p = 0; p < c->fam.size(); ++p) {
    FamDev& F = c->fam[p];
    for (int m = 0; m < 5; ++m) {
      double* dst = m < 3 ? F.buf[m] : (m == 3 ? F.tau : F.dt);
      size_t n = (m < 3 ? (size_t)F.nc : 1) * N;
      if (dst) HIPCHK(hipMemcpyAsync(dst, S.buf[p * 5 + m], n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    F.cur = S.cur[p];
    F.old = S.old[p];
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->d_st, &S.st, sizeof S.st, hipMemcpyHostToDevice));
  c->iter = S.iter;
  return snapshot_i400(c, S, false);
}

void snapshot_free(Snapshot& S) {
  for (double* d : S.buf)
    if (d) (void)hipFree(d);
  S.buf.clear();
}
}  // namespace

int ice_gpu_set_dispatch(ice_gpu_ctx* c, int dispatch) {
  if (dispatch < 0 || dispatch > 2)
    return fail(c, "[ICE::set_dispatch] dispatch is 0 (queues), 1 (items) or 2 (items family by family)");
  c->dispatch = dispatch;
  batched_destroy(c);  // its captured graphs hold the old launches
  return 0;
}

int ice_gpu_autotune(ice_gpu_ctx* c, int steps, int32_t* best, double* ms_out) {
  if (c->rec.empty()) return fail(c, "[ICE::autotune] no boundary table: call ice_gpu_set_boundaries first");
  if (steps < 1) steps = 1;
  int ndim = 1;
  for (const BlkDev& B : c->blk) ndim = std::max(ndim, (B.nj > 1) + (B.nk > 1) + 1);
  const int(*cand)[3] = ndim == 3 ? kTiles3 : (ndim == 2 ? kTiles2 : kTiles1);
  const int ncand = ndim == 3 ? 8 : (ndim == 2 ? 6 : 3);
  int keep[3] = {c->tile[0], c->tile[1], c->tile[2]};
  Snapshot S;
  if (snapshot(c, S, true)) {
    snapshot_free(S);
    return 1;
  }
  double best_ms = std::numeric_limits<double>::infinity();
  int bi = -1, bd = c->dispatch;
  std::string err;
  const int ndisp = c->cfg.mode == ICE_MODE_BATCHED ? 3 : 1;
  for (int k = 0; k < ncand * ndisp; ++k) {
    double ms = std::numeric_limits<double>::infinity();
    if (ndisp > 1) c->dispatch = k / ncand;
    if (ice_gpu_set_tiles(c, cand[k % ncand][0], cand[k % ncand][1], cand[k % ncand][2]) == 0) {
      ice_status_t st;
      // one untimed step (plan, graph capture), then the timed ones
      int rc = ice_gpu_step(c, 1, &st);
      if (rc == 0) rc = ice_gpu_step(c, steps, &st);
      if (rc == 0) ms = c->step_ms / steps;
      else err = c->err;
      if (snapshot(c, S, false)) {
        snapshot_free(S);
        return 1;
      }
    }
    if (ms_out) ms_out[k] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      bi = k % ncand;
      bd = c->dispatch;
    }
  }
  c->dispatch = bd;
  snapshot_free(S);
  if (bi < 0) {
    c->tile[0] = keep[0]; c->tile[1] = keep[1]; c->tile[2] = keep[2];
    batched_destroy(c);
    return fail(c, "[ICE::autotune] no candidate tile ran: %s", err.c_str());
  }
  ice_gpu_set_tiles(c, cand[bi][0], cand[bi][1], cand[bi][2]);
  if (best) {
    best[0] = cand[bi][0]; best[1] = cand[bi][1]; best[2] = cand[bi][2];
    best[3] = bd;
  }
  return 0;
}

int ice_gpu_set_kernel_timing(ice_gpu_ctx* c, int on) {
  c->kernel_timing = on ? 1 : 0;
  return 0;
}

// compute_ghost and fill_second_ghost (Lib_Ghost.f90:29-72, 270-308) of every
// family on its current state, family by family as the stages fill them
int ice_gpu_fill_ghosts(ice_gpu_ctx* c) {
  if (!c) return 1;
  if (c->rec.empty()) return fail(c, "[ICE::fill_ghosts] no boundary table: call ice_gpu_set_boundaries first");
  for (int p = 0; p < c->cfg.ngroups; ++p) {
    const FamDev& F = c->fam[p];
    FamArgs a = fam_args(c, p);
    int cl = c->fams[p].closure;
    if (cl == ICE_MK) launch_ghosts_mk(c->stream, F.grp, F.ngrp, a);
    else if (cl == ICE_IG) launch_ghosts_ig(c->stream, F.grp, F.ngrp, a);
    else launch_ghosts_ag(c->stream, F.grp, F.ngrp, a);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// The rates in force on family p in block b (the imports of clear_import_rates /
// add_birth_rate / add_absorption_rate, Mod_Exchange.f90:242-324), conservative
// slots (nc, ni, nj, nk); every stage adds them over the cell's dt (add_births).
// imports = NULL clears them.
int ice_gpu_put_imports(ice_gpu_ctx* c, int b, int p, const double* imports) {
  if (!c) return 1;
  if (b < 1 || b > (int)c->blk.size() || p < 1 || p > c->cfg.ngroups)
    return fail(c, "[ICE::put_imports] no block %d / family %d", b, p);
  FamDev& F = c->fam[p - 1];
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!F.imports) {
    if (dalloc(c, &F.imports, (size_t)F.nc * c->N) || fill(c, F.imports, (long long)F.nc * c->N, 0.0)) return 1;
    batched_destroy(c);  // captured graphs hold the families' arguments
  }
  if (imports) {
    if (put_interior(c, b, F.imports, F.nc, imports, true)) return 1;
  } else {
    const BlkDev& B = c->blk[b - 1];
    for (int v = 0; v < F.nc; ++v)
      if (fill(c, F.imports + (long long)v * c->N + B.off, (long long)B.NI * B.NJ * B.NK, 0.0)) return 1;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// add_birth (Mod_Exchange.f90:283-300) of n packets of family p: packet m, born in
// cell (b, i, j, k) = cells[4m .. 4m+3] (1-based), is pk[6m .. 6m+5] = [mass,
// momentum (3), energy, number] per unit volume (birth_packet), added to the cell's
// conservative state; the packets of one cell in their order.
int ice_gpu_add_births(ice_gpu_ctx* c, int p, int n, const int32_t* cells, const double* pk) {
  if (!c) return 1;
  if (p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::add_births] no family %d", p);
  if (n <= 0) return 0;
  std::vector<std::pair<long long, int>> at((size_t)n);
  for (int m = 0; m < n; ++m) {
    const int* x = cells + 4 * (size_t)m;
    if (x[0] < 1 || x[0] > (int)c->blk.size()) return fail(c, "[ICE::add_births] packet %d: no block %d", m + 1, x[0]);
    const BlkDev& B = c->blk[x[0] - 1];
    if (x[1] < 1 || x[1] > B.ni || x[2] < 1 || x[2] > B.nj || x[3] < 1 || x[3] > B.nk)
      return fail(c, "[ICE::add_births] packet %d: cell %d %d %d is not in block %d", m + 1, x[1], x[2], x[3], x[0]);
    at[m] = {padded_of(B, x[1], x[2], x[3]), m};
  }
  std::stable_sort(at.begin(), at.end(),
                   [](const std::pair<long long, int>& u, const std::pair<long long, int>& v) { return u.first < v.first; });
  std::vector<long long> cell;
  std::vector<int> first;
  std::vector<double> pks((size_t)n * 6);
  for (size_t m = 0; m < at.size(); ++m) {
    if (m == 0 || at[m].first != at[m - 1].first) {
      cell.push_back(at[m].first);
      first.push_back((int)m);
    }
    std::memcpy(&pks[m * 6], pk + (size_t)at[m].second * 6, 6 * sizeof(double));
  }
  first.push_back(n);
  long long* d_cell = nullptr;
  int* d_first = nullptr;
  double* d_pk = nullptr;
  int rc = dalloc(c, &d_cell, cell.size()) || dalloc(c, &d_first, first.size()) || dalloc(c, &d_pk, pks.size());
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpy(d_cell, cell.data(), cell.size() * sizeof(long long), hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess) e = hipMemcpy(d_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess) e = hipMemcpy(d_pk, pks.data(), pks.size() * sizeof(double), hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess) {
    FamArgs a = fam_args(c, p - 1);
    int cl = c->fams[p - 1].closure, nc = (int)cell.size();
    if (cl == ICE_MK) launch_births_mk(c->stream, a, d_cell, d_first, d_pk, nc);
    else if (cl == ICE_IG) launch_births_ig(c->stream, a, d_cell, d_first, d_pk, nc);
    else launch_births_ag(c->stream, a, d_cell, d_first, d_pk, nc);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (d_cell) (void)hipFree(d_cell);
  if (d_first) (void)hipFree(d_first);
  if (d_pk) (void)hipFree(d_pk);
  if (rc) return 1;
  if (e != hipSuccess) return fail(c, "[ICE::add_births] %s", hipGetErrorString(e));
  return 0;
}

int ice_gpu_set_mode(ice_gpu_ctx* c, int mode) {
  if (!c) return 1;
  if (mode != ICE_MODE_PLAIN && mode != ICE_MODE_BATCHED)
    return fail(c, "[ICE::set_mode] mode is ICE_MODE_PLAIN (0) or ICE_MODE_BATCHED (1)");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->cfg.mode = mode;
  return 0;
}

int ice_gpu_set_surrogate(ice_gpu_ctx* c, int p, int mode) {
  if (!c) return 1;
  if (p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::set_surrogate] no family %d", p);
  if (mode < 0 || mode > 2) return fail(c, "[ICE::set_surrogate] mode is 0 (plain), 1 (network) or 2 (no network)");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->fams[p - 1].surrogate = mode;
  batched_destroy(c);  // captured graphs hold the families' arguments
  return 0;
}

// Restart (Wrap_Setup.f90:158-214): grid%iter, grid%time and res_reference as the
// solution read stored them (IO_Solution.f90:20-30), so that the CFL ramp
// (Mod_dt.f90:83-85), the time of a time-accurate run and the unit of
// res-threshold go on as in the uninterrupted run.
int ice_gpu_set_counters(ice_gpu_ctx* c, int32_t iter, double time, double res_reference) {
  if (!c) return 1;
  if (iter < 0 || !std::isfinite(time) || !(res_reference >= 0.0))
    return fail(c, "[ICE::set_counters] iter >= 0, a finite time and res_reference >= 0");
  HIPCHK(hipStreamSynchronize(c->stream));
  DevStatus s;
  HIPCHK(hipMemcpy(&s, c->d_st, sizeof s, hipMemcpyDeviceToHost));
  s.iter = iter;
  s.time = time;
  HIPCHK(hipMemcpy(c->d_st, &s, sizeof s, hipMemcpyHostToDevice));
  c->iter = iter;
  c->res_reference = res_reference;
  return 0;
}

int ice_gpu_get_counters(ice_gpu_ctx* c, int32_t* iter, double* time, double* res_reference) {
  if (!c) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  DevStatus s;
  HIPCHK(hipMemcpy(&s, c->d_st, sizeof s, hipMemcpyDeviceToHost));
  if (iter) *iter = c->iter;
  if (time) *time = s.time;
  if (res_reference) *res_reference = c->res_reference;
  return 0;
}

// Two contexts one level apart (Check_Multigrid, Lib_Multigrid.f90:12-35): same
// device, families and blocks, each coarse block half the fine one per direction
int level_pair(ice_gpu_ctx* f, ice_gpu_ctx* c) {
  if (!f || !c) return 1;
  if (f->device != c->device) return fail(f, "[ICE::Multigrid] the levels are on different devices");
  if (f->cfg.ngroups != c->cfg.ngroups || f->blk.size() != c->blk.size())
    return fail(f, "[ICE::Multigrid] the levels have different families or blocks");
  for (int p = 0; p < f->cfg.ngroups; ++p)
    if (f->fams[p].closure != c->fams[p].closure) return fail(f, "[ICE::Multigrid] family %d changes closure", p + 1);
  for (size_t b = 0; b < f->blk.size(); ++b) {
    const BlkDev &F = f->blk[b], &C = c->blk[b];
    bool ok = F.ni % 2 == 0 && F.nj % 2 == 0 && C.ni * 2 == F.ni && C.nj * 2 == F.nj &&
              (F.nk == 1 ? C.nk == 1 : (F.nk % 2 == 0 && C.nk * 2 == F.nk));
    if (!ok) return fail(f, " [ERROR] [ICE::Check_Multigrid] block %d not divisible by 2^(MGL-1)", (int)b + 1);
  }
  HIPCHK(hipStreamSynchronize(f->stream));
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(f, "[ICE::Multigrid] %s", hipGetErrorString(e));
  return 0;
}

int ice_gpu_restrict(ice_gpu_ctx* fine, ice_gpu_ctx* coarse) {
  if (level_pair(fine, coarse)) return 1;
  ice_gpu_ctx* c = fine;
  for (int p = 0; p < fine->cfg.ngroups; ++p) {
    FamArgs fa = fam_args(fine, p), ca = fam_args(coarse, p);
    for (size_t b = 0; b < fine->blk.size(); ++b) {
      dim3 grid(grid_of(coarse->blk[b].cells)), blk(kThreads);
      int cl = fine->fams[p].closure;
      if (cl == ICE_MK) hipLaunchKernelGGL(k_restrict<0>, grid, blk, 0, fine->stream, fa, ca, (int)b);
      else if (cl == ICE_IG) hipLaunchKernelGGL(k_restrict<1>, grid, blk, 0, fine->stream, fa, ca, (int)b);
      else hipLaunchKernelGGL(k_restrict<2>, grid, blk, 0, fine->stream, fa, ca, (int)b);
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(fine->stream));
  return 0;
}

int ice_gpu_prolong(ice_gpu_ctx* coarse, ice_gpu_ctx* fine) {
  if (level_pair(fine, coarse)) return 1;
  ice_gpu_ctx* c = fine;
  for (int p = 0; p < fine->cfg.ngroups; ++p) {
    FamArgs fa = fam_args(fine, p), ca = fam_args(coarse, p);
    for (size_t b = 0; b < fine->blk.size(); ++b) {
      dim3 grid(grid_of(coarse->blk[b].cells)), blk(kThreads);
      int nc = nc_of(fine->fams[p].closure);
      if (nc == 6) hipLaunchKernelGGL(k_prolong<6>, grid, blk, 0, fine->stream, ca, fa, (int)b);
      else if (nc == 7) hipLaunchKernelGGL(k_prolong<7>, grid, blk, 0, fine->stream, ca, fa, (int)b);
      else hipLaunchKernelGGL(k_prolong<12>, grid, blk, 0, fine->stream, ca, fa, (int)b);
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(fine->stream));
  return 0;
}

int ice_gpu_set_properties(ice_gpu_ctx* c, int p, int t_min, int n, const double* cols, int32_t flags) {
  if (!c) return 1;
  if (p < 1 || p > c->cfg.ngroups) return fail(c, "[ICE::set_properties] family %d out of range", p);
  if (n < 2) return fail(c, "[ICE::Load_Table] a property table needs at least 2 rows (found %d)", n);
  if (t_min < 0) return fail(c, "[ICE::Load_Table] negative temperatures in the property table");
  if (flags & ~(ICE_PROP_RHO_VARIES | ICE_PROP_CS_VARIES | ICE_PROP_USE_PSAT | ICE_PROP_USE_LV | ICE_PROP_USE_CPV |
                ICE_PROP_USE_EMISS))
    return fail(c, "[ICE::set_properties] unknown flags %d", (int)flags);
  static_assert(ICE_PROP_NCOL == kPropCols, "column count");
  static_assert(ICE_PROP_RHO_VARIES == kRhoVaries && ICE_PROP_CS_VARIES == kCsVaries && ICE_PROP_USE_PSAT == kUsePsat &&
                    ICE_PROP_USE_LV == kUseLv && ICE_PROP_USE_CPV == kUseCpv && ICE_PROP_USE_EMISS == kUseEmiss,
                "flag values");
  FamDev& F = c->fam[p - 1];
  if (F.prop) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(F.prop));
    F.prop = nullptr;
  }
  if (dalloc(c, &F.prop, (size_t)ICE_PROP_NCOL * n)) return 1;
  HIPCHK(hipMemcpy(F.prop, cols, (size_t)ICE_PROP_NCOL * n * sizeof(double), hipMemcpyHostToDevice));
  F.prop_tmin = t_min;
  F.prop_tmax = t_min + n - 1;
  F.prop_flags = flags;
  // mat_rho without rho_varies and mat_cp_const take the table's first row (Lib_Properties.f90:18-53)
  c->fams[p - 1].rho_mat = cols[(size_t)ICE_PROP_RHO * n];
  c->fams[p - 1].cs = cols[(size_t)ICE_PROP_CP * n];
  batched_destroy(c);  // captured graphs hold the old material
  return 0;
}

int ice_gpu_inflow400_lines(ice_gpu_ctx* c, ice_inflow400_line_t* out, int max, int32_t* n, int64_t* dropped) {
  if (!c) return 1;
  *n = 0;
  *dropped = 0;
  if (!c->i400_n) return 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  int logged = 0;
  HIPCHK(hipMemcpy(&logged, c->i400_n, sizeof(int), hipMemcpyDeviceToHost));
  int take = std::min(std::min(logged, ICE_INFLOW400_LOG), std::max(max, 0));
  if (take > 0) HIPCHK(hipMemcpy(out, c->i400_lines, take * sizeof(ice_inflow400_line_t), hipMemcpyDeviceToHost));
  *n = take;
  *dropped = (int64_t)logged - take;
  int zero = 0;
  HIPCHK(hipMemcpy(c->i400_n, &zero, sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

int ice_gpu_inflow400_totals(ice_gpu_ctx* c, ice_inflow400_total_t* out, int max, int32_t* n) {
  if (!c) return 1;
  *n = 0;
  if (!c->i400_n) return 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t nbnd = 6 * c->blk.size();
  std::vector<I400Tot> T(nbnd);
  int k = 0;
  for (size_t p = 0; p < c->i400.size(); ++p) {
    if (!c->i400[p].n) continue;
    HIPCHK(hipMemcpy(T.data(), c->i400[p].tot, nbnd * sizeof(I400Tot), hipMemcpyDeviceToHost));
    for (size_t m = 0; m < nbnd; ++m) {
      if (T[m].nstages == 0) continue;
      if (k >= max) return fail(c, "[ICE::inflow400] more than %d boundaries let particles in", max);
      ice_inflow400_total_t& o = out[k++];
      o.block = (int)(m / 6) + 1;
      o.face = (int)(m % 6) + 1;
      o.family = (int)p + 1;
      o.stages = T[m].nstages;
      o.first_iter = T[m].first_iter;
      o.first_faces = T[m].first_faces;
      o.first_largest = T[m].first_x;
      o.largest = T[m].largest_x;
    }
  }
  *n = k;
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
  for (hipEvent_t e : c->lkev) (void)hipEventDestroy(e);
  auto F = [](void* p) {
    if (p) (void)hipFree(p);
  };
  F(c->d_blk); F(c->d_irs_wg); F(c->d_irs_wl[0]); F(c->d_irs_wl[1]); F(c->vol); F(c->dl); F(c->M); F(c->fa); F(c->gas); F(c->gR); F(c->gG); F(c->gK); F(c->gMu);
  F(c->yinf); F(c->d_rec); F(c->d_donor); F(c->d_weight); F(c->d_st);
  for (auto& f : c->fam) {
    for (int m = 0; m < 3; ++m) F(f.buf[m]);
    F(f.beta); F(f.dt); F(f.tau); F(f.src); F(f.R); F(f.T); F(f.bsum); F(f.connF); F(f.inj); F(f.bidx); F(f.imports);
    F(f.lkq_n); F(f.lkq_g); F(f.lkq_b); F(f.lkq_r);
    F(f.grp); F(f.own); F(f.bfirst); F(f.brec); F(f.bcell_g); F(f.prop);
  }
  for (auto& S : c->surr) {
    F(S.P); F(S.m); F(S.v); F(S.part); F(S.loss); F(S.w); F(S.stats);
  }
  for (auto& I : c->i400) {
    F(I.rec); F(I.acc); F(I.tot); F(I.chunks);
  }
  F(c->i400_lines); F(c->i400_n); F(c->res_part); F(c->res_sums);
  batched_destroy(c);
  if (c->h_flags) (void)hipHostFree(c->h_flags);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"

// ================================================================= batched path
// All blocks and families of a stage in one persistent kernel per closure
// group, fed from per-XCD tile queues; the step's whole launch sequence is
// captured once per buffer rotation in a hipGraph and replayed.
struct BatchGroup {
  int closure = 0, riemann = 0;
  std::vector<int> fams;                 // 0-based family numbers
  WorkItem* d_items = nullptr;
  int* d_qstart = nullptr;
  unsigned long long* d_ticket = nullptr;
  int nitems = 0, nq = 8, grid = 0;
};

struct BatchGraph {
  int key = -1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<FamArgs*> d_args;           // device FamArgs arrays owned by the graph
  std::vector<int> end_cur;              // each family's state buffer after the step
};

struct BatchedPlan {
  std::vector<BatchGroup> groups;
  std::vector<BatchGraph> graphs;
  int cus = 0;
};

namespace {

int build_plan(ice_gpu_ctx* c) {
  BatchedPlan* P = new BatchedPlan();
  c->batched = P;
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, c->device));
  P->cus = prop.multiProcessorCount;
  for (int p = 0; p < c->cfg.ngroups; ++p) {
    int cl = c->fams[p].closure, rs = c->fams[p].riemann;
    BatchGroup* g = nullptr;
    for (auto& x : P->groups)
      if (x.closure == cl && x.riemann == rs) g = &x;
    if (!g) {
      P->groups.emplace_back();
      g = &P->groups.back();
      g->closure = cl;
      g->riemann = rs;
    }
    g->fams.push_back(p);
  }
  for (auto& g : P->groups) {
    // items block by block, tile by tile, the group's families side by side so a
    // tile's geometry and gas are read from L2 by all of them; with dispatch 2 each
    // block's tiles family by family, so that a tile row's cells of one family (not
    // of all) share an XCD's L2
    std::vector<WorkItem> items;
    for (size_t b = 0; b < c->blk.size(); ++b) {
      Tile T = stage_tile(c, c->blk[b]);
      if (T.tx == 0) return fail(c, "[ICE::batched] the batched mode needs a tiled stage kernel");
      const int nt = T.ntx * T.nty * T.ntz, nf = (int)g.fams.size();
      for (int k = 0; k < nt * nf; ++k) {
        const bool fam_major = c->dispatch == 2;
        WorkItem w;
        w.fam = g.fams[fam_major ? k / nt : k % nf];
        w.block = (int)b;
        w.tile = fam_major ? k % nt : k / nf;
        w.pad = 0;
        w.T = T;
        items.push_back(w);
      }
    }
    g.nitems = (int)items.size();
    g.nq = 8;
    std::vector<int> qs(g.nq + 1);
    for (int q = 0; q <= g.nq; ++q) qs[q] = (int)((long long)g.nitems * q / g.nq);
    if (dalloc(c, &g.d_items, items.size()) || dalloc(c, &g.d_qstart, qs.size()) || dalloc(c, &g.d_ticket, g.nq))
      return 1;
    HIPCHK(hipMemcpy(g.d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.d_qstart, qs.data(), qs.size() * sizeof(int), hipMemcpyHostToDevice));
    int cp = 0;
    if (c->cfg.coupled) {
      cp = 1;
      for (int p : g.fams)
        if (c->fams[p].evap_select != 0) cp = 2;
    }
    int wgs = g.closure == ICE_MK ? persistent_wgs_mk(P->cus, cp)
                                  : (g.closure == ICE_IG ? persistent_wgs_ig(P->cus, cp) : persistent_wgs_ag(P->cus, cp));
    g.grid = std::max(g.nq, (wgs / g.nq) * g.nq);
  }
  return 0;
}

// The FamArgs of every pass of one step with the current buffer roles (index 0
// the dt pass, then one per RK stage), uploaded; the roles after the step.
struct StepArgs {
  std::vector<std::vector<FamArgs>> h;
  std::vector<FamArgs*> d;
  std::vector<int> end_cur, start_cur;
};

int prepare_step_args(ice_gpu_ctx* c, StepArgs& A, std::vector<FamArgs*>& owner) {
  int ng = c->cfg.ngroups, nrk = c->cfg.nrk;
  std::vector<int> cur(ng), old(ng);
  for (int p = 0; p < ng; ++p) cur[p] = old[p] = c->fam[p].cur;
  A.start_cur = cur;
  A.h.assign(nrk + 1, std::vector<FamArgs>(ng));
  for (int pass = 0; pass <= nrk; ++pass) {
    std::vector<int> nxt(ng);
    for (int p = 0; p < ng; ++p) {
      FamDev& F = c->fam[p];
      int kc = F.cur, ko = F.old;
      F.cur = cur[p];
      F.old = old[p];
      FamArgs a = fam_args(c, p);
      if (pass > 0) {
        nxt[p] = other_buffer(cur[p], old[p]);
        a.nxt = F.buf[nxt[p]];
        a.srk = pass;
        a.last_stage = pass == nrk;
      }
      A.h[pass][p] = a;
      F.cur = kc;
      F.old = ko;
    }
    if (pass > 0) cur = nxt;
  }
  A.end_cur = cur;
  A.d.assign(nrk + 1, nullptr);
  for (int pass = 0; pass <= nrk; ++pass) {
    if (dalloc(c, &A.d[pass], ng)) return 1;
    owner.push_back(A.d[pass]);
    HIPCHK(hipMemcpy(A.d[pass], A.h[pass].data(), ng * sizeof(FamArgs), hipMemcpyHostToDevice));
  }
  return 0;
}

BatchLaunch group_launch(ice_gpu_ctx* c, const BatchGroup& g, const std::vector<FamArgs>& h, FamArgs* d,
                         std::vector<StageLaunch>& per) {
  BatchLaunch B;
  B.stream = c->stream;
  B.d_fams = d;
  B.h_fams = h.data();
  B.fam_ids = g.fams.data();
  B.nfam = (int)g.fams.size();
  per.resize(g.fams.size());
  for (size_t f = 0; f < g.fams.size(); ++f) {
    const FamDev& F = c->fam[g.fams[f]];
    StageLaunch& L = per[f];
    std::memset(&L, 0, sizeof L);
    L.stream = c->stream;
    L.grp = F.grp; L.ngrp = F.ngrp; L.own = F.own; L.nown = F.nown; L.nbcell = F.nbcell;
  }
  B.per_fam = per.data();
  B.Q.items = g.d_items;
  B.Q.qstart = g.d_qstart;
  B.Q.ticket = g.d_ticket;
  B.Q.nq = g.nq;
  B.grid = g.grid;
  B.dispatch = c->dispatch;
  B.nitems = g.nitems;
  B.irs = c->cfg.irs;
  B.ndir = c->cfg.ndir;
  B.irs_wg = c->d_irs_wg;
  B.irs_nwg = c->irs_nwg;
  for (int d = 0; d < 2; ++d) { B.irs_wl[d] = c->d_irs_wl[d]; B.irs_nwl[d] = c->irs_nwl[d]; }
  B.nb = (int)c->blk.size();
  B.blk = c->blk.data();
  B.riemann = g.riemann;
  B.sd = c->cfg.shock_detector;
  B.coupled = 0;
  if (c->cfg.coupled) {
    B.coupled = 1;
    for (int p : g.fams)
      if (c->fams[p].evap_select != 0) B.coupled = 2;
  }
  B.limiter = c->cfg.limiter;
  B.lk_grid = c->lk_grid;
  return B;
}

// the launches of one step (no allocation: capturable)
void launch_batched_step(ice_gpu_ctx* c, const StepArgs& A, bool timed) {
  BatchedPlan* P = c->batched;
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, c->stream, c->d_st);
  for (auto& g : P->groups) {
    std::vector<StageLaunch> per;
    BatchLaunch B = group_launch(c, g, A.h[0], A.d[0], per);
    if (g.closure == ICE_MK) launch_batched_dt_mk(B);
    else if (g.closure == ICE_IG) launch_batched_dt_ig(B);
    else launch_batched_dt_ag(B);
  }
  hipLaunchKernelGGL(k_dt_end, dim3(1), dim3(64), 0, c->stream, c->d_st, c->cfg.time_accurate);
  for (int srk = 1; srk <= c->cfg.nrk; ++srk) {
    for (auto& g : P->groups) {
      std::vector<StageLaunch> per;
      BatchLaunch B = group_launch(c, g, A.h[srk], A.d[srk], per);
      for (int p : g.fams) launch_i400(c, p, A.h[srk][p]);
      if (timed) {
        (void)hipEventCreate(&B.ev0);
        (void)hipEventCreate(&B.ev1);
        (void)hipEventCreate(&B.ev2);
      }
      if (g.closure == ICE_MK) launch_batched_mk(B);
      else if (g.closure == ICE_IG) launch_batched_ig(B);
      else launch_batched_ag(B);
      if (timed) {
        c->kev.push_back(B.ev0);
        c->kev.push_back(B.ev1);
        c->lkev.push_back(B.ev1);
        c->lkev.push_back(B.ev2);
      }
    }
  }
}

}  // namespace

int batched_step(ice_gpu_ctx* c) {
  if (!c->batched && build_plan(c)) return 1;
  BatchedPlan* P = c->batched;
  c->iter += 1;
  const int ng = c->cfg.ngroups;
  if (c->kernel_timing) {
    StepArgs A;
    if (prepare_step_args(c, A, c->kev_args)) return 1;
    launch_batched_step(c, A, true);
    for (int p = 0; p < ng; ++p) {
      c->fam[p].old = c->fam[p].cur;
      c->fam[p].cur = A.end_cur[p];
    }
    return 0;
  }
  int key = 0;
  for (int p = 0; p < ng; ++p) key = key * 3 + c->fam[p].cur;
  BatchGraph* G = nullptr;
  for (auto& x : P->graphs)
    if (x.key == key) G = &x;
  if (!G) {
    P->graphs.emplace_back();
    G = &P->graphs.back();
    G->key = key;
    StepArgs A;
    if (prepare_step_args(c, A, G->d_args)) return 1;
    G->end_cur = A.end_cur;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    launch_batched_step(c, A, false);
    hipError_t e = hipStreamEndCapture(c->stream, &G->graph);
    if (e != hipSuccess) return fail(c, "hipStreamEndCapture: %s", hipGetErrorString(e));
    HIPCHK(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
  }
  HIPCHK(hipGraphLaunch(G->exec, c->stream));
  for (int p = 0; p < ng; ++p) {
    c->fam[p].old = c->fam[p].cur;
    c->fam[p].cur = G->end_cur[p];
  }
  return 0;
}

void batched_destroy(ice_gpu_ctx* c) {
  BatchedPlan* P = c->batched;
  if (!P) return;
  for (auto& g : P->groups) {
    if (g.d_items) (void)hipFree(g.d_items);
    if (g.d_qstart) (void)hipFree(g.d_qstart);
    if (g.d_ticket) (void)hipFree(g.d_ticket);
  }
  for (auto& G : P->graphs) {
    if (G.exec) (void)hipGraphExecDestroy(G.exec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    for (FamArgs* d : G.d_args) (void)hipFree(d);
  }
  delete P;
  c->batched = nullptr;
}

// the learned seed of the interface solve (same translation unit: it uses the context)
#include "surrogate.hpp"
// the residual norm without allocation (same translation unit: it uses the context)
#include "noalloc.hpp"
