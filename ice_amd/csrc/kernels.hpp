// Kernels of one RK stage of ICE's explicit path (Mod_Explicit.f90:105-177) on
// gfx950.  The state of a family lives in HBM as structure-of-arrays: slot v of
// padded cell g at prim[v * N + g], where N is the padded cell count of all
// blocks and block b's cell (i, j, k), Fortran 1-based with i in [-1, ni+2],
// is g = off_b + (i+1) + NI*((j+1) + NJ*(k+1)), NI = ni+4, NJ = nj+4.
// Geometry and gas use the same index, so every stream is unit-stride in i.
#pragma once
#include "dev_physics.hpp"
#include "dev_evap.hpp"
#include "dev_mlp.hpp"

namespace ice {

constexpr int kNFrozenMax = 6;  // frozen slots of a family: v, w and AG's P12, P13, P23 (Mod_Freeze.f90:117-128)
constexpr int kGuide[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

struct BlkDev {
  int ni, nj, nk, NI, NJ, NK;
  long long off;   // padded offset of the block
  long long cells; // ni*nj*nk
  // bit d: direction d has one cell and only planar (0) faces, so its ghosts
  // are copies of the cell, read by nothing but the sensor, whose term for the
  // direction is then exactly 0 (|(r - 2r + r)/(4r + 1e-30)|)
  int flat;
  int pad;
};

// A boundary record (ICE_bc_type) with its pairing
struct DRec {
  int b, i, j, k, f, type, p, bs, is, js, ks, fs, normal_dir, nchi1, nchi2, donor0, partner, owner, cf, pad;
  double massflux, velocity, temperature, radius, alpha, beta, gfixed, pad2;
};

struct DevStatus {
  unsigned long long dtbits;  // running min of the step's dt (positive doubles order as their bits)
  double dtglobal, time;
  unsigned long long n_floored;
  int bad_state, bad_recon;
  int bad_b, bad_i, bad_j, bad_k, bad_p;
  int iter;                   // grid%iter of the step being computed
  int lk_overflow;            // a deferred interface solve found its family's queue full
};

// Everything a kernel of one family reads or writes
struct FamArgs {
  const BlkDev* blk;
  long long N;
  // geometry
  const double* vol;    // [N]
  const double* dl;     // [3][N]
  const double* M;      // [9][N] row-major entries
  const double* fa;     // [3 dirs][4][N]: N1 N2 N3 A of the face above the cell in each direction
  // gas
  const double* gas;    // [5][N]
  const double* gR;
  const double* gG;
  const double* gK;
  const double* gMu;
  const double* yinf;
  // state
  const double* cur;    // [nc][N]
  const double* old;
  double* nxt;
  double* prim_w;       // ghost kernels write the current buffer through this
  double* beta;         // [N]
  double* dt;           // [N] steady runs
  double* tau;          // [N]
  uint8_t* inj;         // [N]
  double* src;          // [nc][N] kept source (optional)
  const double* imports;  // [nc][N] rates in force a coupler set (births, absorptions), or null
  // cells whose Langmuir-Knudsen interface solve is deferred to k_lk_finish (full
  // waves of such cells): count, padded index, block and flux sum [nc][cap] of each;
  // null when the family has no such solve
  unsigned* lkq_n;
  long long* lkq_g;
  int* lkq_b;
  double* lkq_r;
  int lkq_cap;
  double* R;            // [nc][N] materialised residual (IRS)
  double* T;            // [nc][N] IRS buffer: the directions ping-pong between R and T
  double* bsum;         // [nbcell][nc] boundary-flux sums
  const int* bidx;      // [N] boundary cell of a surface cell, -1 none
  double* connF;        // [nconn][nc]
  // records
  const DRec* rec;
  const int* donor;     // [ndonor][4]
  const double* weight;
  const int* list;      // record list of the launch (family records / owner records)
  int nlist;
  const int* bfirst;    // boundary cells: records bfirst[c]..bfirst[c+1]-1 of brec (0-based)
  const int* brec;
  const long long* bcell_g;  // padded index of each boundary cell
  int nbcell;
  DevStatus* st;
  MatDev mat;
  int limiter, nrk, srk, time_accurate, coupled, drag_sel, heat_sel, blow_sel;
  int nfrozen, frozen[kNFrozenMax];
  int keep_source, last_stage, var_yinf, iter, cfl_ramp_iter, ndir, fam;
  double cfl, dt_max, tau_factor, irs_beta;
  // learned seed of the interface solve (mat.surrogate = 1) and its counters
  MlpDev mlp;
  int mlp_on;
  unsigned long long* lk_stats;  // [4] or nullptr
};

// The material of family a as a launch sees it: with tabulated properties
// possible (TAB), or as constant properties (MatConst), for the launches that
// serve only families without a table.
template <bool TAB> struct MatOf { using type = MatDev; };
template <> struct MatOf<false> { using type = MatConst; };
template <bool TAB> ICE_DEV const typename MatOf<TAB>::type& mat_of(const FamArgs& a) {
  return static_cast<const typename MatOf<TAB>::type&>(a.mat);
}

ICE_DEV long long gidx(const BlkDev& B, int i, int j, int k) {
  return B.off + (long long)(i + 1) + (long long)B.NI * ((long long)(j + 1) + (long long)B.NJ * (long long)(k + 1));
}

template <int NC> ICE_DEV void load_cell(const double* p, long long N, long long g, double* q) {
#pragma unroll
  for (int v = 0; v < NC; ++v) q[v] = p[(long long)v * N + g];
}
template <int NC> ICE_DEV void store_cell(double* p, long long N, long long g, const double* q) {
#pragma unroll
  for (int v = 0; v < NC; ++v) p[(long long)v * N + g] = q[v];
}

// ------------------------------------------------------------ interior face
// Face between padded cell g and g + stride (direction d), its lower cell at
// index m (1-based) of n along d (Mod_Fluxes.f90:599-685); its flux times the
// area is added to (MODE 0), subtracted from (MODE 1) or stored in (MODE 2) r.
template <int M, int S, bool SD, int MODE, int LIM = -1, bool TAB = true>
ICE_DEV bool face_acc(const FamArgs& a, long long g, long long stride, int d, int m, int n, double* r) {
  constexpr int NC = Closure<M>::NC;
  const long long N = a.N;
  const double* dl = a.dl + (long long)d * N;
  double len0 = dl[g - stride], len1 = dl[g], len2 = dl[g + stride], len3 = dl[g + 2 * stride];
  double bf = SD ? dmin2(a.beta[g], a.beta[g + stride]) : 1.0;
  if (m == 1 && a.inj[g - stride]) len0 = 0.0;
  if (m == n - 1 && a.inj[g + 2 * stride]) len3 = 0.0;
  const double* fa = a.fa + (long long)(d * 4) * N;
  double nrm[3] = {fa[g], fa[N + g], fa[2 * N + g]};
  double area = fa[3 * N + g];
  double dl0 = 0.5 * (len0 + len1);
  double dl1 = 0.5 * (len1 + len2);
  double dl2 = 0.5 * (len2 + len3);
  double dll = 0.5 * len1;
  double dlr = 0.5 * len2;
  double ql[NC], qr[NC];
  bool ok = reconstruct_mem<M, LIM>(a.limiter, a.cur + (g - stride), stride, N, dl0, dl1, dl2, dll, dlr, bf, ql, qr);
  riemann_acc<M, S, MODE>(ql, qr, nrm, area, r, mat_of<TAB>(a));
  return ok;
}

// ------------------------------------------------------------ cell source
// compute_source_ (Mod_Sources.f90:465-622) with cell_surface (:230-274), the
// ITC surface (Lib_Drop_Surface.f90:15-43), evaporation with the interface and
// blowing models (Lib_Evaporation.f90), no combustion, no P2T/boiling.  With a
// per-cell far-field vapour fraction its value replaces ep(iYinf).
// The drop radius, slip speed and slip Reynolds number of compute_source_
template <int M, bool TAB = true>
ICE_DEV void drop_state(const FamArgs& a, long long g, const double* q, const double* gp, double& Rp, double& slip,
                        double& Re) {
  constexpr int NC = Closure<M>::NC;
  Rp = pow(0.75 * q[0] / (q[NC - 1] * kPi * mat_rho(mat_of<TAB>(a), q[NC - 2])), 1.0 / 3.0);
  slip = norm3(gp[1] - q[1], gp[2] - q[2], gp[3] - q[3]);
  Re = 2.0 * gp[0] * Rp * slip / a.gMu[g];
}

// The cell's ep at surface temperature Ts: the material's, the per-cell Yinf, the
// table's Lv and Cpv at Ts (Mod_Sources.f90:517-519) and its psat (kPsatX)
template <bool TAB = true> ICE_DEV void cell_ep(const FamArgs& a, long long g, double Ts, double* ep) {
  const auto& mt = mat_of<TAB>(a);
#pragma unroll
  for (int m = 0; m < 8; ++m) ep[m] = mt.ep[m];
  if (a.var_yinf) ep[kYinf] = a.yinf[g];
  if (mt.flags & kUseLv) ep[kLv] = prop_value(mt, kPropLv, Ts);
  if (mt.flags & kUseCpv) ep[kCpv] = prop_value(mt, kPropCpv, Ts);
  ep[kPsatX] = (mt.flags & kUsePsat) ? prop_value(mt, kPropPsat, Ts) : __longlong_as_double(0x7FF8000000000000ll);
}

// seed > 0: the interface solve starts from seed x (equilibrium rate), see evaporation_d
// EVAP = false: a launch whose families do not evaporate (the evaporation code is
// not compiled into it, which frees its registers).  LK = false: a launch whose
// cells needing the Langmuir-Knudsen interface solve are all deferred to
// k_lk_finish (the solve is not compiled into it).
template <int M, bool TAB = true, bool EVAP = true, bool LK = true>
ICE_DEV void cell_source(const FamArgs& a, long long g, const double* q, double* src, double& tau, float seed = 0.0f) {
  constexpr int NC = Closure<M>::NC;
  const long long N = a.N;
  double gp[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) gp[m] = a.gas[(long long)m * N + g];
  double R = a.gR[g], gam = a.gG[g], gk = a.gK[g], mu = a.gMu[g];
  const auto& mt = mat_of<TAB>(a);
  double rho_mat = mat_rho(mt, q[NC - 2]);
  double Rp = pow(0.75 * q[0] / (q[NC - 1] * kPi * rho_mat), 1.0 / 3.0);
  double slip = norm3(gp[1] - q[1], gp[2] - q[2], gp[3] - q[3]);
  double Re = 2.0 * gp[0] * Rp * slip / mu;
  double Ma = slip / sqrt(gam * R * gp[4]);
  double Pr = gam / (gam - 1.0) * R * mu / gk;
  double Nu = nusselt(Re, Pr, Ma, a.heat_sel);
  double Ts = q[NC - 2];
  double Tr = Ts / gp[4];
  double Cd = drag_cd(Re, Ma, gam, Tr, a.drag_sel);
  double F[6];
  F[1] = gp[1] - q[1];
  F[2] = gp[2] - q[2];
  F[3] = gp[3] - q[3];
  double Qconv = 2.0 * Nu * gk * kPi * Rp * (gp[4] - Ts);
  double mdot = 0.0, Qevap = 0.0;
  bool ovr = false;
  if (EVAP && mt.evap != 0) {
    double ep[kNepCell];
    cell_ep<TAB>(a, g, Ts, ep);
    evaporation_d(gp[0], gp[4], gam, R, mu, gk, Ts, 2.0 * Rp, Re, mt.evap, LK ? mt.intf : 0, ep, mdot, Qevap, ovr, seed,
                  LkStats{a.lk_stats});
  }
  if (ovr) Qconv = Qevap;
  else if (a.blow_sel == 1)
    Qconv = Qconv * blowing_factor_d(gam, R, mu, gk, rho_mat, 2.0 * Rp, 4.0 / 3.0 * kPi * (Rp * Rp * Rp) * rho_mat, mdot);
  F[0] = -mdot * q[NC - 1];
  F[4] = Qconv * q[NC - 1];
  // grey-body radiation; with a zero emissivity the term is 0 times a finite
  // product (the temperatures of a state that passed check_prim are finite), so
  // its two powers are not evaluated
  const double emiss = mat_emiss(mt, Ts);
  if (emiss != 0.0)
    F[4] = F[4] + emiss * kSigmaSB * 4.0 * kPi * Rp * Rp * q[NC - 1] * (pow(gp[4], 4.0) - pow(Ts, 4.0));
  if (Cd == 0.0) tau = __longlong_as_double(0x7FF0000000000000ll);
  else tau = 8.0 * rho_mat * Rp / (3.0 * gp[0] * Cd * slip + kTauToll);
  F[5] = tau;
  sourcemake<M>(q, F, src, mt);
}

}  // namespace ice
