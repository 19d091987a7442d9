// ORACLE -- test infrastructure only (see oracle_physics.hpp).  A CPU
// restatement of ICE's Explicit_Step (src/lib/numerics/time/explicit/
// Mod_Explicit.f90:11-245) and of the set-up that feeds it, with the same
// arrays in the same Fortran layouts (variable index fastest, 2 ghost layers),
// the same loops and the same per-cell accumulation order.  OpenMP over cells
// and faces where the reference itself is order-independent.  Exposes the
// ice_gpu C-ABI's argument conventions under an orc_ prefix so one case can be
// run through both.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../include/ice_gpu.h"
#include "oracle_physics.hpp"
#include "oracle_evap.hpp"
#include "oracle_mlp.hpp"

using namespace orc;

namespace {

static const int GUIDE[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};  // Parameters_m.f90:16-21

struct Family {
  int model, solver, nc, nbase, nfrozen, frozen[ICE_NFROZEN_MAX];
  Mat mat;
  std::vector<double> prop;   // the material's property table (mat.tab points here)
  int surrogate = 0;          // 1: network seed, 2: equilibrium seed
  std::vector<float> mlp;     // the network's parameters (surrogate = 1)
};

struct Rec {
  int b, i, j, k, f, type, p, bs, is, js, ks, fs, normal_dir, nchi1, nchi2, donor0;
  double massflux, velocity, temperature, radius, alpha, beta, gfixed;
  int partner = 0, cf = 0;
  bool owner = false;
};

struct Phase {  // ICE_cond_phase_type (Advanced_Types_m.f90:46-84), padded layout
  std::vector<double> prim, prim_old, residual, source, beta, dt, tau;
  std::vector<double> source_step;  // add_stage_source's running sum (Mod_Explicit.f90:257-281)
  std::vector<double> imports;      // rates in force a coupler set (add_births), nc per cell, or empty
  std::vector<char> inj;
};

struct Blk {
  int ni, nj, nk, NI, NJ, NK;
  size_t npad;
  std::vector<double> vol, dl, M, fa[3];  // dl: 3 per cell; M: 9 per cell row-major; fa: {N,A} per face
  std::vector<double> g, gR, gG, gK, gMu, yinf;  // gas prim (5 per cell), R, gam, k, mu
  std::vector<Phase> ph;
  size_t id(int i, int j, int k) const { return (size_t)(i + 1) + (size_t)NI * ((size_t)(j + 1) + (size_t)NJ * (size_t)(k + 1)); }
};

struct BCells {  // ICE_bc_cells_type (Advanced_Types_m.f90:90-95)
  std::vector<int> first, rec;
};

}  // namespace

struct orc_ctx {
  ice_config_t cfg;
  std::vector<Family> fam;
  std::vector<Blk> blk;
  std::vector<Rec> bc;
  std::vector<int> donor;      // 4 per donor
  std::vector<double> weight;
  std::vector<std::vector<int>> grp;    // records of family p in table order (local_bc_grp)
  std::vector<BCells> bcells;           // per family
  std::vector<std::vector<int>> own;    // owner records of family p (collect_owner_records)
  std::vector<std::vector<double>> connF;  // per family: nc * nfaces
  std::vector<int> nconn;
  int iter = 0;
  double time = 0.0, dtglobal = 0.0;
  int64_t n_floored = 0;
  int bad_state = 0, bad_recon = 0;
  int bad_b = 0, bad_i = 0, bad_j = 0, bad_k = 0, bad_p = 0;
  double restot[5] = {0, 0, 0, 0, 0};
  std::vector<double> resfam;
  // count_400_inflow (Mod_Inflow400.f90:42-43), per family, 6 per block
  std::vector<std::vector<int>> i400_nstages, i400_first_iter, i400_first_faces;
  std::vector<std::vector<double>> i400_first_x, i400_largest_x, i400_line_x;
  std::vector<ice_inflow400_line_t> i400_lines;
  int64_t i400_logged = 0;
  int keep_source = 0;                       // 2: hand the step's source (hand_step_source)
  double res_threshold = 0.0, res_reference = 0.0;
  int res_stopped = 0;
  std::string err;
};

namespace {

// ---------------------------------------------------------------- set-up
// Pair_Connections (IO_BC.f90:360-510)
bool pair_connections(orc_ctx* c) {
  int ng = c->cfg.ngroups;
  int nrec = (int)c->bc.size();
  c->nconn.assign(ng, 0);
  for (int p = 1; p <= ng; ++p) {
    // map (b, f, t1, t2) -> record, through a per block face table
    std::vector<std::vector<int>> map(c->blk.size() * 6);
    auto tang = [](int f, int i, int j, int k, int& t1, int& t2) {
      if (f <= 2) { t1 = j; t2 = k; } else if (f <= 4) { t1 = i; t2 = k; } else { t1 = i; t2 = j; }
    };
    auto dims_t = [&](int b, int f, int& n1, int& n2) {
      const Blk& B = c->blk[b - 1];
      tang(f, B.ni, B.nj, B.nk, n1, n2);
    };
    for (int n = p; n <= nrec; n += ng) {
      Rec& r = c->bc[n - 1];
      if (r.type != 101 && r.type != 201) continue;
      int n1, n2, t1, t2;
      dims_t(r.b, r.f, n1, n2);
      auto& m = map[(size_t)(r.b - 1) * 6 + (r.f - 1)];
      if (m.empty()) m.assign((size_t)n1 * n2, 0);
      tang(r.f, r.i, r.j, r.k, t1, t2);
      m[(size_t)(t1 - 1) + (size_t)n1 * (t2 - 1)] = n;
    }
    for (int n = p; n <= nrec; n += ng) {
      Rec& r = c->bc[n - 1];
      if (r.type != 101 && r.type != 201) continue;
      int q = 0;
      if (r.bs >= 1 && r.bs <= (int)c->blk.size() && r.fs >= 1 && r.fs <= 6) {
        auto& m = map[(size_t)(r.bs - 1) * 6 + (r.fs - 1)];
        if (!m.empty()) {
          int n1, n2, t1, t2;
          dims_t(r.bs, r.fs, n1, n2);
          tang(r.fs, r.is, r.js, r.ks, t1, t2);
          if (t1 >= 1 && t2 >= 1 && t1 <= n1 && t2 <= n2) q = m[(size_t)(t1 - 1) + (size_t)n1 * (t2 - 1)];
        }
      }
      if (q != 0) {
        Rec& s = c->bc[q - 1];
        if (s.i != r.is || s.j != r.js || s.k != r.ks) q = 0;
      }
      if (q == 0 || q == n) { c->err = "connection record " + std::to_string(n) + " has no partner"; return false; }
      Rec& s = c->bc[q - 1];
      if (s.type != r.type || s.bs != r.b || s.fs != r.f || s.is != r.i || s.js != r.j || s.ks != r.k) {
        c->err = "connection records " + std::to_string(n) + " and " + std::to_string(q) + " do not point at each other";
        return false;
      }
      const Blk& B = c->blk[r.bs - 1];
      int across = (r.fs + 1) / 2 == 1 ? B.ni : ((r.fs + 1) / 2 == 2 ? B.nj : B.nk);
      if (across < ICE_GC) { c->err = "connection across a block thinner than gc"; return false; }
      r.partner = q;
      if (r.f % 2 == 0 && s.f % 2 == 1) r.owner = true;
      else if (r.f % 2 == 1 && s.f % 2 == 0) r.owner = false;
      else {
        int k1[5] = {r.b, r.f, r.k, r.j, r.i}, k2[5] = {s.b, s.f, s.k, s.j, s.i};
        bool less = false;
        for (int m = 0; m < 5; ++m) if (k1[m] != k2[m]) { less = k1[m] < k2[m]; break; }
        r.owner = less;
      }
    }
    int nf = 0;
    for (int n = p; n <= nrec; n += ng) {
      Rec& r = c->bc[n - 1];
      if (r.type != 101 && r.type != 201) continue;
      if (!r.owner) continue;
      ++nf;
      r.cf = nf;
      c->bc[r.partner - 1].cf = nf;
    }
    c->nconn[p - 1] = nf;
  }
  return true;
}

// build_local_bc_index, collect_boundary_cells, collect_owner_records
// (Mod_GhostExchange.f90:103-360), one rank owning every block
void build_local_index(orc_ctx* c) {
  int ng = c->cfg.ngroups;
  int nrec = (int)c->bc.size();
  c->grp.assign(ng, {});
  for (int n = 1; n <= nrec; ++n) c->grp[c->bc[n - 1].p - 1].push_back(n);
  c->bcells.assign(ng, {});
  c->own.assign(ng, {});
  c->connF.assign(ng, {});
  for (int p = 1; p <= ng; ++p) {
    BCells& cells = c->bcells[p - 1];
    for (size_t b = 1; b <= c->blk.size(); ++b) {
      const Blk& B = c->blk[b - 1];
      std::vector<std::vector<int>> at((size_t)B.ni * B.nj * B.nk);
      for (int n = 1; n <= nrec; ++n) {
        const Rec& r = c->bc[n - 1];
        if (r.b != (int)b || r.p != p || r.type == 0) continue;
        at[(size_t)(r.i - 1) + (size_t)B.ni * ((r.j - 1) + (size_t)B.nj * (r.k - 1))].push_back(n);
      }
      for (int k = 1; k <= B.nk; ++k)
        for (int j = 1; j <= B.nj; ++j)
          for (int i = 1; i <= B.ni; ++i) {
            auto& v = at[(size_t)(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (k - 1))];
            if (v.empty()) continue;
            cells.first.push_back((int)cells.rec.size() + 1);
            // records of the cell by face, stably
            std::stable_sort(v.begin(), v.end(), [&](int a, int bb) { return c->bc[a - 1].f < c->bc[bb - 1].f; });
            for (int n : v) cells.rec.push_back(n);
          }
    }
    cells.first.push_back((int)cells.rec.size() + 1);
    for (int n : c->grp[p - 1]) {
      const Rec& r = c->bc[n - 1];
      if ((r.type == 101 || r.type == 201) && r.owner) c->own[p - 1].push_back(n);
    }
    c->connF[p - 1].assign((size_t)c->fam[p - 1].nc * std::max(c->nconn[p - 1], 1), 0.0);
  }
}

// ---------------------------------------------------------------- ghosts
double* P(orc_ctx* c, int b, int p, int i, int j, int k) {
  Blk& B = c->blk[b - 1];
  return &B.ph[p - 1].prim[B.id(i, j, k) * c->fam[p - 1].nc];
}
const double* FN(const Blk& B, int d, int i, int j, int k) { return &B.fa[d][B.id(i, j, k) * 4]; }

enum { GHOST_COPY = 0, GHOST_MIRROR = 1, GHOST_INJECT = 2 };

// Lib_Ghost.f90:395-413
int ghost_rule(const Rec& r, double veln) {
  int rule = GHOST_COPY;
  if (r.type == 200) rule = GHOST_MIRROR;
  else if (r.type == 300 || r.type == 301) {
    if (veln * (double)(1 - 2 * (r.f % 2)) <= 0.0 || r.f == 3) rule = GHOST_MIRROR;
  } else if (r.type >= 401 && r.type <= 403) {
    if (!(veln * (double)(1 - 2 * (r.f % 2)) > 0.0)) {
      rule = GHOST_INJECT;
      if (r.massflux == 0.0) rule = GHOST_MIRROR;
    }
  }
  return rule;
}

// Lib_Ghost.f90:418-432
void injected_velocity(double speed, const Rec& r, const double* inward, double* v) {
  if (r.normal_dir) {
    v[0] = speed * inward[0]; v[1] = speed * inward[1]; v[2] = speed * inward[2];
  } else {
    v[0] = speed * std::cos(r.beta) * std::cos(r.alpha);
    v[1] = speed * std::cos(r.beta) * std::sin(r.alpha);
    v[2] = speed * std::sin(r.beta);
  }
}

const double* bnd_face(const Blk& B, const Rec& r) {
  int im = r.i, jm = r.j, km = r.k, fm = r.f;
  if (fm <= 2) return FN(B, 0, im - fm % 2, jm, km);
  if (fm <= 4) return FN(B, 1, im, jm - fm % 2, km);
  return FN(B, 2, im, jm, km - fm % 2);
}

// count_400_inflow (Mod_Inflow400.f90:52-128): one serial pass over family p's
// 400 records in table order, folded into the boundaries' counts; the lines it
// would print are logged (at most ICE_INFLOW400_LOG kept between two reads)
void count_400_inflow(orc_ctx* c, int p) {
  const int nb = (int)c->blk.size(), model = c->fam[p - 1].model;
  if (c->i400_nstages.empty()) {
    int ng = c->cfg.ngroups;
    c->i400_nstages.assign(ng, std::vector<int>(6 * nb, 0));
    c->i400_first_iter = c->i400_nstages;
    c->i400_first_faces = c->i400_nstages;
    c->i400_first_x.assign(ng, std::vector<double>(6 * nb, 0.0));
    c->i400_largest_x = c->i400_first_x;
    c->i400_line_x = c->i400_first_x;
  }
  std::vector<int> nface(6 * nb, 0), nabove(6 * nb, 0);
  std::vector<double> xs(6 * nb, 0.0), xa(6 * nb, 0.0);
  std::vector<double>& line_x = c->i400_line_x[p - 1];
  for (int n : c->grp[p - 1]) {
    const Rec& r = c->bc[n - 1];
    if (r.type != 400) continue;
    const Blk& B = c->blk[r.b - 1];
    const double* q = P(c, r.b, p, r.i, r.j, r.k);
    if (!(q[0] > RHO_EMPTY)) continue;
    const double* fn = bnd_face(B, r);
    double sg = (double)(1 - 2 * (r.f % 2)), normal[3] = {sg * fn[0], sg * fn[1], sg * fn[2]};
    double w = q[1] * normal[0] + q[2] * normal[1] + q[3] * normal[2];
    if (!(w < -1e-25)) continue;
    double x;
    if (model == ICE_IG) {
      x = -w * std::sqrt(q[0] / q[4]);
      if (!(x > 1e-12)) continue;
    } else if (model == ICE_AG) {
      double pnn = q[4] * (normal[0] * normal[0]) + q[7] * (normal[1] * normal[1]) + q[9] * (normal[2] * normal[2]) +
                   2.0 * (q[5] * normal[0] * normal[1] + q[6] * normal[0] * normal[2] + q[8] * normal[1] * normal[2]);
      x = -w * std::sqrt(q[0] / fmax2(pnn, 2.2250738585072014e-308));
      if (!(x > 1e-12)) continue;
    } else {
      x = -w;
      if (!(x > 1e-12 * norm2_3(q[1], q[2], q[3]))) continue;
    }
    int m = (r.b - 1) * 6 + (r.f - 1);
    nface[m] += 1;
    xs[m] = fmax2(xs[m], x);
    if (x > 10.0 * line_x[m]) {
      nabove[m] += 1;
      xa[m] = fmax2(xa[m], x);
    }
  }
  for (int m = 0; m < 6 * nb; ++m) {
    if (nface[m] == 0) continue;
    if (c->i400_nstages[p - 1][m] == 0) {
      c->i400_first_iter[p - 1][m] = c->iter;
      c->i400_first_faces[p - 1][m] = nface[m];
      c->i400_first_x[p - 1][m] = xs[m];
    }
    c->i400_nstages[p - 1][m] += 1;
    c->i400_largest_x[p - 1][m] = fmax2(c->i400_largest_x[p - 1][m], xs[m]);
    if (nabove[m] == 0) continue;
    ice_inflow400_line_t L;
    L.iter = c->iter;
    L.block = m / 6 + 1;
    L.face = m % 6 + 1;
    L.family = p;
    L.faces = nabove[m];
    L.relative = line_x[m] > 0.0 ? 1 : 0;
    L.largest = xa[m];
    if (c->i400_logged < ICE_INFLOW400_LOG) c->i400_lines.push_back(L);
    c->i400_logged += 1;
    line_x[m] = xa[m];
  }
}

// Lib_Ghost.f90:439-466
void ghost_chimera(orc_ctx* c, const Rec& r, int g) {
  int p = r.p, nv = c->fam[p - 1].nc;
  const Family& F = c->fam[p - 1];
  int c1, c2;
  if (g == 1) { c1 = 1; c2 = r.nchi1; } else { c1 = r.nchi1 + 1; c2 = r.nchi1 + r.nchi2; }
  double consg[ICE_NCOND_MAX] = {0};
  for (int cc = c1; cc <= c2; ++cc) {
    const int* d = &c->donor[(size_t)(r.donor0 + cc - 1) * 4];
    double cons[ICE_NCOND_MAX];
    p2c(F.model, P(c, d[0], p, d[1], d[2], d[3]), cons, F.mat);
    double w = c->weight[r.donor0 + cc - 1];
    for (int v = 0; v < nv; ++v) consg[v] = consg[v] + w * cons[v];
  }
  int ig = r.i - g * GUIDE[r.f - 1][0], jg = r.j - g * GUIDE[r.f - 1][1], kg = r.k - g * GUIDE[r.f - 1][2];
  c2p(F.model, consg, P(c, r.b, p, ig, jg, kg), F.mat);
}

// Lib_Ghost.f90:76-267
void ghost_record(orc_ctx* c, int n) {
  const Rec& r = c->bc[n - 1];
  int p = r.p;
  const Family& F = c->fam[p - 1];
  int nc = F.nc, nb = F.nbase;
  Blk& B = c->blk[r.b - 1];
  int ig = r.i - GUIDE[r.f - 1][0], jg = r.j - GUIDE[r.f - 1][1], kg = r.k - GUIDE[r.f - 1][2];
  double* gh = P(c, r.b, p, ig, jg, kg);
  const double* in = P(c, r.b, p, r.i, r.j, r.k);
  switch (r.type) {
    case 101: case 201: {
      const double* s = P(c, r.bs, p, r.is, r.js, r.ks);
      for (int v = 0; v < nc; ++v) gh[v] = s[v];
      break;
    }
    case 200: case 300: case 301: {
      const double* N = bnd_face(B, r);
      double nrm[3] = {N[0], N[1], N[2]};
      for (int v = 0; v < nc; ++v) gh[v] = in[v];
      double vel[3] = {in[1], in[2], in[3]};
      double veln = vel[0] * nrm[0] + vel[1] * nrm[1] + vel[2] * nrm[2];
      if (ghost_rule(r, veln) == GHOST_MIRROR) {
        for (int m = 0; m < 3; ++m) vel[m] = vel[m] - 2.0 * veln * nrm[m];
        if (nb == 12) mirror_tensor(&in[4], nrm, &gh[4]);
      }
      gh[1] = vel[0]; gh[2] = vel[1]; gh[3] = vel[2];
      break;
    }
    case 0: case 400:
      for (int v = 0; v < nc; ++v) gh[v] = in[v];
      break;
    case 401: case 402: case 403: {
      const double* N = bnd_face(B, r);
      double nrm[3] = {N[0], N[1], N[2]};
      double vel[3] = {in[1], in[2], in[3]};
      double veln = vel[0] * nrm[0] + vel[1] * nrm[1] + vel[2] * nrm[2];
      int rule = ghost_rule(r, veln);
      B.ph[p - 1].inj[B.id(ig, jg, kg)] = (rule == GHOST_INJECT);
      if (rule == GHOST_COPY) {
        for (int v = 0; v < nc; ++v) gh[v] = in[v];
        return;
      }
      if (rule == GHOST_MIRROR) {
        for (int v = 0; v < nc; ++v) gh[v] = in[v];
        for (int m = 0; m < 3; ++m) vel[m] = vel[m] - 2.0 * veln * nrm[m];
        gh[1] = vel[0]; gh[2] = vel[1]; gh[3] = vel[2];
        if (nb == 12) mirror_tensor(&in[4], nrm, &gh[4]);
      } else {
        for (int v = 0; v < nc; ++v) gh[v] = 0.0;
        double sgn = (double)(2 * (r.f % 2) - 1);
        double inward[3] = {sgn * nrm[0], sgn * nrm[1], sgn * nrm[2]};
        size_t ci = B.id(r.i, r.j, r.k);
        if (r.type == 401) {
          const double* gp = &B.g[ci * 5];
          double vg = gp[1] * inward[0] + gp[2] * inward[1] + gp[3] * inward[2];
          injected_velocity(r.velocity * vg, r, inward, &gh[1]);
          gh[0] = r.massflux * gp[0] / r.velocity;
          double vn = std::fabs(gh[1] * nrm[0] + gh[2] * nrm[1] + gh[3] * nrm[2]);
          if (r.gfixed > 0.0 && vn > 0.0) gh[0] = gh[0] + r.massflux * r.gfixed / vn;
          gh[nb - 2] = r.temperature * gp[4];
        } else if (r.type == 402) {
          injected_velocity(r.velocity, r, inward, &gh[1]);
          double vn = gh[1] * nrm[0] + gh[2] * nrm[1] + gh[3] * nrm[2];
          gh[0] = r.massflux / std::fabs(vn);
          gh[nb - 2] = r.temperature;
        } else {
          const double* gp = &B.g[ci * 5];
          double vg = norm2_3(gp[1], gp[2], gp[3]);
          injected_velocity(r.velocity * vg, r, inward, &gh[1]);
          double vn = gh[1] * nrm[0] + gh[2] * nrm[1] + gh[3] * nrm[2];
          gh[0] = r.massflux / std::fabs(vn);
          gh[nb - 2] = r.temperature;
        }
        // N from rho and the inlet radius, with the material density at the inlet
        // temperature (Lib_Ghost.f90:230-232; radius**3._I4 is a real power)
        gh[nb - 1] = gh[0] / mat_rho(F.mat, gh[nb - 2]) / (4.0 / 3.0 * PI * pow_libm(r.radius, 3.0));
        if (F.model == IG) gh[4] = 1e-6;
        else if (F.model == AG) { gh[4] = 1e-6; gh[7] = 1e-6; gh[9] = 1e-6; }
      }
      break;
    }
    case 102:
      ghost_chimera(c, r, 1);
      break;
    default:
      break;
  }
}

// Lib_Ghost.f90:312-387
void second_ghost_record(orc_ctx* c, int n) {
  const Rec& r = c->bc[n - 1];
  if (r.type == 0) return;
  int p = r.p;
  const Family& F = c->fam[p - 1];
  int nc = F.nc, nb = F.nbase;
  Blk& B = c->blk[r.b - 1];
  const int* G = GUIDE[r.f - 1];
  int ig = r.i - G[0], jg = r.j - G[1], kg = r.k - G[2];
  int ig2 = r.i - 2 * G[0], jg2 = r.j - 2 * G[1], kg2 = r.k - 2 * G[2];
  int ip = r.i + G[0], jp = r.j + G[1], kp = r.k + G[2];
  double* g2 = P(c, r.b, p, ig2, jg2, kg2);
  switch (r.type) {
    case 101: case 201: {
      const int* Gs = GUIDE[r.fs - 1];
      const double* s = P(c, r.bs, p, r.is + Gs[0], r.js + Gs[1], r.ks + Gs[2]);
      for (int v = 0; v < nc; ++v) g2[v] = s[v];
      break;
    }
    case 102:
      ghost_chimera(c, r, 2);
      break;
    case 200: case 300: case 301: case 401: case 402: case 403: {
      const double* N = bnd_face(B, r);
      double nrm[3] = {N[0], N[1], N[2]};
      const double* in = P(c, r.b, p, r.i, r.j, r.k);
      double veln = in[1] * nrm[0] + in[2] * nrm[1] + in[3] * nrm[2];
      int rule = ghost_rule(r, veln);
      if (rule == GHOST_INJECT) {
        const double* g1 = P(c, r.b, p, ig, jg, kg);
        for (int v = 0; v < nc; ++v) g2[v] = g1[v];
      } else {
        const double* q = P(c, r.b, p, ip, jp, kp);
        for (int v = 0; v < nc; ++v) g2[v] = q[v];
        if (rule == GHOST_MIRROR) {
          double vel[3] = {q[1], q[2], q[3]};
          double d = vel[0] * nrm[0] + vel[1] * nrm[1] + vel[2] * nrm[2];
          for (int m = 0; m < 3; ++m) g2[1 + m] = vel[m] - 2.0 * d * nrm[m];
          if (nb == 12) mirror_tensor(&q[4], nrm, &g2[4]);
        }
      }
      break;
    }
    default: {  // 400: 2nd-order extrapolation 3 g1 - 3 m + m+1
      const double* g1 = P(c, r.b, p, ig, jg, kg);
      const double* m = P(c, r.b, p, r.i, r.j, r.k);
      const double* q = P(c, r.b, p, ip, jp, kp);
      for (int v = 0; v < nc; ++v) g2[v] = 3.0 * g1[v] - 3.0 * m[v] + q[v];
      break;
    }
  }
}

// ---------------------------------------------------------------- fluxes
// Mod_Fluxes.f90:599-652: face between (i,j,k) and its upper neighbour in d
bool face_flux(orc_ctx* c, int b, int p, int d, int i, int j, int k, double* flux) {
  const Family& F = c->fam[p - 1];
  Blk& B = c->blk[b - 1];
  Phase& ph = B.ph[p - 1];
  const int D[3] = {d == 0, d == 1, d == 2};
  const double* st[4];
  double len[4];
  for (int s = 0; s < 4; ++s) {
    int ii = i + (s - 1) * D[0], jj = j + (s - 1) * D[1], kk = k + (s - 1) * D[2];
    st[s] = P(c, b, p, ii, jj, kk);
    len[s] = B.dl[B.id(ii, jj, kk) * 3 + d];
  }
  double bf = fmin2(ph.beta[B.id(i, j, k)], ph.beta[B.id(i + D[0], j + D[1], k + D[2])]);
  int idx = d == 0 ? i : (d == 1 ? j : k);
  int nd = d == 0 ? B.ni : (d == 1 ? B.nj : B.nk);
  if (idx == 1 && ph.inj[B.id(i - D[0], j - D[1], k - D[2])]) len[0] = 0.0;
  if (idx == nd - 1 && ph.inj[B.id(i + 2 * D[0], j + 2 * D[1], k + 2 * D[2])]) len[3] = 0.0;
  const double* fn = FN(B, d, i, j, k);
  double nrm[3] = {fn[0], fn[1], fn[2]};
  return compute_flux_(F.model, F.solver, c->cfg.limiter, st, len, nrm, fn[3], bf, flux, F.mat);
}

// Mod_BC_Fluxes.f90:180-272
bool record_flux(orc_ctx* c, int p, int n, double* flux, const double* beta_face) {
  const Rec& r = c->bc[n - 1];
  const Family& F = c->fam[p - 1];
  int nc = F.nc;
  Blk& B = c->blk[r.b - 1];
  Phase& ph = B.ph[p - 1];
  const int* G = GUIDE[r.f - 1];
  int i = r.i, j = r.j, k = r.k;
  int ig = i - G[0], jg = j - G[1], kg = k - G[2];
  int ig2 = i - 2 * G[0], jg2 = j - 2 * G[1], kg2 = k - 2 * G[2];
  int ip = i + G[0], jp = j + G[1], kp = k + G[2];
  int dir = (r.f + 1) / 2 - 1;
  double beta_val = beta_face ? *beta_face : ph.beta[B.id(i, j, k)];
  double ql[ICE_NCOND_MAX], qr[ICE_NCOND_MAX], fl[ICE_NCOND_MAX];
  bool ok;
  if (r.f % 2 == 1) {
    const double* fn = FN(B, dir, ig, jg, kg);
    double nrm[3] = {fn[0], fn[1], fn[2]};
    double area = fn[3];
    double dl_g1 = B.dl[B.id(ig, jg, kg) * 3 + dir];
    double dl_m = B.dl[B.id(i, j, k) * 3 + dir];
    double dl_4 = B.dl[B.id(ip, jp, kp) * 3 + dir];
    double dl0 = dl_g1, dl1 = 0.5 * (dl_g1 + dl_m), dl2 = 0.5 * (dl_m + dl_4);
    double dll = 0.5 * dl_g1, dlr = 0.5 * dl_m;
    if (ph.inj[B.id(ig, jg, kg)]) { dl1 = 0.5 * dl_m; dll = 0.0; }
    if (ph.inj[B.id(ip, jp, kp)]) dl2 = 0.5 * dl_m;
    ok = reconstruct(F.model, c->cfg.limiter, P(c, r.b, p, ig2, jg2, kg2), P(c, r.b, p, ig, jg, kg),
                     P(c, r.b, p, i, j, k), P(c, r.b, p, ip, jp, kp), dl0, dl1, dl2, dll, dlr, beta_val, ql, qr);
    riemann(F.solver, F.model, ql, qr, nrm, fl, F.mat);
    for (int v = 0; v < nc; ++v) flux[v] = fl[v] * area;
  } else {
    const double* fn = FN(B, dir, i, j, k);
    double nrm[3] = {fn[0], fn[1], fn[2]};
    double area = fn[3];
    double dl_4 = B.dl[B.id(ip, jp, kp) * 3 + dir];
    double dl_m = B.dl[B.id(i, j, k) * 3 + dir];
    double dl_g1 = B.dl[B.id(ig, jg, kg) * 3 + dir];
    double dl0 = 0.5 * (dl_4 + dl_m), dl1 = 0.5 * (dl_m + dl_g1), dl2 = dl_g1;
    double dll = 0.5 * dl_m, dlr = 0.5 * dl_g1;
    if (ph.inj[B.id(ig, jg, kg)]) { dl1 = 0.5 * dl_m; dlr = 0.0; }
    if (ph.inj[B.id(ip, jp, kp)]) dl0 = 0.5 * dl_m;
    ok = reconstruct(F.model, c->cfg.limiter, P(c, r.b, p, ip, jp, kp), P(c, r.b, p, i, j, k),
                     P(c, r.b, p, ig, jg, kg), P(c, r.b, p, ig2, jg2, kg2), dl0, dl1, dl2, dll, dlr, beta_val, ql, qr);
    riemann(F.solver, F.model, ql, qr, nrm, fl, F.mat);
    for (int v = 0; v < nc; ++v) flux[v] = -(fl[v] * area);
  }
  return ok;
}

// compute_bound (Mod_BC_Fluxes.f90:27-100), owner_flux (:151-170), bound_cell (:105-144)
void compute_bound(orc_ctx* c, int p) {
  const Family& F = c->fam[p - 1];
  int nc = F.nc;
  std::vector<double>& CF = c->connF[p - 1];
  const std::vector<int>& own = c->own[p - 1];
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : bad)
  for (int jj = 0; jj < (int)own.size(); ++jj) {
    int n = own[jj];
    const Rec& r = c->bc[n - 1];
    const Rec& q = c->bc[r.partner - 1];
    Blk& Bq = c->blk[q.b - 1];
    Blk& Bn = c->blk[r.b - 1];
    double bq = Bq.ph[p - 1].beta[Bq.id(q.i, q.j, q.k)];
    double bn = Bn.ph[p - 1].beta[Bn.id(r.i, r.j, r.k)];
    double bf = fmin2(bn, bq);
    double fl[ICE_NCOND_MAX];
    if (!record_flux(c, p, n, fl, &bf)) bad = 1;
    for (int v = 0; v < nc; ++v) CF[(size_t)(r.cf - 1) * nc + v] = fl[v];
  }
  const BCells& cells = c->bcells[p - 1];
  int ncell = (int)cells.first.size() - 1;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : bad)
  for (int cc = 0; cc < ncell; ++cc) {
    for (int rr = cells.first[cc]; rr < cells.first[cc + 1]; ++rr) {
      int n = cells.rec[rr - 1];
      const Rec& r = c->bc[n - 1];
      if (r.type == 200 && F.model == MK) continue;
      double fl[ICE_NCOND_MAX];
      if (r.type == 101 || r.type == 201) {
        for (int v = 0; v < nc; ++v) {
          double x = CF[(size_t)(r.cf - 1) * nc + v];
          fl[v] = r.owner ? x : -x;
        }
      } else {
        if (!record_flux(c, p, n, fl, nullptr)) bad = 1;
      }
      Blk& B = c->blk[r.b - 1];
      double* res = &B.ph[p - 1].residual[B.id(r.i, r.j, r.k) * nc];
      for (int v = 0; v < nc; ++v) res[v] = res[v] + fl[v];
    }
  }
  if (bad) c->bad_recon = 1;
}

// compute_flux/tile_sweep (Mod_Fluxes.f90:159-243, 421-521): each face once,
// each cell its two faces per direction in the odd-then-even order
void compute_flux(orc_ctx* c, int p) {
  const Family& F = c->fam[p - 1];
  int nc = F.nc;
  int bad = 0;
  for (size_t b = 1; b <= c->blk.size(); ++b) {
    Blk& B = c->blk[b - 1];
    std::vector<double> fx[3];
    const int nd[3] = {B.ni, B.nj, B.nk};
    for (int d = 0; d < 3; ++d) {
      if (nd[d] < 2) continue;
      // face index (i,j,k) for i < ni in direction d, stored on the cell below
      fx[d].assign(B.npad * nc, 0.0);
      int ie = B.ni - (d == 0), je = B.nj - (d == 1), ke = B.nk - (d == 2);
#pragma omp parallel for collapse(2) schedule(static) reduction(| : bad)
      for (int k = 1; k <= ke; ++k)
        for (int j = 1; j <= je; ++j)
          for (int i = 1; i <= ie; ++i)
            if (!face_flux(c, (int)b, p, d, i, j, k, &fx[d][B.id(i, j, k) * nc])) bad = 1;
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i) {
          double* r = &B.ph[p - 1].residual[B.id(i, j, k) * nc];
          const int idx[3] = {i, j, k};
          for (int d = 0; d < 3; ++d) {
            if (nd[d] < 2) continue;
            const int D[3] = {d == 0, d == 1, d == 2};
            const double* up = &fx[d][B.id(i, j, k) * nc];
            const double* lo = &fx[d][B.id(i - D[0], j - D[1], k - D[2]) * nc];
            int m = idx[d], n = nd[d];
            if (m % 2 == 1) {
              if (m < n) for (int v = 0; v < nc; ++v) r[v] = r[v] - up[v];
              if (m > 1) for (int v = 0; v < nc; ++v) r[v] = r[v] + lo[v];
            } else {
              for (int v = 0; v < nc; ++v) r[v] = r[v] + lo[v];
              if (m < n) for (int v = 0; v < nc; ++v) r[v] = r[v] - up[v];
            }
          }
        }
  }
  if (bad) c->bad_recon = 1;
}

// ---------------------------------------------------------------- sources
// Mod_Sources.f90:465-622 with cell_surface (:230-274), ITC surface
// (Lib_Drop_Surface.f90:15-43) and no evaporation model
// compute_source_ (Mod_Sources.f90:465-622) with cell_surface (:228-267) and an
// isothermal drop (drop_surface, Lib_Drop_Surface.f90:15-43: Ts = T without P2T
// or boiling), constant material properties and no combustion.  `yinf` replaces
// ep(iYinf) when the far-field vapour fraction is given per cell.
void source_cell(orc_ctx* c, const Family& F, const double* q, const double* gp, double R, double gam, double gk,
                 double mu, double yinf, double* tau, double* src) {
  int n = F.nbase;
  double rho_mat = mat_rho(F.mat, q[n - 2]);
  double Rp = std::pow(0.75 * q[0] / (q[n - 1] * PI * rho_mat), 1.0 / 3.0);
  double slip = norm2_3(gp[1] - q[1], gp[2] - q[2], gp[3] - q[3]);
  double Re = 2.0 * gp[0] * Rp * slip / mu;
  double Ma = slip / std::sqrt(gam * R * gp[4]);
  double Pr = gam / (gam - 1.0) * R * mu / gk;
  double Nu = heat(Re, Pr, Ma, c->cfg.heat_select);
  double Ts = q[n - 2];
  double Tr = Ts / gp[4];
  double Cd = drag(Re, Ma, gam, Tr, c->cfg.drag_select);
  double force[6];
  force[1] = gp[1] - q[1]; force[2] = gp[2] - q[2]; force[3] = gp[3] - q[3];
  double Qconv = 2.0 * Nu * gk * PI * Rp * (gp[4] - Ts);
  double mdot = 0.0, Qevap = 0.0;
  bool ovr = false;
  double ep[8];
  for (int m = 0; m < 8; ++m) ep[m] = F.mat.ep[m];
  if (c->cfg.var_yinf) ep[EP_YINF] = yinf;
  // the table's Lv, Cpv and Psat at the surface temperature (Mod_Sources.f90:517-535)
  if (F.mat.flags & ICE_PROP_USE_LV) ep[EP_LV] = table_value(F.mat, ICE_PROP_LV, Ts);
  if (F.mat.flags & ICE_PROP_USE_CPV) ep[EP_CPV] = table_value(F.mat, ICE_PROP_CPV, Ts);
  double psat_tab = 0.0;
  const double* psat_ext = nullptr;
  if (F.mat.flags & ICE_PROP_USE_PSAT) {
    psat_tab = table_value(F.mat, ICE_PROP_PSAT, Ts);
    psat_ext = &psat_tab;
  }
  double seed = 0.0;
  if (F.surrogate == 2) {
    seed = 1.0;
  } else if (F.surrogate == 1 && !F.mlp.empty() && F.mat.evap != 0 && F.mat.intf == 1) {
    // the network's rate ratio, clamped as the device clamps it
    float x[MLP_IN];
    mlp_features(gp[4], gp[0] * R * gp[4], ep[EP_YINF], Ts, 2.0 * Rp, Re, x);
    MlpCache mc;
    mlp_forward(F.mlp.data(), x, mc);
    float r = mc.out[0];
    seed = r < 1e-2f ? 1e-2f : (r > 1.0f ? 1.0f : r);
  }
  evaporation(gp[0], gp[4], gam, R, mu, gk, Ts, 2.0 * Rp, Re, F.mat.evap, F.mat.intf, ep, mdot, Qevap, ovr, nullptr,
              seed, psat_ext);
  if (ovr) Qconv = Qevap;
  else if (c->cfg.blow_select == 1)
    Qconv = Qconv * blowing_factor(gam, R, mu, gk, rho_mat, 2.0 * Rp, 4.0 / 3.0 * PI * (Rp * Rp * Rp) * rho_mat, mdot);
  force[0] = -mdot * q[n - 1];
  force[4] = Qconv * q[n - 1];
  force[4] = force[4] + mat_emiss(F.mat, Ts) * SIGMA_SB * 4.0 * PI * Rp * Rp * q[n - 1] *
                            (std::pow(gp[4], 4.0) - std::pow(Ts, 4.0));
  if (Cd == 0.0) *tau = std::numeric_limits<double>::infinity();
  else *tau = 8.0 * rho_mat * Rp / (3.0 * gp[0] * Cd * slip + TAU_TOLL);
  force[5] = *tau;
  source_make(F.model, q, force, src, F.mat);
}

void compute_source(orc_ctx* c, int p) {
  const Family& F = c->fam[p - 1];
  int nc = F.nc;
  for (auto& B : c->blk) {
    Phase& ph = B.ph[p - 1];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i) {
          size_t id = B.id(i, j, k);
          double* s = &ph.source[id * nc];
          source_cell(c, F, &ph.prim[id * nc], &B.g[id * 5], B.gR[id], B.gG[id], B.gK[id], B.gMu[id], B.yinf[id],
                      &ph.tau[id], s);
          for (int f = 0; f < F.nfrozen; ++f) s[F.frozen[f] - 1] = 0.0;
        }
  }
}

// ---------------------------------------------------------------- dt
// compute_dt_ (Mod_dt.f90:256-275)
void compute_dt_(const Family& F, int ndir, const double* q, const double* M, double& dt) {
  double rate = 0.0;
  for (int d = 0; d < ndir; ++d) {
    const double* row = &M[d * 3];
    double gn = norm2_3(row[0], row[1], row[2]);
    double vs[3] = {row[0] / gn, row[1] / gn, row[2] / gn};
    double un = q[1] * vs[0] + q[2] * vs[1] + q[3] * vs[2];
    rate = rate + (std::fabs(un) + wavespeed(F.model, q, vs)) * gn;
  }
  if (rate > 0.0) dt = fmin2(dt, 1.0 / rate);
}

// compute_dt (Mod_dt.f90:11-112)
double compute_dt(orc_ctx* c, int p) {
  const Family& F = c->fam[p - 1];
  const ice_config_t& cf = c->cfg;
  bool tau_limit = cf.coupled && cf.tau_factor > 0.0;
  double dtmin = std::numeric_limits<double>::max();
  for (auto& B : c->blk) {
    Phase& ph = B.ph[p - 1];
#pragma omp parallel for collapse(2) schedule(static) reduction(min : dtmin)
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i) {
          size_t id = B.id(i, j, k);
          double dt = 1e+5;
          compute_dt_(F, cf.ndir, &ph.prim[id * F.nc], &B.M[id * 9], dt);
          dt = dt * cf.cfl;
          if (c->iter <= cf.cfl_ramp_iter) dt = dt * (double)c->iter / (double)cf.cfl_ramp_iter;
          dt = fmin2(dt, cf.dt_max);
          if (tau_limit) {
            double tau = ph.tau[id];
            if (ph.prim[id * F.nc] > RHO_EMPTY && !not_finite(tau) && tau > 0.0) dt = fmin2(dt, cf.tau_factor * tau);
          }
          ph.dt[id] = dt;
          dtmin = fmin2(dtmin, dt);
        }
  }
  return dtmin;
}

// ---------------------------------------------------------------- IRS
// Lib_IRS.f90:34-86, one direction d of one block
void irs_block(Blk& B, int nc, int d, double beta_irs, std::vector<double>& res) {
  int n[3] = {B.ni, B.nj, B.nk};
  // residual_star over 0..n+1 in every direction (padded storage), ghosts fixed
  std::vector<double> star(res.size(), 0.0), rnew(res.size(), 0.0);
  for (int k = 1; k <= B.nk; ++k)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i)
        for (int v = 0; v < nc; ++v) star[B.id(i, j, k) * nc + v] = res[B.id(i, j, k) * nc + v];
  const int D[3] = {d == 0, d == 1, d == 2};
  for (int k = 1; k <= B.nk; ++k)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i) {
        int idx[3] = {i, j, k};
        if (idx[d] == 1)
          for (int v = 0; v < nc; ++v) star[B.id(i - D[0], j - D[1], k - D[2]) * nc + v] = star[B.id(i, j, k) * nc + v];
        if (idx[d] == n[d])
          for (int v = 0; v < nc; ++v) star[B.id(i + D[0], j + D[1], k + D[2]) * nc + v] = star[B.id(i, j, k) * nc + v];
      }
  for (int sjb = 0; sjb < 2; ++sjb) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i) {
          size_t id = B.id(i, j, k);
          size_t i1 = B.id(i - D[0], j - D[1], k - D[2]), i2 = B.id(i + D[0], j + D[1], k + D[2]);
          for (int v = 0; v < nc; ++v)
            rnew[id * nc + v] = (res[id * nc + v] + beta_irs * (star[i1 * nc + v] + star[i2 * nc + v])) /
                                (1.0 + 2.0 * beta_irs);
        }
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i)
          for (int v = 0; v < nc; ++v) star[B.id(i, j, k) * nc + v] = rnew[B.id(i, j, k) * nc + v];
  }
  for (int k = 1; k <= B.nk; ++k)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i)
        for (int v = 0; v < nc; ++v) res[B.id(i, j, k) * nc + v] = rnew[B.id(i, j, k) * nc + v];
}

// ---------------------------------------------------------------- the step
void one_step(orc_ctx* c) {
  const ice_config_t& cf = c->cfg;
  int ng = cf.ngroups;
  c->iter += 1;
  c->dtglobal = 1e+5;
  for (int p = 1; p <= ng; ++p) c->dtglobal = fmin2(c->dtglobal, compute_dt(c, p));
  if (cf.time_accurate) c->time = c->time + c->dtglobal;
  // set_dt_global and state_copy (Lib_Newstate.f90:19-60)
  for (auto& B : c->blk)
    for (int p = 1; p <= ng; ++p) {
      Phase& ph = B.ph[p - 1];
      int nc = c->fam[p - 1].nc;
      for (int k = 1; k <= B.nk; ++k)
        for (int j = 1; j <= B.nj; ++j)
          for (int i = 1; i <= B.ni; ++i) {
            size_t id = B.id(i, j, k);
            if (cf.time_accurate) ph.dt[id] = c->dtglobal;
            for (int v = 0; v < nc; ++v) ph.prim_old[id * nc + v] = ph.prim[id * nc + v];
          }
    }
  for (int p = 1; p <= ng; ++p) {
    const Family& F = c->fam[p - 1];
    int nc = F.nc;
    for (int srk = 1; srk <= cf.nrk; ++srk) {
      for (auto& B : c->blk) {
        Phase& ph = B.ph[p - 1];
        for (int k = 1; k <= B.nk; ++k)
          for (int j = 1; j <= B.nj; ++j)
            for (int i = 1; i <= B.ni; ++i)
              for (int v = 0; v < nc; ++v) ph.residual[B.id(i, j, k) * nc + v] = 0.0;
      }
      if (cf.coupled) {
        compute_source(c, p);
        // add_stage_source with RK_accumulate (Mod_Explicit.f90:257-281, Lib_RK.f90:33-44)
        if (c->keep_source == 2) {
          double w = rkc(cf.nrk, srk);
          for (auto& B : c->blk) {
            Phase& ph = B.ph[p - 1];
            if (ph.source_step.size() != ph.source.size()) ph.source_step.assign(ph.source.size(), 0.0);
            for (size_t m = 0; m < ph.source.size(); ++m)
              ph.source_step[m] = srk == 1 ? ph.source[m] : (ph.source_step[m] + ph.source[m]) * w;
          }
        }
      }
      {
        const std::vector<int>& g = c->grp[p - 1];
#pragma omp parallel for schedule(dynamic, 64)
        for (int ii = 0; ii < (int)g.size(); ++ii) ghost_record(c, g[ii]);
#pragma omp parallel for schedule(dynamic, 64)
        for (int ii = 0; ii < (int)g.size(); ++ii) second_ghost_record(c, g[ii]);
      }
      count_400_inflow(c, p);
      // compute_beta (Mod_Fluxes.f90:71-105, 366-389)
      for (auto& B : c->blk) {
        Phase& ph = B.ph[p - 1];
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= B.nk; ++k)
          for (int j = 1; j <= B.nj; ++j)
            for (int i = 1; i <= B.ni; ++i) {
              size_t id = B.id(i, j, k);
              if (cf.shock_detector) {
                auto rho = [&](int a, int bb, int cc) { return ph.prim[B.id(a, bb, cc) * nc]; };
                ph.beta[id] = sd_rho(rho(i, j, k), rho(i - 1, j, k), rho(i + 1, j, k), rho(i, j - 1, k),
                                     rho(i, j + 1, k), rho(i, j, k - 1), rho(i, j, k + 1));
              } else {
                ph.beta[id] = 1.0;
              }
            }
      }
      compute_bound(c, p);
      compute_flux(c, p);
      // compute_residual (Lib_Residual.f90:46-82, 147-156)
      for (auto& B : c->blk) {
        Phase& ph = B.ph[p - 1];
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= B.nk; ++k)
          for (int j = 1; j <= B.nj; ++j)
            for (int i = 1; i <= B.ni; ++i) {
              size_t id = B.id(i, j, k);
              double* r = &ph.residual[id * nc];
              const double* s = &ph.source[id * nc];
              double vol = B.vol[id], dt = ph.dt[id];
              for (int v = 0; v < nc; ++v) r[v] = (r[v] / vol + (cf.coupled ? s[v] : 0.0)) * dt;
              // add_births (Lib_Residual.f90:164-210)
              if (!ph.imports.empty())
                for (int v = 0; v < nc; ++v) r[v] = r[v] + ph.imports[id * nc + v] * dt;
              for (int f = 0; f < F.nfrozen; ++f) r[F.frozen[f] - 1] = 0.0;
            }
      }
      if (cf.irs)
        for (int d = 0; d < cf.ndir; ++d)
          for (auto& B : c->blk) irs_block(B, nc, d, cf.irs_beta, B.ph[p - 1].residual);
      // state_update (Lib_Newstate.f90:63-173)
      for (size_t b = 1; b <= c->blk.size(); ++b) {
        Blk& B = c->blk[b - 1];
        Phase& ph = B.ph[p - 1];
        int64_t nfl = 0;
        int bad = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : nfl) reduction(| : bad)
        for (int k = 1; k <= B.nk; ++k)
          for (int j = 1; j <= B.nj; ++j)
            for (int i = 1; i <= B.ni; ++i) {
              size_t id = B.id(i, j, k);
              double co[ICE_NCOND_MAX], cu[ICE_NCOND_MAX];
              double* q = &ph.prim[id * nc];
              p2c(F.model, &ph.prim_old[id * nc], co, F.mat);
              p2c(F.model, q, cu, F.mat);
              const double* r = &ph.residual[id * nc];
              double w = rkc(cf.nrk, srk);
              for (int v = 0; v < nc; ++v) {
                double rr = (cu[v] - co[v] + r[v]) * w;
                cu[v] = co[v] + rr;
              }
              if (cu[0] < 0.0) ++nfl;
              c2p(F.model, cu, q, F.mat);
              if (!check_prim(F.model, q)) bad = 1;
            }
        c->n_floored += nfl;
        if (bad && !c->bad_state) {
          c->bad_state = 1;
          for (int k = 1; k <= B.nk && !c->bad_b; ++k)
            for (int j = 1; j <= B.nj && !c->bad_b; ++j)
              for (int i = 1; i <= B.ni && !c->bad_b; ++i)
                if (!check_prim(F.model, &ph.prim[B.id(i, j, k) * nc])) {
                  c->bad_b = (int)b; c->bad_i = i; c->bad_j = j; c->bad_k = k; c->bad_p = p;
                }
        }
      }
    }
  }
  // residual_norm (Mod_Diagnostic.f90:22-69, 106-127): per (block, family)
  // sums of squared changes of rho, u, v, w, T, cells in k, j, i order; the
  // totals over blocks then families, the family sums over blocks
  size_t nbk = c->blk.size();
  std::vector<double> part(5 * nbk * ng, 0.0);
  for (size_t b = 0; b < nbk; ++b) {
    Blk& B = c->blk[b];
    for (int p = 1; p <= ng; ++p) {
      const Family& F = c->fam[p - 1];
      int nc = F.nc;
      Phase& ph = B.ph[p - 1];
      double* pt = &part[5 * (b * ng + (p - 1))];
      for (int k = 1; k <= B.nk; ++k)
        for (int j = 1; j <= B.nj; ++j)
          for (int i = 1; i <= B.ni; ++i) {
            size_t id = B.id(i, j, k);
            const double* q = &ph.prim[id * nc];
            const double* o = &ph.prim_old[id * nc];
            double rn[5];
            rn[0] = std::fabs(q[0] - o[0]);
            for (int m = 1; m <= 3; ++m) rn[m] = std::fabs(q[m] - o[m]);
            rn[4] = std::fabs(q[F.nbase - 2] - o[F.nbase - 2]);
            for (int m = 0; m < 5; ++m) pt[m] = pt[m] + rn[m] * rn[m];
          }
    }
  }
  c->resfam.assign(5 * ng, 0.0);
  for (int m = 0; m < 5; ++m) c->restot[m] = 0.0;
  for (size_t b = 0; b < nbk; ++b)
    for (int p = 1; p <= ng; ++p)
      for (int m = 0; m < 5; ++m) {
        double x = part[5 * (b * ng + (p - 1)) + m];
        c->restot[m] = c->restot[m] + x;
        c->resfam[(p - 1) * 5 + m] = c->resfam[(p - 1) * 5 + m] + x;
      }
  for (int m = 0; m < 5; ++m) c->restot[m] = std::sqrt(c->restot[m]);
  for (auto& x : c->resfam) x = std::sqrt(x);
}

}  // namespace

// ================================================================= C API
extern "C" {

int orc_create(int /*device*/, const ice_config_t* cfg, const ice_family_t* fams, int nb, const int32_t* dims, orc_ctx** out) {
  orc_ctx* c = new orc_ctx();
  c->cfg = *cfg;
  for (int p = 0; p < cfg->ngroups; ++p) {
    Family F;
    F.model = fams[p].closure;
    F.solver = fams[p].riemann;
    F.nc = ncond_of(F.model);
    F.nbase = F.nc;
    F.nfrozen = fams[p].nfrozen;
    for (int m = 0; m < ICE_NFROZEN_MAX; ++m) F.frozen[m] = fams[p].frozen[m];
    F.mat.rho = fams[p].rho_mat; F.mat.cs = fams[p].cs; F.mat.lv = fams[p].lv; F.mat.emiss = fams[p].emiss;
    for (int m = 0; m < 8; ++m) F.mat.ep[m] = fams[p].ep[m];
    F.mat.evap = fams[p].evap_select; F.mat.intf = fams[p].intf_select;
    F.surrogate = fams[p].surrogate;
    c->fam.push_back(F);
  }
  const double nan = std::numeric_limits<double>::quiet_NaN();
  const double inf = std::numeric_limits<double>::infinity();
  for (int b = 0; b < nb; ++b) {
    Blk B;
    B.ni = dims[3 * b]; B.nj = dims[3 * b + 1]; B.nk = dims[3 * b + 2];
    B.NI = B.ni + 4; B.NJ = B.nj + 4; B.NK = B.nk + 4;
    B.npad = (size_t)B.NI * B.NJ * B.NK;
    B.vol.assign(B.npad, nan); B.dl.assign(B.npad * 3, nan); B.M.assign(B.npad * 9, nan);
    for (int d = 0; d < 3; ++d) B.fa[d].assign(B.npad * 4, nan);
    B.g.assign(B.npad * 5, nan); B.gR.assign(B.npad, nan); B.gG.assign(B.npad, nan);
    B.gK.assign(B.npad, nan); B.gMu.assign(B.npad, nan); B.yinf.assign(B.npad, nan);
    for (int p = 0; p < cfg->ngroups; ++p) {
      Phase ph;
      int nc = c->fam[p].nc;
      ph.prim.assign(B.npad * nc, nan); ph.prim_old.assign(B.npad * nc, nan);
      ph.residual.assign(B.npad * nc, 0.0); ph.source.assign(B.npad * nc, 0.0);
      ph.beta.assign(B.npad, 1.0); ph.dt.assign(B.npad, 0.0); ph.tau.assign(B.npad, inf);
      ph.inj.assign(B.npad, 0);
      B.ph.push_back(std::move(ph));
    }
    c->blk.push_back(std::move(B));
  }
  *out = c;
  return 0;
}

int orc_set_geometry(orc_ctx* c, int b, const double* vol, const double* dl, const double* M, const double* fi,
                     const double* fj, const double* fk) {
  Blk& B = c->blk[b - 1];
  int ni = B.ni, nj = B.nj, nk = B.nk;
  for (int k = 1; k <= nk; ++k)
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) {
        size_t s = (size_t)(i - 1) + (size_t)ni * ((j - 1) + (size_t)nj * (k - 1));
        B.vol[B.id(i, j, k)] = vol[s];
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) B.M[B.id(i, j, k) * 9 + r * 3 + cc] = M[s * 9 + r + 3 * cc];
      }
  for (int k = 0; k <= nk + 1; ++k)
    for (int j = 0; j <= nj + 1; ++j)
      for (int i = 0; i <= ni + 1; ++i) {
        size_t s = (size_t)i + (size_t)(ni + 2) * (j + (size_t)(nj + 2) * k);
        for (int m = 0; m < 3; ++m) B.dl[B.id(i, j, k) * 3 + m] = dl[s * 3 + m];
      }
  for (int k = 1; k <= nk; ++k)
    for (int j = 1; j <= nj; ++j)
      for (int i = 0; i <= ni; ++i) {
        size_t s = (size_t)i + (size_t)(ni + 1) * ((j - 1) + (size_t)nj * (k - 1));
        for (int m = 0; m < 4; ++m) B.fa[0][B.id(i, j, k) * 4 + m] = fi[s * 4 + m];
      }
  for (int k = 1; k <= nk; ++k)
    for (int j = 0; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) {
        size_t s = (size_t)(i - 1) + (size_t)ni * (j + (size_t)(nj + 1) * (k - 1));
        for (int m = 0; m < 4; ++m) B.fa[1][B.id(i, j, k) * 4 + m] = fj[s * 4 + m];
      }
  for (int k = 0; k <= nk; ++k)
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) {
        size_t s = (size_t)(i - 1) + (size_t)ni * ((j - 1) + (size_t)nj * k);
        for (int m = 0; m < 4; ++m) B.fa[2][B.id(i, j, k) * 4 + m] = fk[s * 4 + m];
      }
  return 0;
}

int orc_set_boundaries(orc_ctx* c, int nrec, const int32_t* ri, const double* rd, int ndonor, const int32_t* donor,
                       const double* weight) {
  c->bc.assign(nrec, Rec());
  for (int n = 0; n < nrec; ++n) {
    const int32_t* a = &ri[(size_t)n * ICE_REC_NI];
    const double* d = &rd[(size_t)n * ICE_REC_ND];
    Rec& r = c->bc[n];
    r.b = a[ICE_R_B]; r.i = a[ICE_R_I]; r.j = a[ICE_R_J]; r.k = a[ICE_R_K]; r.f = a[ICE_R_F];
    r.type = a[ICE_R_TYPE]; r.p = a[ICE_R_P]; r.bs = a[ICE_R_BS]; r.is = a[ICE_R_IS]; r.js = a[ICE_R_JS];
    r.ks = a[ICE_R_KS]; r.fs = a[ICE_R_FS]; r.normal_dir = a[ICE_R_NORMAL_DIR];
    r.nchi1 = a[ICE_R_NCHI1]; r.nchi2 = a[ICE_R_NCHI2]; r.donor0 = a[ICE_R_DONOR0];
    r.massflux = d[ICE_RD_MASSFLUX]; r.velocity = d[ICE_RD_VELOCITY]; r.temperature = d[ICE_RD_TEMPERATURE];
    r.radius = d[ICE_RD_RADIUS]; r.alpha = d[ICE_RD_ALPHA]; r.beta = d[ICE_RD_BETA]; r.gfixed = d[ICE_RD_GFIXED];
  }
  c->donor.assign(donor, donor + (size_t)ndonor * 4);
  c->weight.assign(weight, weight + ndonor);
  if (!pair_connections(c)) return 1;
  build_local_index(c);
  return 0;
}

int orc_get_bc_maps(orc_ctx* c, int32_t* partner, int32_t* owner, int32_t* cf, int p, int32_t* ncell, int32_t* first,
                    int32_t* rec, int cap) {
  for (size_t n = 0; n < c->bc.size(); ++n) {
    partner[n] = c->bc[n].partner; owner[n] = c->bc[n].owner; cf[n] = c->bc[n].cf;
  }
  const BCells& cells = c->bcells[p - 1];
  *ncell = (int)cells.first.size() - 1;
  if ((int)cells.first.size() > cap || (int)cells.rec.size() > cap) return 1;
  for (size_t m = 0; m < cells.first.size(); ++m) first[m] = cells.first[m];
  for (size_t m = 0; m < cells.rec.size(); ++m) rec[m] = cells.rec[m];
  return 0;
}

int orc_put_state(orc_ctx* c, int b, int p, const double* prim) {
  Blk& B = c->blk[b - 1];
  int nc = c->fam[p - 1].nc;
  std::memcpy(B.ph[p - 1].prim.data(), prim, B.npad * nc * sizeof(double));
  return 0;
}

int orc_get_state(orc_ctx* c, int b, int p, double* prim) {
  Blk& B = c->blk[b - 1];
  int nc = c->fam[p - 1].nc;
  std::memcpy(prim, B.ph[p - 1].prim.data(), B.npad * nc * sizeof(double));
  return 0;
}

int orc_put_gas(orc_ctx* c, int b, const double* gprim, const double* R, const double* gam, const double* k,
                const double* mu) {
  Blk& B = c->blk[b - 1];
  for (int kk = 1; kk <= B.nk; ++kk)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i) {
        size_t s = (size_t)(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (kk - 1));
        size_t id = B.id(i, j, kk);
        for (int m = 0; m < 5; ++m) B.g[id * 5 + m] = gprim[s * 5 + m];
        B.gR[id] = R[s]; B.gG[id] = gam[s]; B.gK[id] = k[s]; B.gMu[id] = mu[s];
      }
  return 0;
}

int orc_put_gas_yinf(orc_ctx* c, int b, const double* yinf) {
  Blk& B = c->blk[b - 1];
  for (int kk = 1; kk <= B.nk; ++kk)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i)
        B.yinf[B.id(i, j, kk)] = yinf[(size_t)(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (kk - 1))];
  return 0;
}

int orc_put_tau(orc_ctx* c, int b, int p, const double* tau) {
  Blk& B = c->blk[b - 1];
  for (int kk = 1; kk <= B.nk; ++kk)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i)
        B.ph[p - 1].tau[B.id(i, j, kk)] = tau[(size_t)(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (kk - 1))];
  return 0;
}

int orc_get_tau(orc_ctx* c, int b, int p, double* tau) {
  if (b < 1 || b > (int)c->blk.size() || p < 1 || p > c->cfg.ngroups) {
    c->err = "[ICE::get_tau] no block / family";
    return 1;
  }
  Blk& B = c->blk[b - 1];
  for (int kk = 1; kk <= B.nk; ++kk)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i)
        tau[(size_t)(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (kk - 1))] = B.ph[p - 1].tau[B.id(i, j, kk)];
  return 0;
}

int orc_step(orc_ctx* c, int nsteps, ice_status_t* st) {
  c->res_stopped = 0;
  for (int s = 0; s < nsteps; ++s) {
    one_step(c);
    if (c->bad_state || c->bad_recon) break;
    // res-threshold of a steady run (Mod_Explicit.f90:196-225)
    if (c->res_threshold > 0.0 && !c->cfg.time_accurate) {
      if (c->res_reference == 0.0) c->res_reference = c->restot[0];
      if (c->res_reference > 0.0 && c->restot[0] <= c->res_threshold * c->res_reference) {
        c->res_stopped = 1;
        break;
      }
    }
  }
  if (st) {
    st->n_floored = c->n_floored;
    st->bad_state = c->bad_state; st->bad_block = c->bad_b; st->bad_i = c->bad_i; st->bad_j = c->bad_j;
    st->bad_k = c->bad_k; st->bad_p = c->bad_p; st->bad_recon = c->bad_recon; st->iter = c->iter;
    st->dtglobal = c->dtglobal; st->time = c->time;
  }
  return (c->bad_state || c->bad_recon) ? 2 : 0;
}

// compute_ghost and fill_second_ghost (Lib_Ghost.f90:29-72, 270-308) of every family
int orc_fill_ghosts(orc_ctx* c) {
  for (int p = 1; p <= c->cfg.ngroups; ++p) {
    const std::vector<int>& g = c->grp[p - 1];
    for (int ii = 0; ii < (int)g.size(); ++ii) ghost_record(c, g[ii]);
    for (int ii = 0; ii < (int)g.size(); ++ii) second_ghost_record(c, g[ii]);
  }
  return 0;
}

// the rates in force on family p in block b (Mod_Exchange.f90:242-324), (nc, ni, nj, nk)
int orc_put_imports(orc_ctx* c, int b, int p, const double* imports) {
  if (b < 1 || b > (int)c->blk.size() || p < 1 || p > c->cfg.ngroups) {
    c->err = "[ICE::put_imports] no block / family";
    return 1;
  }
  Blk& B = c->blk[b - 1];
  Phase& ph = B.ph[p - 1];
  int nc = c->fam[p - 1].nc;
  for (auto& BB : c->blk)
    if (BB.ph[p - 1].imports.empty()) BB.ph[p - 1].imports.assign(BB.npad * nc, 0.0);
  for (int k = 1; k <= B.nk; ++k)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i) {
        size_t s = (size_t)(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (k - 1));
        for (int v = 0; v < nc; ++v) ph.imports[B.id(i, j, k) * nc + v] = imports ? imports[s * nc + v] : 0.0;
      }
  return 0;
}

// add_birth (Mod_Exchange.f90:283-300) of n packets, in their order
int orc_add_births(orc_ctx* c, int p, int n, const int32_t* cells, const double* pk) {
  if (p < 1 || p > c->cfg.ngroups) {
    c->err = "[ICE::add_births] no family";
    return 1;
  }
  const Family& F = c->fam[p - 1];
  int nc = F.nc;
  for (int m = 0; m < n; ++m) {
    const int* x = cells + 4 * (size_t)m;
    if (x[0] < 1 || x[0] > (int)c->blk.size()) {
      c->err = "[ICE::add_births] no block";
      return 1;
    }
    Blk& B = c->blk[x[0] - 1];
    if (x[1] < 1 || x[1] > B.ni || x[2] < 1 || x[2] > B.nj || x[3] < 1 || x[3] > B.nk) {
      c->err = "[ICE::add_births] cell out of its block";
      return 1;
    }
    double* q = &B.ph[p - 1].prim[B.id(x[1], x[2], x[3]) * nc];
    double u[ICE_NCOND_MAX];
    p2c(F.model, q, u, F.mat);
    const double* k = pk + 6 * (size_t)m;
    for (int v = 0; v < 4; ++v) u[v] = u[v] + k[v];
    u[nc - 2] = u[nc - 2] + k[4];
    u[nc - 1] = u[nc - 1] + k[5];
    c2p(F.model, u, q, F.mat);
  }
  return 0;
}

// the oracle has one path: the execution mode changes nothing
int orc_set_mode(orc_ctx* c, int mode) {
  if (mode != ICE_MODE_PLAIN && mode != ICE_MODE_BATCHED) {
    c->err = "[ICE::set_mode] mode is ICE_MODE_PLAIN (0) or ICE_MODE_BATCHED (1)";
    return 1;
  }
  c->cfg.mode = mode;
  return 0;
}

int orc_set_surrogate(orc_ctx* c, int p, int mode) {
  if (p < 1 || p > c->cfg.ngroups || mode < 0 || mode > 2) {
    c->err = "[ICE::set_surrogate] family or mode out of range";
    return 1;
  }
  c->fam[p - 1].surrogate = mode;
  return 0;
}

// restart counters (Wrap_Setup.f90:158-214, IO_Solution.f90:20-30)
int orc_set_counters(orc_ctx* c, int32_t iter, double time, double res_reference) {
  if (iter < 0 || !std::isfinite(time) || !(res_reference >= 0.0)) {
    c->err = "[ICE::set_counters] iter >= 0, a finite time and res_reference >= 0";
    return 1;
  }
  c->iter = iter;
  c->time = time;
  c->res_reference = res_reference;
  return 0;
}

int orc_get_counters(orc_ctx* c, int32_t* iter, double* time, double* res_reference) {
  if (iter) *iter = c->iter;
  if (time) *time = c->time;
  if (res_reference) *res_reference = c->res_reference;
  return 0;
}

int orc_get_source(orc_ctx* c, int b, int p, double* source, double* tau) {
  Blk& B = c->blk[b - 1];
  int nc = c->fam[p - 1].nc;
  for (int k = 1; k <= B.nk; ++k)
    for (int j = 1; j <= B.nj; ++j)
      for (int i = 1; i <= B.ni; ++i) {
        size_t s = (size_t)(i - 1) + (size_t)B.ni * ((j - 1) + (size_t)B.nj * (k - 1));
        size_t id = B.id(i, j, k);
        const std::vector<double>& src = c->keep_source == 2 ? B.ph[p - 1].source_step : B.ph[p - 1].source;
        for (int v = 0; v < nc; ++v) source[s * nc + v] = src[id * nc + v];
        tau[s] = B.ph[p - 1].tau[id];
      }
  return 0;
}

// Check_Multigrid (Lib_Multigrid.f90:12-35) for two contexts one level apart
static int level_pair(orc_ctx* f, orc_ctx* c) {
  if (f->cfg.ngroups != c->cfg.ngroups || f->blk.size() != c->blk.size()) {
    f->err = "[ICE::Multigrid] the levels have different families or blocks";
    return 1;
  }
  for (size_t b = 0; b < f->blk.size(); ++b) {
    const Blk &F = f->blk[b], &C = c->blk[b];
    bool ok = F.ni % 2 == 0 && F.nj % 2 == 0 && C.ni * 2 == F.ni && C.nj * 2 == F.nj &&
              (F.nk == 1 ? C.nk == 1 : (F.nk % 2 == 0 && C.nk * 2 == F.nk));
    if (!ok) {
      f->err = " [ERROR] [ICE::Check_Multigrid] block not divisible by 2^(MGL-1)";
      return 1;
    }
  }
  return 0;
}

// fine2coarse_prim (Lib_Multigrid.f90:40-109)
int orc_restrict(orc_ctx* fine, orc_ctx* coarse) {
  if (level_pair(fine, coarse)) return 1;
  for (size_t b = 0; b < fine->blk.size(); ++b) {
    const Blk& F = fine->blk[b];
    Blk& Cb = coarse->blk[b];
    for (int p = 1; p <= fine->cfg.ngroups; ++p) {
      const Family& fam = fine->fam[p - 1];
      const int nc = fam.nc;
      // fCons of every fine cell
      std::vector<double> fcons((size_t)F.ni * F.nj * F.nk * nc);
      auto fid = [&](int i, int j, int k) { return ((size_t)(i - 1) + (size_t)F.ni * ((j - 1) + (size_t)F.nj * (k - 1))) * nc; };
      for (int k = 1; k <= F.nk; ++k)
        for (int j = 1; j <= F.nj; ++j)
          for (int i = 1; i <= F.ni; ++i)
            p2c(fam.model, &F.ph[p - 1].prim[F.id(i, j, k) * nc], &fcons[fid(i, j, k)], fam.mat);
      std::vector<double> cc(nc), q(nc);
      for (int k = 1; k <= Cb.nk; ++k)
        for (int j = 1; j <= Cb.nj; ++j)
          for (int i = 1; i <= Cb.ni; ++i) {
            int i2 = 2 * i, j2 = 2 * j, k2 = 2 * k, i2d = i2 - 1, j2d = j2 - 1, k2d = k2 - 1;
            auto term = [&](int a, int bb, int cc3, int v) { return fcons[fid(a, bb, cc3) + v] * F.vol[F.id(a, bb, cc3)]; };
            for (int v = 0; v < nc; ++v) {
              if (F.nk == 1) {
                cc[v] = term(i2d, j2d, 1, v) + term(i2, j2d, 1, v) + term(i2d, j2, 1, v) + term(i2, j2, 1, v);
              } else {
                cc[v] = term(i2d, j2d, k2d, v) + term(i2, j2d, k2d, v) + term(i2d, j2, k2d, v) + term(i2, j2, k2d, v) +
                        term(i2d, j2d, k2, v) + term(i2, j2d, k2, v) + term(i2d, j2, k2, v) + term(i2, j2, k2, v);
              }
            }
            for (int v = 0; v < nc; ++v) cc[v] = cc[v] / Cb.vol[Cb.id(i, j, k)];
            c2p(fam.model, cc.data(), &Cb.ph[p - 1].prim[Cb.id(i, j, k) * nc], coarse->fam[p - 1].mat);
          }
    }
  }
  return 0;
}

// coarse2fine_prim (Lib_Multigrid.f90:113-191)
int orc_prolong(orc_ctx* coarse, orc_ctx* fine) {
  if (level_pair(fine, coarse)) return 1;
  for (size_t b = 0; b < fine->blk.size(); ++b) {
    Blk& F = fine->blk[b];
    const Blk& Cb = coarse->blk[b];
    double co[8];
    if (F.nk == 1) {
      double a1 = 9.0 / 16.0, a2 = 3.0 / 16.0, a3 = 1.0 / 16.0, a4 = 0.0;
      double t[8] = {a1, a2, a2, a4, a3, a4, a4, a4};
      std::copy(t, t + 8, co);
    } else {
      double a1 = 27.0 / 64.0, a2 = 9.0 / 64.0, a3 = 3.0 / 64.0, a4 = 1.0 / 64.0;
      double t[8] = {a1, a2, a2, a2, a3, a3, a3, a4};
      std::copy(t, t + 8, co);
    }
    static const int MASK[8][3] = {{1, 2, 3}, {4, 2, 3}, {1, 5, 3}, {4, 5, 3}, {1, 2, 6}, {4, 2, 6}, {1, 5, 6}, {4, 5, 6}};
    for (int p = 1; p <= fine->cfg.ngroups; ++p) {
      const int nc = fine->fam[p - 1].nc;
      const std::vector<double>& cp = Cb.ph[p - 1].prim;
      std::vector<double>& fp = F.ph[p - 1].prim;
      auto C = [&](int i, int j, int k, int v) { return cp[Cb.id(i, j, k) * nc + v]; };
      for (int k = 1; k <= Cb.nk; ++k)
        for (int j = 1; j <= Cb.nj; ++j)
          for (int i = 1; i <= Cb.ni; ++i) {
            int i2 = 2 * i, j2 = 2 * j, k2 = 2 * k, i2d = i2 - 1, j2d = j2 - 1, k2d = k2 - 1;
            int im = std::max(1, i - 1), ip = std::min(Cb.ni, i + 1);
            int jm = std::max(1, j - 1), jp = std::min(Cb.nj, j + 1);
            int km = std::max(1, k - 1), kp = std::min(Cb.nk, k + 1);
            if (F.nk == 1) { k2 = 1; k2d = 1; }
            int id[6] = {im, jm, km, ip, jp, kp};
            int counter = 0;
            for (int kk = k2d; kk <= k2; ++kk)
              for (int jj = j2d; jj <= j2; ++jj)
                for (int ii = i2d; ii <= i2; ++ii) {
                  const int* m = MASK[counter];
                  int a = id[m[0] - 1], bb = id[m[1] - 1], cc3 = id[m[2] - 1];
                  for (int v = 0; v < nc; ++v)
                    fp[F.id(ii, jj, kk) * nc + v] = co[0] * C(i, j, k, v) + co[1] * C(a, j, k, v) + co[2] * C(i, bb, k, v) +
                                                    co[3] * C(i, j, cc3, v) + co[4] * C(a, bb, k, v) +
                                                    co[5] * C(a, j, cc3, v) + co[6] * C(i, bb, cc3, v) +
                                                    co[7] * C(a, bb, cc3, v);
                  counter += 1;
                }
          }
    }
  }
  return 0;
}

int orc_set_keep_source(orc_ctx* c, int on) {
  c->keep_source = on == 2 ? 2 : (on ? 1 : 0);
  return 0;
}
int orc_set_res_threshold(orc_ctx* c, double threshold) {
  c->res_threshold = threshold;
  c->res_reference = 0.0;
  c->res_stopped = 0;
  return 0;
}
int orc_res_stop(orc_ctx* c, int32_t* stopped, double* reference) {
  if (stopped) *stopped = c->res_stopped;
  if (reference) *reference = c->res_reference;
  return 0;
}

int orc_residual_norm(orc_ctx* c, double* total, double* fam) {
  for (int m = 0; m < 5; ++m) total[m] = c->restot[m];
  for (size_t m = 0; m < c->resfam.size(); ++m) fam[m] = c->resfam[m];
  return 0;
}

int orc_set_properties(orc_ctx* c, int p, int t_min, int n, const double* cols, int32_t flags) {
  if (p < 1 || p > c->cfg.ngroups) { c->err = "[ICE::set_properties] family out of range"; return 1; }
  if (n < 2 || t_min < 0) { c->err = "[ICE::Load_Table] a property table needs 2 rows on T >= 0"; return 1; }
  Family& F = c->fam[p - 1];
  F.prop.assign(cols, cols + (size_t)ICE_PROP_NCOL * n);
  F.mat.tab = F.prop.data();
  F.mat.tmin = t_min;
  F.mat.tmax = t_min + n - 1;
  F.mat.flags = flags;
  F.mat.rho = cols[(size_t)ICE_PROP_RHO * n];
  F.mat.cs = cols[(size_t)ICE_PROP_CP * n];
  return 0;
}

int orc_inflow400_lines(orc_ctx* c, ice_inflow400_line_t* out, int max, int32_t* n, int64_t* dropped) {
  int take = (int)std::min<int64_t>((int64_t)c->i400_lines.size(), std::max(max, 0));
  for (int k = 0; k < take; ++k) out[k] = c->i400_lines[k];
  *n = take;
  *dropped = c->i400_logged - take;
  c->i400_lines.clear();
  c->i400_logged = 0;
  return 0;
}

// report_400_inflow's order (Mod_Inflow400.f90:156-169): family, block, face
int orc_inflow400_totals(orc_ctx* c, ice_inflow400_total_t* out, int max, int32_t* n) {
  int k = 0;
  for (size_t p = 0; p < c->i400_nstages.size(); ++p)
    for (size_t m = 0; m < c->i400_nstages[p].size(); ++m) {
      if (c->i400_nstages[p][m] == 0) continue;
      if (k >= max) {
        c->err = "[ICE::inflow400] more boundaries let particles in than the output holds";
        return 1;
      }
      ice_inflow400_total_t& o = out[k++];
      o.block = (int)(m / 6) + 1;
      o.face = (int)(m % 6) + 1;
      o.family = (int)p + 1;
      o.stages = c->i400_nstages[p][m];
      o.first_iter = c->i400_first_iter[p][m];
      o.first_faces = c->i400_first_faces[p][m];
      o.first_largest = c->i400_first_x[p][m];
      o.largest = c->i400_largest_x[p][m];
    }
  *n = k;
  return 0;
}

int orc_last_error(orc_ctx* c, char* buf, int len) {
  if (!c || len <= 0) return 1;
  std::snprintf(buf, (size_t)len, "%s", c->err.c_str());
  return 0;
}

// the network of the seeded interface solve (surrogate = 1)
int orc_surrogate_set_params(orc_ctx* c, int p, const float* params) {
  if (p < 1 || p > (int)c->fam.size()) return 1;
  c->fam[p - 1].mlp.assign(params, params + MLP_PARAMS);
  return 0;
}
int orc_mlp_grad(orc_ctx* c, int n, const float* x, const float* y, const float* params, float* grad, float* loss) {
  (void)c;
  if (n < 1) return 1;
  *loss = mlp_grad(params, n, x, y, grad);
  return 0;
}
int orc_mlp_forward(orc_ctx* c, int n, const float* x, const float* params, float* out) {
  (void)c;
  MlpCache mc;
  for (int s = 0; s < n; ++s) {
    mlp_forward(params, x + (size_t)s * MLP_IN, mc);
    for (int o = 0; o < MLP_OUT; ++o) out[(size_t)s * MLP_OUT + o] = mc.out[o];
  }
  return 0;
}
void orc_destroy(orc_ctx* c) { delete c; }

}  // extern "C"

// ================================================================= per-function entry points
// (for pinning each restated function against the reference on random inputs)
extern "C" {
static Mat g_mat;
static int g_model = IG;
static std::vector<double> g_tab;
void orc_fn_setup(int model, double rho, double cs, double lv, double emiss) {
  g_model = model; g_mat.rho = rho; g_mat.cs = cs; g_mat.lv = lv; g_mat.emiss = emiss;
  g_mat.tab = nullptr; g_mat.flags = 0;
}
// a property table for the per-function entry points (ice_gpu_set_properties' layout)
void orc_fn_setup_table(int t_min, int n, const double* cols, int flags) {
  g_tab.assign(cols, cols + (size_t)ICE_PROP_NCOL * n);
  g_mat.tab = g_tab.data(); g_mat.tmin = t_min; g_mat.tmax = t_min + n - 1; g_mat.flags = flags;
  g_mat.rho = cols[(size_t)ICE_PROP_RHO * n]; g_mat.cs = cols[(size_t)ICE_PROP_CP * n];
}
double orc_fn_table_value(int col, double T) { return table_value(g_mat, col, T); }
double orc_fn_mat_e(double T) { return mat_e(g_mat, T); }
double orc_fn_mat_T_from_e(double e) { return mat_T_from_e(g_mat, e); }
double orc_fn_mat_T_from_rhoe(double rhoe, double rho, double eps) { return mat_T_from_rhoe(g_mat, rhoe, rho, eps); }
double orc_fn_mat_rho(double T) { return mat_rho(g_mat, T); }
void orc_fn_p2c(const double* q, double* c) { p2c(g_model, q, c, g_mat); }
void orc_fn_c2p(const double* c, double* q) { c2p(g_model, c, q, g_mat); }
int orc_fn_check(const double* q) { return check_prim(g_model, q) ? 1 : 0; }
double orc_fn_wavespeed(const double* q, const double* n) { return wavespeed(g_model, q, n); }
void orc_fn_flux(const double* q, const double* n, double* f) { flux_make(g_model, q, n, f, g_mat); }
void orc_fn_source_make(const double* q, const double* F, double* s) { source_make(g_model, q, F, s, g_mat); }
void orc_fn_riemann(int solver, const double* q1, const double* q4, const double* n, double* f) {
  riemann(solver, g_model, q1, q4, n, f, g_mat);
}
double orc_fn_limiter(int id, double x, double y, double cx, double cy) { return limiter(id, x, y, cx, cy); }
int orc_fn_reconstruct(int lim, const double* a, const double* b, const double* c, const double* d, double dl0,
                       double dl1, double dl2, double dll, double dlr, double beta, double* ql, double* qr) {
  return reconstruct(g_model, lim, a, b, c, d, dl0, dl1, dl2, dll, dlr, beta, ql, qr) ? 1 : 0;
}
double orc_fn_sd(const double* r27) {
  // r27 in Fortran order (0:2,0:2,0:2): index a + 3 b + 9 c
  auto R = [&](int a, int b, int c) { return r27[a + 3 * b + 9 * c]; };
  return sd_rho(R(1, 1, 1), R(0, 1, 1), R(2, 1, 1), R(1, 0, 1), R(1, 2, 1), R(1, 1, 0), R(1, 1, 2));
}
double orc_fn_drag(double Re, double Ma, double G, double Tr, int sel) { return drag(Re, Ma, G, Tr, sel); }
double orc_fn_heat(double Re, double Pr, double Ma, int sel) { return heat(Re, Pr, Ma, sel); }
void orc_fn_rk_stage(int irk, int nrk, int nc, const double* cu, const double* co, const double* r, double* out) {
  double w = rkc(nrk, irk);
  for (int v = 0; v < nc; ++v) {
    double rr = (cu[v] - co[v] + r[v]) * w;
    out[v] = co[v] + rr;
  }
}
double orc_fn_norm2(double a, double b, double c) { return norm2_3(a, b, c); }
// residual_smoothing of one block (Lib_IRS.f90:15-31): directions 1..ndir in
// turn on res(nc, ni, nj, nk), interior only, variable fastest
void orc_fn_irs(int ni, int nj, int nk, int nc, int ndir, double beta, double* res) {
  Blk B;
  B.ni = ni; B.nj = nj; B.nk = nk; B.NI = ni + 4; B.NJ = nj + 4; B.NK = nk + 4;
  B.npad = (size_t)B.NI * B.NJ * B.NK;
  std::vector<double> r(B.npad * nc, 0.0);
  for (int k = 1; k <= nk; ++k)
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i)
        for (int v = 0; v < nc; ++v)
          r[B.id(i, j, k) * nc + v] = res[(((size_t)(k - 1) * nj + (j - 1)) * ni + (i - 1)) * nc + v];
  for (int d = 0; d < ndir; ++d) irs_block(B, nc, d, beta, r);
  for (int k = 1; k <= nk; ++k)
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i)
        for (int v = 0; v < nc; ++v)
          res[(((size_t)(k - 1) * nj + (j - 1)) * ni + (i - 1)) * nc + v] = r[B.id(i, j, k) * nc + v];
}
void orc_fn_evaporation(double rhog, double Tg, double gam, double Rg, double mug, double kg, double Tp, double dp,
                        double Re, int evap, int intf, const double* ep, double* mdot, double* qdot, int* ovr,
                        int* iters) {
  bool o = false;
  evaporation(rhog, Tg, gam, Rg, mug, kg, Tp, dp, Re, evap, intf, ep, *mdot, *qdot, o, iters);
  *ovr = o ? 1 : 0;
}
// the same with the property table's saturation pressure (psatExt)
void orc_fn_evaporation_psat(double rhog, double Tg, double gam, double Rg, double mug, double kg, double Tp, double dp,
                             double Re, int evap, int intf, const double* ep, double psat, double* mdot, double* qdot,
                             int* ovr) {
  bool o = false;
  evaporation(rhog, Tg, gam, Rg, mug, kg, Tp, dp, Re, evap, intf, ep, *mdot, *qdot, o, nullptr, 0.0, &psat);
  *ovr = o ? 1 : 0;
}
double orc_fn_blowing(double gam, double Rg, double mug, double kg, double rhop, double dp, double mp, double mdot) {
  return blowing_factor(gam, Rg, mug, kg, rhop, dp, mp, mdot);
}
}
