// Device evaporation of one drop (src/lib/physics/Lib_Evaporation.f90): the
// Clausius-Clapeyron surface state, the gas-side models d2-law, CEM, CEM-B, ASM
// and TC, the Langmuir-Knudsen non-equilibrium interface solved by the
// reference's damped Picard iteration, and the Stefan-blowing reduction of the
// convective heat.
//
// Each model's factors that do not depend on the surface state (film
// properties, Sherwood and Nusselt numbers, the product prefix of mdot) are
// formed once per cell, before the interface iteration; they are the very
// subexpressions the reference evaluates first, left to right, so hoisting
// them changes no bit.
#pragma once
#include "dev_physics.hpp"

namespace ice {

constexpr double kRu = 8314.46;     // Lib_Evaporation.f90:21
constexpr double kPatm = 101325.0;  // :22
constexpr double kXsCap = 1e-12;    // :24
constexpr double kHuge = 1.7976931348623157e308;
// ep slots (Lib_Evaporation.f90:15-19), 0-based
enum { kMv = 0, kLv = 1, kCpv = 2, kLe = 3, kYinf = 4, kLvMvRu = 5, kInvTb = 6, kAlphaE = 7 };
// A cell's ep carries one more slot: the property table's saturation pressure at
// the surface temperature (psatExt, Mod_Sources.f90:531-535), NaN for the
// Clausius-Clapeyron curve (psat_CC, Lib_Evaporation.f90:738-745)
enum { kPsatX = 8, kNepCell = 9 };

ICE_DEV double surface_psat(const double* ep, double Tp) {
  return ep[kPsatX] == ep[kPsatX] ? ep[kPsatX] : kPatm * exp(-ep[kLvMvRu] * (1.0 / Tp - ep[kInvTb]));
}

ICE_DEV double molar2mass_d(double Xs, double Mv, double Mg) { return Xs * Mv / (Xs * Mv + (1.0 - Xs) * Mg); }
ICE_DEV double mass2molar_d(double Ys, double Mv, double Mg) { return Ys * Mg / (Ys * Mg + (1.0 - Ys) * Mv); }

// F_correction :514-524
ICE_DEV double f_correction_d(double B) {
  if (B > 1.0e-10) return pow(1.0 + B, 0.7) * log(1.0 + B) / B;
  return 1.0;
}

// A drop's gas side: the state evaporation() shares with its models, and the
// invariants of the selected model.
struct GasSide {
  double rhog, Tg, kg, Tp, dp, cpg, p, Mg, Pr, Sc, Re05, mug;
  double Mv, Yinf, Le, cpv_eff;
  int evap;
  double K;         // d2-law: mdot; CEM/CEM-B/TC: prefix of mdot; ASM: prefix before Sh*
  double Sh0, Nu0;  // ASM: Frossling numbers; TC: Sh, Nu
  double Xinf, Minf, Tts, Lev;  // TC
};

ICE_DEV void gas_side_setup(GasSide& s, const double* ep) {
  s.Mv = ep[kMv];
  s.Yinf = ep[kYinf];
  s.Le = ep[kLe];
  double cpv = ep[kCpv];
  s.cpv_eff = cpv > 0.0 ? cpv : s.cpg;
  switch (s.evap) {
    case 1: {  // d2law :287-300, independent of the surface state
      double BT = s.cpg * (s.Tg - s.Tp) / ep[kLv];
      s.K = BT <= 0.0 ? 0.0 : -2.0 * kPi * s.dp * s.kg / s.cpg * log(1.0 + BT);
      break;
    }
    case 2: {  // CEM_model :334-345
      double Dv = s.kg / (s.rhog * s.cpg * s.Le);
      double Sh = 2.0 + 0.6 * s.Re05 * pow(s.Sc, 1.0 / 3.0);
      s.K = -kPi * s.dp * s.rhog * Dv * Sh;
      break;
    }
    case 3: {  // CEMB_model :353-374
      double Tf = s.Tp + (s.Tg - s.Tp) / 3.0;
      double rho_f = s.rhog * s.Tg / Tf;
      double mu_f = s.mug * pow(Tf / s.Tg, 0.7);
      double kl_f = s.kg * pow(Tf / s.Tg, 0.7);
      double cpg_f = s.cpg;
      double Dv_f = kl_f / (rho_f * cpg_f * s.Le);
      double Sc_f = mu_f * cpg_f * s.Le / kl_f;
      double Re_f = s.Re05 * s.Re05 * rho_f / s.rhog * s.mug / mu_f;
      double Sh_f = 2.0 + 0.6 * sqrt(dmax2(Re_f, 0.0)) * pow(Sc_f, 1.0 / 3.0);
      s.K = -kPi * s.dp * rho_f * Dv_f * Sh_f;
      break;
    }
    case 4: {  // ASM_model :382-439
      double Dv = s.kg / (s.rhog * s.cpg * s.Le);
      s.Sh0 = 2.0 + 0.552 * s.Re05 * pow(s.Sc, 1.0 / 3.0);
      s.Nu0 = 2.0 + 0.552 * s.Re05 * pow(s.Pr, 1.0 / 3.0);
      s.K = -kPi * s.dp * s.rhog * Dv;
      break;
    }
    case 5: {  // TC_model :448-508
      double Dv = s.kg / (s.rhog * s.cpg * s.Le);
      s.Xinf = mass2molar_d(s.Yinf, s.Mv, s.Mg);
      s.Minf = s.Xinf * s.Mv + (1.0 - s.Xinf) * s.Mg;
      s.Tts = s.Tp / s.Tg;
      s.Lev = s.kg / (s.cpv_eff * Dv * s.rhog);
      s.Sh0 = 2.0 + 0.6 * s.Re05 * pow(s.Sc, 1.0 / 3.0);
      s.Nu0 = 2.0 + 0.6 * s.Re05 * pow(s.Pr, 1.0 / 3.0);
      s.K = -kPi * s.dp * s.rhog * Dv * s.Sh0;
      break;
    }
    default: s.K = 0.0; break;
  }
}

// gasSideRate :207-226 at surface state (Ys, BM)
ICE_DEV void gas_side_rate_d(const GasSide& s, double Ys, double BM, double& md, double& Qd, bool& ovr) {
  md = 0.0;
  Qd = 0.0;
  ovr = false;
  switch (s.evap) {
    case 1: md = s.K; break;
    case 2:
    case 3: md = s.K * log(1.0 + BM); break;
    case 4: {
      double lnBM = log(1.0 + BM);
      double FM = f_correction_d(BM);
      double Sh_star = 2.0 + (s.Sh0 - 2.0) / FM;
      md = s.K * Sh_star * lnBM;
      double Nu_star = s.Nu0, BT = 0.0;
      for (int it = 1; it <= 5; ++it) {
        double phi = s.cpv_eff / s.cpg * Sh_star / Nu_star / s.Le;
        BT = pow(1.0 + BM, phi) - 1.0;
        if (BT <= 0.0) {
          BT = BM;
          Nu_star = s.Nu0;
          break;
        }
        double FT = f_correction_d(BT);
        Nu_star = 2.0 + (s.Nu0 - 2.0) / FT;
      }
      Qd = BT > 0.0 ? -md * s.cpv_eff * (s.Tg - s.Tp) / BT : 0.0;
      ovr = true;
      break;
    }
    case 5: {
      double Xs = dmin2(mass2molar_d(Ys, s.Mv, s.Mg), 1.0 - kXsCap);
      double rhs0 = s.Mv / s.Minf * log((1.0 - s.Xinf) / (1.0 - Xs));
      if (rhs0 <= 0.0) break;
      double mlo = rhs0 / dmax2(s.Tts, 1.0);
      double mhi = rhs0 / dmin2(dmax2(s.Tts, 1.0e-3), 1.0);
      double mhat = rhs0;
      for (int it = 1; it <= 30; ++it) {
        double x = mhat / s.Lev, f, fp;
        if (x > 1.0e-6) {
          double E = exp(-x);
          f = x / (1.0 - E);
          fp = (1.0 - E - x * E) / ((1.0 - E) * (1.0 - E));
        } else {
          f = 1.0 + 0.5 * x + x * x / 12.0;
          fp = 0.5 + x / 6.0;
        }
        double G = mhat + (s.Tts - 1.0) * s.Lev * (f - 1.0) - rhs0;
        double dG = 1.0 + (s.Tts - 1.0) * fp;
        if (fabs(G) <= 1.e-13 * dmax2(1.0, rhs0)) break;
        if (G > 0.0) mhi = mhat;
        else mlo = mhat;
        mhat = mhat - G / dG;
        if (mhat <= mlo || mhat >= mhi) mhat = 0.5 * (mlo + mhi);
      }
      md = s.K * mhat;
      ovr = true;
      double chi = dmin2(-md * s.cpv_eff / (kPi * s.dp * s.kg * s.Nu0), 700.0);
      if (chi > 1.e-9) Qd = -md * s.cpv_eff * (s.Tg - s.Tp) / (exp(chi) - 1.0);
      else Qd = kPi * s.dp * s.kg * s.Nu0 * (s.Tg - s.Tp);
      break;
    }
    default: break;
  }
}

// The gas-side rate at the Langmuir-Knudsen surface state that a rate md
// implies (one Picard map of lkCorrection, :236-244).
ICE_DEV void lk_map(const GasSide& s, double XsEq, double LKd2, double md, double& mdNew, double& QdNew, bool& ovrNew) {
  double beta = -md * s.Pr / (2.0 * kPi * s.mug * s.dp);
  double XsN = dmax2(XsEq - LKd2 * beta, 0.0);
  double YsN = molar2mass_d(XsN, s.Mv, s.Mg);
  mdNew = 0.0;
  QdNew = 0.0;
  ovrNew = false;
  if (YsN > s.Yinf) {
    double BMN = (YsN - s.Yinf) / (1.0 - YsN);
    if (BMN > 0.0) gas_side_rate_d(s, YsN, BMN, mdNew, QdNew, ovrNew);
  }
}

// lkCorrection :230-259: Picard on mdot, damped by one half when it does not
// contract; `evals` returns the map evaluations made
ICE_DEV void lk_picard_counted(const GasSide& s, double XsEq, double LKd2, double& md, double& Qd, bool& ovr,
                               int& evals) {
  double diffPrev = kHuge;
  double mdNew = 0.0, QdNew = 0.0;
  bool ovrNew = false;
  for (int it = 1; it <= 30; ++it) {
    evals = it;
    lk_map(s, XsEq, LKd2, md, mdNew, QdNew, ovrNew);
    double diff = fabs(mdNew - md);
    if (diff <= 1.e-12 * fabs(mdNew)) {
      md = mdNew;
      Qd = QdNew;
      ovr = ovrNew;
      return;
    }
    if (diff < diffPrev) md = mdNew;
    else md = 0.5 * (md + mdNew);
    diffPrev = diff;
  }
  md = mdNew;
  Qd = QdNew;
  ovr = ovrNew;
}

// The surface state of evaporation() (:168-200) before the gas-side rate.
// Returns false when the drop does not evaporate (Ys <= Yinf or BM <= 0).
ICE_DEV bool evap_surface(GasSide& s, double rhog, double Tg, double gamma, double Rg, double mug, double kg,
                          double Tp, double dp, double Re, const double* ep, double& Xs, double& Ys, double& BM) {
  s.rhog = rhog; s.Tg = Tg; s.kg = kg; s.Tp = Tp; s.dp = dp; s.mug = mug;
  s.cpg = gamma * Rg / (gamma - 1.0);
  s.p = rhog * Rg * Tg;
  s.Mg = kRu / Rg;
  double psat = surface_psat(ep, Tp);
  Xs = psat >= s.p ? 1.0 - kXsCap : dmin2(psat / s.p, 1.0 - kXsCap);
  Ys = molar2mass_d(Xs, ep[kMv], s.Mg);
  if (Ys <= ep[kYinf]) return false;
  BM = (Ys - ep[kYinf]) / (1.0 - Ys);
  if (BM <= 0.0) return false;
  s.Re05 = sqrt(dmax2(Re, 0.0));
  s.Pr = mug * s.cpg / kg;
  s.Sc = s.Pr * ep[kLe];
  return true;
}

ICE_DEV double lk_factor(const GasSide& s, const double* ep) {
  return 2.0 * s.mug * sqrt(2.0 * kPi * s.Tp * kRu / ep[kMv]) / (ep[kAlphaE] * s.Sc * s.p) / s.dp;
}

// The converged solve of lkCorrection's fixed point m = map(m), map the Picard
// map of lk_map.  G(m) = map(m) - m is strictly decreasing on [md_eq, 0] (a more
// negative rate lowers the surface vapour fraction and with it the rate), with
// G(md_eq) >= 0 > G(0) = md_eq, so that interval brackets the single root.
// Secant steps from the seed m0 and its Picard image, each kept inside the
// bracket (else the bracket's midpoint), converged by lkCorrection's own test
// (the map moves m by at most 1e-12 |map(m)|), whose map evaluation is then
// the result.  A secant that keeps landing on one side of the root shrinks the
// bracket slowly: every second step that has not halved the bracket is replaced
// by a bisection, so the bracket at least halves every three evaluations.  False
// after kSeedEvals evaluations.
constexpr int kSeedEvals = 64;
ICE_DEV bool lk_secant(const GasSide& s, double XsEq, double LKd2, double md_eq, double m0, double& md, double& Qd,
                       bool& ovr, int& evals) {
  double lo = md_eq, hi = 0.0;
  double f0, f1, Q;
  bool o;
  lk_map(s, XsEq, LKd2, m0, f0, Q, o);
  evals = 1;
  double g0 = f0 - m0;
  if (fabs(g0) <= 1.e-12 * fabs(f0)) {
    md = f0;
    Qd = Q;
    ovr = o;
    return true;
  }
  if (g0 > 0.0) lo = m0;
  else hi = m0;
  double m1 = f0;
  if (!(m1 > lo && m1 < hi)) m1 = 0.5 * (lo + hi);
  double wref = hi - lo;
  int since = 0;
  while (evals < kSeedEvals) {
    lk_map(s, XsEq, LKd2, m1, f1, Q, o);
    ++evals;
    double g1 = f1 - m1;
    if (fabs(g1) <= 1.e-12 * fabs(f1)) {
      md = f1;
      Qd = Q;
      ovr = o;
      return true;
    }
    if (g1 > 0.0) lo = m1;
    else hi = m1;
    double den = g1 - g0;
    double m2 = den != 0.0 ? m1 - g1 * (m1 - m0) / den : 0.5 * (lo + hi);
    if (!(m2 > lo && m2 < hi)) m2 = 0.5 * (lo + hi);
    if (++since == 2) {
      if (hi - lo > 0.5 * wref) m2 = 0.5 * (lo + hi);
      wref = hi - lo;
      since = 0;
    }
    m0 = m1;
    g0 = g1;
    m1 = m2;
  }
  return false;
}

// Counters of the interface solves of a stage (optional): solves, map
// evaluations, fallbacks to the plain solve, largest evaluation count.
struct LkStats {
  unsigned long long* c;  // [4] or nullptr
};

// evaporation :145-203 (Clausius-Clapeyron psat).  seed <= 0: the plain solve,
// the reference's damped Picard iteration from the equilibrium rate, at most 30
// evaluations (it can stop short of the fixed point where the map's slope is
// near -1).  seed > 0, the converged solve: lk_secant from seed x (equilibrium
// rate), the plain solve when it does not converge.
ICE_DEV void evaporation_d(double rhog, double Tg, double gamma, double Rg, double mug, double kg, double Tp, double dp,
                           double Re, int evap, int intf, const double* ep, double& mdot, double& Qd, bool& ovr,
                           float seed = 0.0f, LkStats stats = LkStats{nullptr}) {
  mdot = 0.0;
  Qd = 0.0;
  ovr = false;
  if (evap == 0) return;
  GasSide s;
  s.evap = evap;
  double Xs, Ys, BM;
  if (!evap_surface(s, rhog, Tg, gamma, Rg, mug, kg, Tp, dp, Re, ep, Xs, Ys, BM)) return;
  gas_side_setup(s, ep);
  gas_side_rate_d(s, Ys, BM, mdot, Qd, ovr);
  if (intf != 1) return;
  int evals = 0;
  bool fell = false;
  if (seed > 0.0f) {
    double md = 0.0, Q = 0.0;
    bool o = false;
    if (lk_secant(s, Xs, lk_factor(s, ep), mdot, (double)seed * mdot, md, Q, o, evals)) {
      mdot = md;
      Qd = Q;
      ovr = o;
    } else {
      fell = true;
    }
  }
  if (seed <= 0.0f || fell) {
    int it = 0;
    lk_picard_counted(s, Xs, lk_factor(s, ep), mdot, Qd, ovr, it);
    evals += it;
  }
  if (stats.c) {
    atomicAdd(&stats.c[0], 1ull);
    atomicAdd(&stats.c[1], (unsigned long long)evals);
    if (fell) atomicAdd(&stats.c[2], 1ull);
    atomicMax(&stats.c[3], (unsigned long long)evals);
  }
}

// Whether evaporation() will run the interface solve for this drop (the
// surface state admits evaporation and the interface is Langmuir-Knudsen),
// and the gas pressure it uses.
ICE_DEV bool lk_needed(double rhog, double Tg, double Rg, double Tp, int evap, int intf, const double* ep, double& p) {
  p = rhog * Rg * Tg;
  if (evap == 0 || intf != 1) return false;
  double Mg = kRu / Rg;
  double psat = surface_psat(ep, Tp);
  double Xs = psat >= p ? 1.0 - kXsCap : dmin2(psat / p, 1.0 - kXsCap);
  double Ys = molar2mass_d(Xs, ep[kMv], Mg);
  if (Ys <= ep[kYinf]) return false;
  return (Ys - ep[kYinf]) / (1.0 - Ys) > 0.0;
}

// blowingFactor :129-143
ICE_DEV double blowing_factor_d(double gamma, double Rg, double mug, double kg, double rhop, double dp, double mp,
                                double mdot) {
  if (mdot >= 0.0 || mp <= 0.0 || mug <= 0.0 || kg <= 0.0) return 1.0;
  if (dp <= 0.0 || rhop <= 0.0 || gamma <= 1.0) return 1.0;
  double cpg = gamma * Rg / (gamma - 1.0);
  double Pr = mug * cpg / kg;
  double taud = rhop * dp * dp / (18.0 * mug);
  double b = -1.5 * Pr * taud * mdot / mp;
  if (b != b || b <= 1.e-10) return 1.0;
  if (b > 500.0) return 0.0;
  return b / (exp(b) - 1.0);
}

}  // namespace ice
