// Oracle (test infrastructure only): the evaporation rate of one drop and the
// Stefan-blowing reduction of its convective heat, restated from
// src/lib/physics/Lib_Evaporation.f90 operation for operation (left-to-right
// products, integer powers as products, real powers as pow, libm exp/log).
// Pinned bit for bit against the flang-built reference module (ref_evaporation,
// ref_blowing in oracle/ref_shim.f90) over every model and interface selector.
#pragma once
#include <cmath>

#include "oracle_physics.hpp"

namespace orc {

// Lib_Evaporation.f90:15-25: indices into ep (0-based here) and constants
enum { EP_MV = 0, EP_LV = 1, EP_CPV = 2, EP_LE = 3, EP_YINF = 4, EP_LVMV_RU = 5, EP_INVTB = 6, EP_ALPHAE = 7 };
static const double RU = 8314.46;
static const double PATM = 101325.0;
static const double XSCAP = 1e-12;
static const double HUGE_R8 = 1.7976931348623157e308;  // huge(1._R8)

// psat_CC :738-745
static inline double psat_cc(double T, double lvmv_ru, double invtb) {
  return PATM * std::exp(-lvmv_ru * (1.0 / T - invtb));
}
// molar2mass :265-270, mass2molar :274-279
static inline double molar2mass(double Xs, double Mv, double Mg) { return Xs * Mv / (Xs * Mv + (1.0 - Xs) * Mg); }
static inline double mass2molar(double Ys, double Mv, double Mg) { return Ys * Mg / (Ys * Mg + (1.0 - Ys) * Mv); }

// F_correction :514-524
static inline double f_correction(double B) {
  if (B > 1.0e-10) return pow_libm(1.0 + B, 0.7) * std::log(1.0 + B) / B;
  return 1.0;
}

// d2law :287-300
static inline double d2law(double Tg, double kg, double Tp, double dp, double cpg, double Lv) {
  double BT = cpg * (Tg - Tp) / Lv;
  if (BT <= 0.0) return 0.0;
  return -2.0 * PI * dp * kg / cpg * std::log(1.0 + BT);
}

// CEM_model :334-345
static inline double cem(double rhog, double kg, double dp, double cpg, double Sc, double Re05, double BM, double Le) {
  double Dv = kg / (rhog * cpg * Le);
  double Sh = 2.0 + 0.6 * Re05 * pow_libm(Sc, 1.0 / 3.0);
  return -PI * dp * rhog * Dv * Sh * std::log(1.0 + BM);
}

// CEMB_model :353-374
static inline double cemb(double rhog, double Tg, double mug, double kg, double Tp, double dp, double cpg, double Re05,
                          double BM, double Le) {
  double Tf = Tp + (Tg - Tp) / 3.0;
  double rho_f = rhog * Tg / Tf;
  double mu_f = mug * pow_libm(Tf / Tg, 0.7);
  double kl_f = kg * pow_libm(Tf / Tg, 0.7);
  double cpg_f = cpg;
  double Dv_f = kl_f / (rho_f * cpg_f * Le);
  double Sc_f = mu_f * cpg_f * Le / kl_f;
  double Re_f = Re05 * Re05 * rho_f / rhog * mug / mu_f;
  double Sh_f = 2.0 + 0.6 * std::sqrt(fmax2(Re_f, 0.0)) * pow_libm(Sc_f, 1.0 / 3.0);
  return -PI * dp * rho_f * Dv_f * Sh_f * std::log(1.0 + BM);
}

// ASM_model :382-439 (override_Qdot is set)
static inline void asm_model(double rhog, double Tg, double kg, double Tp, double dp, double cpg, double Pr, double Sc,
                             double Re05, double BM, double Le, double cpv, double& mdot, double& Qd) {
  double Dv = kg / (rhog * cpg * Le);
  double lnBM = std::log(1.0 + BM);
  double FM = f_correction(BM);
  double Sh0 = 2.0 + 0.552 * Re05 * pow_libm(Sc, 1.0 / 3.0);
  double Sh_star = 2.0 + (Sh0 - 2.0) / FM;
  mdot = -PI * dp * rhog * Dv * Sh_star * lnBM;
  double cpv_eff = cpv > 0.0 ? cpv : cpg;
  double Nu0 = 2.0 + 0.552 * Re05 * pow_libm(Pr, 1.0 / 3.0);
  double Nu_star = Nu0, BT = 0.0;
  for (int it = 1; it <= 5; ++it) {
    double phi = cpv_eff / cpg * Sh_star / Nu_star / Le;
    BT = pow_libm(1.0 + BM, phi) - 1.0;
    if (BT <= 0.0) {
      BT = BM;
      Nu_star = Nu0;
      break;
    }
    double FT = f_correction(BT);
    Nu_star = 2.0 + (Nu0 - 2.0) / FT;
  }
  Qd = BT > 0.0 ? -mdot * cpv_eff * (Tg - Tp) / BT : 0.0;
}

// TC_model :448-508
static inline void tc_model(double rhog, double Tg, double kg, double Tp, double dp, double Mg, double cpg, double Pr,
                            double Sc, double Re05, double Ys, double Le, double cpv, double Yinf, double Mv,
                            double& mdot, double& Qd, bool& ovr) {
  mdot = 0.0;
  Qd = 0.0;
  ovr = false;
  double Dv = kg / (rhog * cpg * Le);
  double cpv_eff = cpv > 0.0 ? cpv : cpg;
  double Xs = fmin2(mass2molar(Ys, Mv, Mg), 1.0 - XSCAP);
  double Xinf = mass2molar(Yinf, Mv, Mg);
  double Minf = Xinf * Mv + (1.0 - Xinf) * Mg;
  double rhs0 = Mv / Minf * std::log((1.0 - Xinf) / (1.0 - Xs));
  if (rhs0 <= 0.0) return;
  double Tts = Tp / Tg;
  double Lev = kg / (cpv_eff * Dv * rhog);
  double mlo = rhs0 / fmax2(Tts, 1.0);
  double mhi = rhs0 / fmin2(fmax2(Tts, 1.0e-3), 1.0);
  double mhat = rhs0;
  for (int it = 1; it <= 30; ++it) {
    double x = mhat / Lev, f, fp;
    if (x > 1.0e-6) {
      double E = std::exp(-x);
      f = x / (1.0 - E);
      fp = (1.0 - E - x * E) / ((1.0 - E) * (1.0 - E));
    } else {
      f = 1.0 + 0.5 * x + x * x / 12.0;
      fp = 0.5 + x / 6.0;
    }
    double G = mhat + (Tts - 1.0) * Lev * (f - 1.0) - rhs0;
    double dG = 1.0 + (Tts - 1.0) * fp;
    if (std::fabs(G) <= 1.e-13 * fmax2(1.0, rhs0)) break;
    if (G > 0.0) mhi = mhat;
    else mlo = mhat;
    mhat = mhat - G / dG;
    if (mhat <= mlo || mhat >= mhi) mhat = 0.5 * (mlo + mhi);
  }
  double Sh = 2.0 + 0.6 * Re05 * pow_libm(Sc, 1.0 / 3.0);
  mdot = -PI * dp * rhog * Dv * Sh * mhat;
  ovr = true;
  double Nu = 2.0 + 0.6 * Re05 * pow_libm(Pr, 1.0 / 3.0);
  double chi = fmin2(-mdot * cpv_eff / (PI * dp * kg * Nu), 700.0);
  if (chi > 1.e-9) Qd = -mdot * cpv_eff * (Tg - Tp) / (std::exp(chi) - 1.0);
  else Qd = PI * dp * kg * Nu * (Tg - Tp);
}

// the film properties evaporation() shares with its gas-side models (host association, :168-203)
struct EvapState {
  double rhog, Tg, gamma, Rg, mug, kg, Tp, dp, cpg, p, Mg, Pr, Sc, Re05;
  int evap;
  const double* ep;
};

// gasSideRate :207-226
static inline void gas_side_rate(const EvapState& s, double Ys, double BM, double& md, double& Qd, bool& ovr) {
  md = 0.0;
  Qd = 0.0;
  ovr = false;
  const double* ep = s.ep;
  switch (s.evap) {
    case 1: md = d2law(s.Tg, s.kg, s.Tp, s.dp, s.cpg, ep[EP_LV]); break;
    case 2: md = cem(s.rhog, s.kg, s.dp, s.cpg, s.Sc, s.Re05, BM, ep[EP_LE]); break;
    case 3: md = cemb(s.rhog, s.Tg, s.mug, s.kg, s.Tp, s.dp, s.cpg, s.Re05, BM, ep[EP_LE]); break;
    case 4:
      asm_model(s.rhog, s.Tg, s.kg, s.Tp, s.dp, s.cpg, s.Pr, s.Sc, s.Re05, BM, ep[EP_LE], ep[EP_CPV], md, Qd);
      ovr = true;
      break;
    case 5:
      tc_model(s.rhog, s.Tg, s.kg, s.Tp, s.dp, s.Mg, s.cpg, s.Pr, s.Sc, s.Re05, Ys, ep[EP_LE], ep[EP_CPV],
               ep[EP_YINF], ep[EP_MV], md, Qd, ovr);
      break;
    default: break;
  }
}

// lkCorrection :230-259: Picard on mdot, damped when not contracting; `iters`
// returns the gas-side evaluations made (for the surrogate's statistics)
static inline void lk_correction(const EvapState& s, double XsEq, double& md, double& Qd, bool& ovr, int* iters) {
  const double* ep = s.ep;
  double LKd2 = 2.0 * s.mug * std::sqrt(2.0 * PI * s.Tp * RU / ep[EP_MV]) / (ep[EP_ALPHAE] * s.Sc * s.p) / s.dp;
  double diffPrev = HUGE_R8;
  double mdNew = 0.0, QdNew = 0.0;
  bool ovrNew = false;
  for (int it = 1; it <= 30; ++it) {
    if (iters) *iters = it;
    double beta = -md * s.Pr / (2.0 * PI * s.mug * s.dp);
    double XsN = fmax2(XsEq - LKd2 * beta, 0.0);
    double YsN = molar2mass(XsN, ep[EP_MV], s.Mg);
    mdNew = 0.0;
    QdNew = 0.0;
    ovrNew = false;
    if (YsN > ep[EP_YINF]) {
      double BMN = (YsN - ep[EP_YINF]) / (1.0 - YsN);
      if (BMN > 0.0) gas_side_rate(s, YsN, BMN, mdNew, QdNew, ovrNew);
    }
    double diff = std::fabs(mdNew - md);
    if (diff <= 1.e-12 * std::fabs(mdNew)) {
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

// One Picard map of lkCorrection (:236-244): the gas-side rate at the surface
// state that the rate md implies
static inline void lk_map(const EvapState& s, double XsEq, double LKd2, double md, double& mdNew, double& QdNew,
                          bool& ovrNew) {
  const double* ep = s.ep;
  double beta = -md * s.Pr / (2.0 * PI * s.mug * s.dp);
  double XsN = fmax2(XsEq - LKd2 * beta, 0.0);
  double YsN = molar2mass(XsN, ep[EP_MV], s.Mg);
  mdNew = 0.0;
  QdNew = 0.0;
  ovrNew = false;
  if (YsN > ep[EP_YINF]) {
    double BMN = (YsN - ep[EP_YINF]) / (1.0 - YsN);
    if (BMN > 0.0) gas_side_rate(s, YsN, BMN, mdNew, QdNew, ovrNew);
  }
}

// The converged solve of the surrogate path (no counterpart in the reference):
// G(m) = map(m) - m decreases strictly on [md_eq, 0], G(md_eq) >= 0 > G(0), so
// that interval brackets its root; secant steps from the seed m0 and its
// Picard image, kept inside the bracket (else its midpoint), converged by
// lkCorrection's own test; every second step that has not halved the bracket
// is replaced by a bisection; false after SEED_EVALS evaluations.
static const int SEED_EVALS = 64;
static inline bool lk_secant(const EvapState& s, double XsEq, double LKd2, double md_eq, double m0, double& md,
                             double& Qd, bool& ovr, int& evals) {
  double lo = md_eq, hi = 0.0;
  double f0, f1, Q;
  bool o;
  lk_map(s, XsEq, LKd2, m0, f0, Q, o);
  evals = 1;
  double g0 = f0 - m0;
  if (std::fabs(g0) <= 1.e-12 * std::fabs(f0)) {
    md = f0; Qd = Q; ovr = o;
    return true;
  }
  if (g0 > 0.0) lo = m0;
  else hi = m0;
  double m1 = f0;
  if (!(m1 > lo && m1 < hi)) m1 = 0.5 * (lo + hi);
  double wref = hi - lo;
  int since = 0;
  while (evals < SEED_EVALS) {
    lk_map(s, XsEq, LKd2, m1, f1, Q, o);
    ++evals;
    double g1 = f1 - m1;
    if (std::fabs(g1) <= 1.e-12 * std::fabs(f1)) {
      md = f1; Qd = Q; ovr = o;
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
    m0 = m1; g0 = g1; m1 = m2;
  }
  return false;
}

// evaporation :145-203 (psat_ext: the property table's psatExt): mdot [kg/s] <= 0, Qd [W], ovr
// when the model's gas-side heat replaces the Nusselt one
static inline void evaporation(double rhog, double Tg, double gamma, double Rg, double mug, double kg, double Tp,
                               double dp, double Re, int evap, int intf, const double* ep, double& mdot, double& Qd,
                               bool& ovr, int* iters = nullptr, double seed = 0.0, const double* psat_ext = nullptr) {
  mdot = 0.0;
  Qd = 0.0;
  ovr = false;
  if (iters) *iters = 0;
  if (evap == 0) return;
  EvapState s;
  s.rhog = rhog; s.Tg = Tg; s.gamma = gamma; s.Rg = Rg; s.mug = mug; s.kg = kg; s.Tp = Tp; s.dp = dp;
  s.evap = evap; s.ep = ep;
  s.cpg = gamma * Rg / (gamma - 1.0);
  s.p = rhog * Rg * Tg;
  s.Mg = RU / Rg;
  double psat = psat_ext ? *psat_ext : psat_cc(Tp, ep[EP_LVMV_RU], ep[EP_INVTB]);
  double Xs = psat >= s.p ? 1.0 - XSCAP : fmin2(psat / s.p, 1.0 - XSCAP);
  double Ys = molar2mass(Xs, ep[EP_MV], s.Mg);
  if (Ys <= ep[EP_YINF]) return;
  double BM = (Ys - ep[EP_YINF]) / (1.0 - Ys);
  if (BM <= 0.0) return;
  s.Re05 = std::sqrt(fmax2(Re, 0.0));
  s.Pr = mug * s.cpg / kg;
  s.Sc = s.Pr * ep[EP_LE];
  gas_side_rate(s, Ys, BM, mdot, Qd, ovr);
  if (intf != 1) return;
  if (seed > 0.0) {
    // the surrogate path: seeded secant, the plain solve when it does not converge
    double LKd2 = 2.0 * s.mug * std::sqrt(2.0 * PI * s.Tp * RU / ep[EP_MV]) / (ep[EP_ALPHAE] * s.Sc * s.p) / s.dp;
    double md, Q;
    bool o;
    int ev = 0;
    bool ok = lk_secant(s, Xs, LKd2, mdot, seed * mdot, md, Q, o, ev);
    int ev2 = 0;
    if (ok) {
      mdot = md; Qd = Q; ovr = o;
      if (iters) *iters = ev + ev2;
      return;
    }
  }
  lk_correction(s, Xs, mdot, Qd, ovr, iters);
}

// blowingFactor :129-143
static inline double blowing_factor(double gamma, double Rg, double mug, double kg, double rhop, double dp, double mp,
                                    double mdot) {
  if (mdot >= 0.0 || mp <= 0.0 || mug <= 0.0 || kg <= 0.0) return 1.0;
  if (dp <= 0.0 || rhop <= 0.0 || gamma <= 1.0) return 1.0;
  double cpg = gamma * Rg / (gamma - 1.0);
  double Pr = mug * cpg / kg;
  double taud = rhop * dp * dp / (18.0 * mug);
  double b = -1.5 * Pr * taud * mdot / mp;
  if (b != b || b <= 1.e-10) return 1.0;
  if (b > 500.0) return 0.0;
  return b / (std::exp(b) - 1.0);
}

}  // namespace orc
