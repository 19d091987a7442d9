// Device arithmetic of ICE's explicit residual path for gfx950, FP64.
//
// Per-face and per-cell functions of the reference, restated for the device:
// closures (src/lib/physics/Lib_Model_{MK,IG,AG}.f90), Riemann solvers
// (src/lib/numerics/fluxes/riemann/Lib_Riemann.f90), limiters
// (src/lib/numerics/space/Lib_Limiters.f90), MUSCL reconstruction
// (src/lib/numerics/space/Lib_Reconstruction.f90:17-99), the Jameson sensor
// (src/lib/numerics/space/Lib_Shock_Detector.f90:17-44), drag and heat laws
// (src/lib/physics/Lib_Drag.f90, Lib_Heat.f90) and the cell source
// (src/lib/physics/Mod_Sources.f90:230-274, 465-622).
//
// The closure is a template parameter (NC = 6, 7, 12 values live in
// registers).  Operations keep the reference's order: Fortran evaluates + - * /
// of one precedence left to right, and the build passes -ffp-contract=off so no
// product is fused into a sum.  Literals the reference writes as single reals
// in double expressions are promoted from float (eps = 1e-25, toll = 1e-20).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ice {

#define ICE_DEV __device__ __forceinline__

// ------------------------------------------------------------ shared divisor
// Quotients x / d of several numerators over one divisor, each bit for bit the
// IEEE quotient that gfx950's division gives.  That division scales the divisor
// (v_div_scale), refines its reciprocal (v_rcp and two Newton steps), forms and
// corrects the quotient (v_div_fmas) and fixes the special cases (v_div_fixup).
// Only the scaled divisor and its refined reciprocal are shared here; the divisor
// is scaled for the numerator x only at extreme exponents, and such a numerator
// gets the plain division.
struct SharedDiv {
  double d, d0, r;
};
ICE_DEV SharedDiv div_prepare(double d) {
  bool f;
  SharedDiv s;
  s.d = d;
  s.d0 = __builtin_amdgcn_div_scale(1.0, d, false, &f);
  const double r0 = __builtin_amdgcn_rcp(s.d0);
  const double e0 = fma(-s.d0, r0, 1.0);
  const double r1 = fma(r0, e0, r0);
  const double e1 = fma(-s.d0, r1, 1.0);
  s.r = fma(r1, e1, r1);
  return s;
}
ICE_DEV double div_shared(double x, const SharedDiv& s) {
  bool f, vcc;
  const double d0 = __builtin_amdgcn_div_scale(x, s.d, false, &f);
  if (__builtin_expect(__double_as_longlong(d0) != __double_as_longlong(s.d0), 0)) return x / s.d;
  const double n0 = __builtin_amdgcn_div_scale(x, s.d, true, &vcc);
  const double q0 = n0 * s.r;
  const double e = fma(-s.d0, q0, n0);
  return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(e, s.r, q0, vcc), s.d, x);
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kSigmaSB = 5.670374419e-8;
constexpr double kRhoEmpty = 1e-6;
constexpr double kThetaCap = 1.0;
constexpr double kEps = (double)1e-25f;
constexpr double kTollDrag = (double)1e-20f;
constexpr double kTollHeat = 1.e-20;
constexpr double kTauToll = (double)1e-20f;

template <int MODEL> struct Closure;
template <> struct Closure<0> { static constexpr int NC = 6; };
template <> struct Closure<1> { static constexpr int NC = 7; };
template <> struct Closure<2> { static constexpr int NC = 12; };

// Material of a family: constant properties, and optionally a property table
// on the integer kelvins tmin..tmax (Load_Table.f90:131-147), one column of
// tmax - tmin + 1 values per property in the kProp* order, with flags saying
// which tabulated properties the path takes.  rho and cs hold the table's
// values at tmin when a table is set (mat_rho without rho_varies, mat_cp_const).
enum : int { kPropRho = 0, kPropCp, kPropE, kPropPsat, kPropLv, kPropCpv, kPropEmiss, kPropCols };
enum : int { kRhoVaries = 1, kCsVaries = 2, kUsePsat = 4, kUseLv = 8, kUseCpv = 16, kUseEmiss = 32 };
struct MatDev {
  double rho, cs, lv, emiss;
  double ep[8];
  int evap, intf, surrogate, flags;
  const double* tab;  // kPropCols columns; null without a table
  int tmin, tmax;
};
// The material seen by the launches that serve families without tabulated
// properties: its flags are the constant 0, so every table branch of the functions
// below, which take the material's type as a template parameter, folds away.
struct MatConst : MatDev {
  static constexpr int flags = 0;
};

// NaN-ignoring two-argument min / max (Fortran MIN, MAX, MINVAL, MAXVAL)
ICE_DEV double dmin2(double a, double b) { return (b < a || a != a) ? b : a; }
ICE_DEV double dmax2(double a, double b) { return (b > a || a != a) ? b : a; }

// NaN or infinity by the exponent bits (Finite_m.f90:15-23)
ICE_DEV bool nonfinite(double x) {
  uint64_t u = __double_as_longlong(x);
  return ((u >> 52) & 0x7FFull) == 0x7FFull;
}

// NORM2 of three values, the scaled sum of squares
ICE_DEV double norm3(double a, double b, double c) {
  double scale = 0.0, ssq = 1.0;
  double x[3] = {a, b, c};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    if (x[m] != 0.0) {
      double ax = fabs(x[m]);
      if (scale < ax) {
        double t = scale / ax;
        ssq = 1.0 + ssq * t * t;
        scale = ax;
      } else {
        double t = ax / scale;
        ssq = ssq + t * t;
      }
    }
  }
  return scale * sqrt(ssq);
}

// ------------------------------------------------------------- properties
// Lib_Properties.f90:163-250.  A column of the table, index T - tmin.
template <class MT> ICE_DEV const double* prop_col(const MT& m, int c) { return m.tab + (long long)c * (m.tmax - m.tmin + 1); }

// table_value: linear between the nodes, the end value outside, NaN for NaN
ICE_DEV double table_value(const double* t, int lo, int hi, double T) {
  if (T != T) return T;
  double Tc = fmin(fmax(T, (double)lo), (double)hi);
  if (Tc >= (double)hi) return t[hi - lo];
  int i = (int)Tc;
  return t[i - lo] + (t[i + 1 - lo] - t[i - lo]) * (Tc - (double)i);
}
template <class MT> ICE_DEV double prop_value(const MT& m, int c, double T) { return table_value(prop_col(m, c), m.tmin, m.tmax, T); }

// mat_e: the energy column between the nodes, the line through e(0) = 0 below
// tmin (the first segment when tmin = 0), the last segment's slope above tmax
template <class MT> ICE_DEV double mat_e(const MT& m, double T) {
  const double* t = prop_col(m, kPropE);
  const int lo = m.tmin, hi = m.tmax, n = hi - lo;
  if (T != T) return T;
  if (T < (double)lo) return lo > 0 ? t[0] * (T / (double)lo) : (t[1] - t[0]) * T;
  if (T >= (double)hi) return t[n] + (t[n] - t[n - 1]) * (T - (double)hi);
  int i = (int)T;
  return t[i - lo] + (t[i + 1 - lo] - t[i - lo]) * (T - (double)i);
}

// mat_T_from_e: the inverse of mat_e, the bracketing node by bisection
template <class MT> ICE_DEV double mat_T_from_e(const MT& m, double e) {
  const double* t = prop_col(m, kPropE);
  const int lo = m.tmin, hi = m.tmax, n = hi - lo;
  if (e != e) return e;
  if (e < t[0]) return lo > 0 ? (double)lo * (e / t[0]) : e / (t[1] - t[0]);
  if (e >= t[n]) return (double)hi + (e - t[n]) / (t[n] - t[n - 1]);
  int i = lo, j = hi;
  while (j - i > 1) {
    int k = (i + j) / 2;
    if (t[k - lo] <= e) i = k;
    else j = k;
  }
  return (double)i + (e - t[i - lo]) / (t[i + 1 - lo] - t[i - lo]);
}

// The energy of the closures per unit mass, cs T or e(T) (cs_varies)
template <class MT> ICE_DEV double energy_of(const MT& m, double T) { return (m.flags & kCsVaries) ? mat_e(m, T) : m.cs * T; }
// a rho cs T / rho e(T) product in the reference's operand order
template <class MT> ICE_DEV double energy_times(const MT& m, double a, double T) {
  return (m.flags & kCsVaries) ? a * mat_e(m, T) : a * m.cs * T;
}
// T from the internal energy per unit volume: mat_T_from_rhoe (cs_varies) or rhoe / (cs rho + eps)
template <class MT> ICE_DEV double temperature_of(const MT& m, double rhoe, double rho) {
  if (m.flags & kCsVaries) {
    const double* t = prop_col(m, kPropE);
    return mat_T_from_e(m, rhoe / (rho + kEps / (t[1] - t[0])));
  }
  return rhoe / (m.cs * rho + kEps);
}
template <class MT> ICE_DEV double mat_rho(const MT& m, double T) { return (m.flags & kRhoVaries) ? prop_value(m, kPropRho, T) : m.rho; }
template <class MT> ICE_DEV double mat_lv(const MT& m, double T) { return (m.flags & kUseLv) ? prop_value(m, kPropLv, T) : m.lv; }
template <class MT> ICE_DEV double mat_emiss(const MT& m, double T) {
  return (m.flags & kUseEmiss) ? prop_value(m, kPropEmiss, T) : m.emiss;
}

// ------------------------------------------------------------- closures
template <int M, class MT> ICE_DEV void prim2cons(const double* q, double* c, const MT& m) {
  double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  c[0] = q[0];
  c[1] = q[0] * q[1];
  c[2] = q[0] * q[2];
  c[3] = q[0] * q[3];
  if constexpr (M == 0) {
    c[4] = energy_times(m, q[0], q[4]) + 0.5 * q[0] * n2;
    c[5] = q[5];
  } else if constexpr (M == 1) {
    c[4] = q[0] * n2 + 3.0 * q[4];
    c[5] = energy_times(m, q[0], q[5]) + 0.5 * c[4];
    c[6] = q[6];
  } else {
    c[4] = q[0] * q[1] * q[1] + q[4];
    c[5] = q[0] * q[1] * q[2] + q[5];
    c[6] = q[0] * q[1] * q[3] + q[6];
    c[7] = q[0] * q[2] * q[2] + q[7];
    c[8] = q[0] * q[2] * q[3] + q[8];
    c[9] = q[0] * q[3] * q[3] + q[9];
    c[10] = energy_times(m, q[0], q[10]) + 0.5 * (c[4] + c[7] + c[9]);
    c[11] = q[11];
  }
}

template <int M, class MT> ICE_DEV void cons2prim(const double* c, double* q, const MT& m) {
  constexpr int NC = Closure<M>::NC;
  q[0] = c[0];
  q[1] = c[1] / (q[0] + kEps);
  q[2] = c[2] / (q[0] + kEps);
  q[3] = c[3] / (q[0] + kEps);
  if constexpr (M == 0) {
    double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    q[4] = temperature_of(m, c[4] - 0.5 * q[0] * n2, q[0]);
    q[5] = c[5];
    if (q[0] < kEps) {
#pragma unroll
      for (int v = 0; v < NC; ++v) q[v] = kEps;
    }
    if (q[4] < kEps) q[4] = kEps;
    if (q[5] < kEps) q[5] = kEps;
  } else if constexpr (M == 1) {
    double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    q[4] = 1.0 / 3.0 * (c[4] - q[0] * n2);
    if (q[0] <= kRhoEmpty && q[4] > q[0] * kThetaCap) q[4] = q[0] * kThetaCap;
    q[5] = temperature_of(m, c[5] - 0.5 * c[4], q[0]);
    q[6] = c[6];
    if (q[0] < kEps) {
#pragma unroll
      for (int v = 0; v < NC; ++v) q[v] = kEps;
    }
    if (q[4] < kEps) q[4] = kEps;
    if (q[5] < kEps) q[5] = kEps;
    if (q[6] < kEps) q[6] = kEps;
  } else {
    q[4] = c[4] - q[0] * q[1] * q[1];
    q[5] = c[5] - q[0] * q[1] * q[2];
    q[6] = c[6] - q[0] * q[1] * q[3];
    q[7] = c[7] - q[0] * q[2] * q[2];
    q[8] = c[8] - q[0] * q[2] * q[3];
    q[9] = c[9] - q[0] * q[3] * q[3];
    q[10] = temperature_of(m, c[10] - 0.5 * (c[4] + c[7] + c[9]), q[0]);
    q[11] = c[11];
    if (q[0] <= kRhoEmpty && q[4] + q[7] + q[9] > 3.0 * q[0] * kThetaCap) {
      double f = 3.0 * q[0] * kThetaCap / (q[4] + q[7] + q[9]);
#pragma unroll
      for (int v = 4; v < 10; ++v) q[v] = q[v] * f;
    }
    if (q[0] < kEps) {
#pragma unroll
      for (int v = 0; v < NC; ++v) q[v] = kEps;
      q[5] = 0.0;
      q[6] = 0.0;
      q[8] = 0.0;
    }
    if (q[4] < kEps) q[4] = kEps;
    if (fabs(q[5]) < kEps) q[5] = 0.0;
    if (fabs(q[6]) < kEps) q[6] = 0.0;
    if (q[7] < kEps) q[7] = kEps;
    if (fabs(q[8]) < kEps) q[8] = 0.0;
    if (q[9] < kEps) q[9] = kEps;
    if (q[10] < kEps) q[10] = kEps;
    if (q[11] < kEps) q[11] = kEps;
  }
}

template <int M> ICE_DEV bool checkprim(const double* q) {
  constexpr int NC = Closure<M>::NC;
  bool bad = false;
#pragma unroll
  for (int v = 0; v < NC; ++v) bad = bad || nonfinite(q[v]);
  if constexpr (M == 0) return !(q[0] < 0.0 || bad);
  else if constexpr (M == 1) return !(q[0] < 0.0 || q[4] < 0.0 || bad);
  else return !(q[0] < 0.0 || q[4] < 0.0 || q[7] < 0.0 || q[9] < 0.0 || bad);
}

template <int M> ICE_DEV double wavespeed(const double* q, const double* n) {
  if constexpr (M == 0) {
    return 0.0;
  } else if constexpr (M == 1) {
    double pr = (q[4] + q[4] + q[4]) / 3.0;
    return sqrt(3.0 * pr / (q[0] + kEps));
  } else {
    double pnn = n[0] * (q[4] * n[0] + q[5] * n[1] + q[6] * n[2]) +
                 n[1] * (q[5] * n[0] + q[7] * n[1] + q[8] * n[2]) +
                 n[2] * (q[6] * n[0] + q[8] * n[1] + q[9] * n[2]);
    return sqrt(3.0 * dmax2(pnn, 0.0) / (q[0] + kEps));
  }
}

template <int M, class MT> ICE_DEV void physflux(const double* q, const double* n, double* f, const MT& m) {
  double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  double un = q[1] * n[0] + q[2] * n[1] + q[3] * n[2];
  f[0] = q[0] * un;
  if constexpr (M == 0) {
    f[1] = f[0] * q[1];
    f[2] = f[0] * q[2];
    f[3] = f[0] * q[3];
    f[4] = f[0] * (0.5 * n2 + energy_of(m, q[4]));
    f[5] = q[5] * un;
  } else if constexpr (M == 1) {
    f[1] = f[0] * q[1] + q[4] * n[0];
    f[2] = f[0] * q[2] + q[4] * n[1];
    f[3] = f[0] * q[3] + q[4] * n[2];
    f[4] = f[0] * n2 + 5.0 * q[4] * un;
    f[5] = f[0] * (0.5 * n2 + energy_of(m, q[5])) + 2.5 * q[4] * un;
    f[6] = q[6] * un;
  } else {
    double p1n = q[4] * n[0] + q[5] * n[1] + q[6] * n[2];
    double p2n = q[5] * n[0] + q[7] * n[1] + q[8] * n[2];
    double p3n = q[6] * n[0] + q[8] * n[1] + q[9] * n[2];
    f[1] = f[0] * q[1] + p1n;
    f[2] = f[0] * q[2] + p2n;
    f[3] = f[0] * q[3] + p3n;
    f[4] = f[0] * q[1] * q[1] + un * q[4] + q[1] * p1n + q[1] * p1n;
    f[5] = f[0] * q[1] * q[2] + un * q[5] + q[1] * p2n + q[2] * p1n;
    f[6] = f[0] * q[1] * q[3] + un * q[6] + q[1] * p3n + q[3] * p1n;
    f[7] = f[0] * q[2] * q[2] + un * q[7] + q[2] * p2n + q[2] * p2n;
    f[8] = f[0] * q[2] * q[3] + un * q[8] + q[2] * p3n + q[3] * p2n;
    f[9] = f[0] * q[3] * q[3] + un * q[9] + q[3] * p3n + q[3] * p3n;
    f[10] = f[0] * (0.5 * n2 + 0.5 * (q[4] + q[7] + q[9]) / (q[0] + kEps) + energy_of(m, q[10])) + q[1] * p1n +
            q[2] * p2n + q[3] * p3n;
    f[11] = q[11] * un;
  }
}

// source_make of the closure; F = force(1:6) of compute_source_
template <int M, class MT> ICE_DEV void sourcemake(const double* q, const double* F, double* s, const MT& m) {
  double it = F[5] + kEps;
  s[0] = -F[0];
  s[1] = -F[0] * q[1] + q[0] * F[1] / it;
  s[2] = -F[0] * q[2] + q[0] * F[2] / it;
  s[3] = -F[0] * q[3] + q[0] * F[3] / it;
  if constexpr (M == 0) {
    double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    s[4] = -(F[0] * (0.5 * n2 + energy_of(m, q[4]))) - F[0] * mat_lv(m, q[4]) + F[4] + q[0] * (F[1] * q[1] + F[2] * q[2] + F[3] * q[3]) / it;
    s[5] = 0.0;
  } else if constexpr (M == 1) {
    double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    s[4] = -(F[0] * (n2 + 3.0 * q[4] / q[0])) + 2.0 * q[0] * (F[1] * q[1] + F[2] * q[2] + F[3] * q[3]) / it -
           6.0 * q[4] / it;
    s[5] = -energy_times(m, F[0], q[5]) - F[0] * mat_lv(m, q[5]) + F[4] + 0.5 * s[4];
    s[6] = 0.0;
  } else {
    s[4] = -(F[0] * (q[1] * q[1] + q[4] / q[0])) + q[0] * (F[1] * q[1] + F[1] * q[1]) / it - 2.0 * q[4] / it;
    s[5] = -(F[0] * (q[1] * q[2] + q[5] / q[0])) + q[0] * (F[2] * q[1] + F[1] * q[2]) / it - 2.0 * q[5] / it;
    s[6] = -(F[0] * (q[1] * q[3] + q[6] / q[0])) + q[0] * (F[3] * q[1] + F[1] * q[3]) / it - 2.0 * q[6] / it;
    s[7] = -(F[0] * (q[2] * q[2] + q[7] / q[0])) + q[0] * (F[2] * q[2] + F[2] * q[2]) / it - 2.0 * q[7] / it;
    s[8] = -(F[0] * (q[2] * q[3] + q[8] / q[0])) + q[0] * (F[3] * q[2] + F[2] * q[3]) / it - 2.0 * q[8] / it;
    s[9] = -(F[0] * (q[3] * q[3] + q[9] / q[0])) + q[0] * (F[3] * q[3] + F[3] * q[3]) / it - 2.0 * q[9] / it;
    s[10] = -energy_times(m, F[0], q[10]) - F[0] * mat_lv(m, q[10]) + F[4] + 0.5 * s[4] + 0.5 * s[7] + 0.5 * s[9];
    s[11] = 0.0;
  }
}

// H P H with H = I - 2 n n^T on (xx, xy, xz, yy, yz, zz)  (Lib_Model_AG.f90:181-199)
ICE_DEV void mirror6(const double* p, const double* n, double* pm) {
  double t0 = p[0] * n[0] + p[1] * n[1] + p[2] * n[2];
  double t1 = p[1] * n[0] + p[3] * n[1] + p[4] * n[2];
  double t2 = p[2] * n[0] + p[4] * n[1] + p[5] * n[2];
  double s = n[0] * t0 + n[1] * t1 + n[2] * t2;
  pm[0] = p[0] - 4.0 * n[0] * t0 + 4.0 * s * n[0] * n[0];
  pm[1] = p[1] - 2.0 * (n[0] * t1 + t0 * n[1]) + 4.0 * s * n[0] * n[1];
  pm[2] = p[2] - 2.0 * (n[0] * t2 + t0 * n[2]) + 4.0 * s * n[0] * n[2];
  pm[3] = p[3] - 4.0 * n[1] * t1 + 4.0 * s * n[1] * n[1];
  pm[4] = p[4] - 2.0 * (n[1] * t2 + t1 * n[2]) + 4.0 * s * n[1] * n[2];
  pm[5] = p[5] - 4.0 * n[2] * t2 + 4.0 * s * n[2] * n[2];
}

// ------------------------------------------------------------- Riemann
// S: 0 Saurel (MK only), 1 Rusanov, 2 HLLE
template <int M, int S, class MT>
ICE_DEV void riemann(const double* q1, const double* q4, const double* n, double* f, const MT& m) {
  constexpr int NC = Closure<M>::NC;
  double vn1 = q1[1] * n[0] + q1[2] * n[1] + q1[3] * n[2];
  double vn4 = q4[1] * n[0] + q4[2] * n[1] + q4[3] * n[2];
  if constexpr (S == 0) {
    static_assert(M == 0, "Saurel serves the MK closure");
    double vn = 0.5 * (vn1 + vn4);
    const double* qu = vn > 0.0 ? q1 : q4;
    double vu = vn > 0.0 ? vn1 : vn4;
    if (vn > 0.0 || vn < 0.0) {
      double ut = qu[1] - vu * n[0], vt = qu[2] - vu * n[1], wt = qu[3] - vu * n[2];
      double uu = vu * n[0] + ut, vv = vu * n[1] + vt, ww = vu * n[2] + wt;
      f[0] = qu[0] * vn;
      f[1] = uu * f[0];
      f[2] = vv * f[0];
      f[3] = ww * f[0];
      f[4] = f[0] * (0.5 * (uu * uu + vv * vv + ww * ww) + energy_of(m, qu[4]));
      f[5] = qu[5] * vn;
    } else {
#pragma unroll
      for (int v = 0; v < 6; ++v) f[v] = 0.0;
    }
  } else if constexpr (S == 1) {
    double a1 = wavespeed<M>(q1, n), a4 = wavespeed<M>(q4, n);
    double c1[NC], c4[NC], f1[NC], f4[NC];
    prim2cons<M>(q1, c1, m);
    prim2cons<M>(q4, c4, m);
    physflux<M>(q1, n, f1, m);
    physflux<M>(q4, n, f4, m);
    double A = dmax2(dmax2(dmax2(fabs(vn1 - a1), fabs(vn4 - a4)), fabs(vn1 + a1)), fabs(vn4 + a4));
#pragma unroll
    for (int v = 0; v < NC; ++v) f[v] = 0.5 * (f1[v] + f4[v] - A * (c4[v] - c1[v]));
  } else {
    double a1 = wavespeed<M>(q1, n), a4 = wavespeed<M>(q4, n);
    double s1 = sqrt(q1[0]), s4 = sqrt(q4[0]);
    double uR = (s1 * q1[1] + s4 * q4[1]) / (s1 + s4);
    double vR = (s1 * q1[2] + s4 * q4[2]) / (s1 + s4);
    double wR = (s1 * q1[3] + s4 * q4[3]) / (s1 + s4);
    double vnR = uR * n[0] + vR * n[1] + wR * n[2];
    double aR = (s1 * a1 + s4 * a4) / (s1 + s4);
    double S1 = dmin2(dmin2(0.0, vn1 - a1), vnR - aR);
    double S4 = dmax2(dmax2(0.0, vn4 + a4), vnR + aR);
    double k0 = -S1, k1 = S1 * S4, k2 = S4;
    int loc = 0;
    double mn = k0;
    if (k1 < mn) { loc = 1; mn = k1; }
    if (k2 < mn) { loc = 2; }
    if (loc != 1) {
      physflux<M>(loc == 0 ? q1 : q4, n, f, m);
    } else {
      double f1[NC], f4[NC], c1[NC], c4[NC];
      physflux<M>(q1, n, f1, m);
      physflux<M>(q4, n, f4, m);
      prim2cons<M>(q1, c1, m);
      prim2cons<M>(q4, c4, m);
#pragma unroll
      for (int v = 0; v < NC; ++v) f[v] = (S4 * f1[v] - S1 * f4[v] + S1 * S4 * (c4[v] - c1[v])) / (S4 - S1);
    }
  }
}

// ------------------------------------------------------------- limiters
ICE_DEV double limiter(int id, double x, double y, double cx, double cy) {
  if (id == 0) return 0.0;
  double ru = y / (x + 1e-30);
  double phi;
  switch (id) {
    case 1:
      return dmax2(0.0, dmin2(1.0, ru)) * x;
    case 2: phi = (ru * ru + ru) / (ru * ru + 1.0); break;
    case 3: phi = (ru + fabs(ru)) / (1.0 + fabs(ru)); break;
    case 4: phi = (1.5 * (ru * ru + ru)) / (ru * ru + ru + 1.0); break;
    case 5: phi = dmax2(0.0, dmin2(dmin2(dmin2(2.0 * ru, 0.25 + 0.75 * ru), 0.75 + 0.25 * ru), 2.0)); break;
    case 6: phi = dmax2(0.0, dmin2(ru, 1.5)); break;
    case 7: phi = dmax2(dmax2(0.0, dmin2(1.5 * ru, 1.0)), dmin2(ru, 1.5)); break;
    case 8: phi = dmax2(0.0, dmin2(dmin2(2.0 * ru, 0.5 * (1.0 + ru)), 2.0)); break;
    case 9: phi = dmax2(0.0, dmin2(2.0 * ru, dmin2((1.0 + 2.0 * ru) / 3.0, 2.0))); break;
    default: phi = dmax2(dmax2(0.0, dmin2(2.0 * ru, 1.0)), dmin2(ru, 2.0)); break;
  }
  if (ru > 0.0) phi = dmin2(dmin2(phi, cx), cy * ru);
  return phi * x;
}

// MUSCL reconstruction of one face from the four cells of its stencil; false
// where the stencil is unphysical even at first order.
template <int M>
ICE_DEV bool reconstruct(int lim, const double* prev, const double* local, const double* next, const double* next2,
                         double dl0, double dl1, double dl2, double dll, double dlr, double beta, double* ql,
                         double* qr) {
  constexpr int NC = Closure<M>::NC;
  double cxl, cyl, cxr, cyr;
  if (dll > 0.0) { cxl = dl1 / dll; cyl = dl0 / dll; } else { cxl = 2.0; cyl = 2.0; }
  if (dlr > 0.0) { cxr = dl2 / dlr; cyr = dl1 / dlr; } else { cxr = 2.0; cyr = 2.0; }
  bool first_l = prev[0] <= kRhoEmpty || local[0] <= kRhoEmpty || next[0] <= kRhoEmpty;
  bool first_r = local[0] <= kRhoEmpty || next[0] <= kRhoEmpty || next2[0] <= kRhoEmpty;
  if (first_l) {
#pragma unroll
    for (int v = 0; v < NC; ++v) ql[v] = local[v];
  }
  if (first_r) {
#pragma unroll
    for (int v = 0; v < NC; ++v) qr[v] = next[v];
  }
  bool ok_l = first_l, ok_r = first_r;
  double limval = beta;
  while (!ok_l || !ok_r) {
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      double s1 = (next[v] - local[v]) / dl1;
      if (!first_l) {
        double s0 = (local[v] - prev[v]) / dl0;
        ql[v] = local[v] + limiter(lim, s1, s0, cxl, cyl) * limval * dll;
      }
      if (!first_r) {
        double s2 = (next2[v] - next[v]) / dl2;
        qr[v] = next[v] - limiter(lim, s2, s1, cxr, cyr) * limval * dlr;
      }
    }
    if (!first_l) ok_l = checkprim<M>(ql);
    if (!first_r) ok_r = checkprim<M>(qr);
    limval = 0.5 * limval;
    if (limval < 1.e-18 && !(ok_l && ok_r)) {
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        ql[v] = local[v];
        qr[v] = next[v];
      }
      return checkprim<M>(ql) && checkprim<M>(qr);
    }
  }
  return true;
}

// Jameson-type density sensor weight of a cell from its six neighbours
ICE_DEV double sd_weight(double r, double r1, double r2, double r3, double r4, double r5, double r6) {
  double kp1 = fabs((r2 - 2.0 * r + r1) / (r2 + 2.0 * r + r1 + 1e-30));
  double kp2 = fabs((r4 - 2.0 * r + r3) / (r4 + 2.0 * r + r3 + 1e-30));
  double kp3 = fabs((r6 - 2.0 * r + r5) / (r6 + 2.0 * r + r5 + 1e-30));
  double sw = dmax2(dmax2(kp1, kp2), kp3);
  if (sw < 1.0 / 20.0) {
    double phi = dmax2(sw * 20.0, 0.0);
    return 1.0 - tanh(10.0 * phi * phi * phi);
  }
  return 0.0;
}

// ------------------------------------------------------------- drag / heat
ICE_DEV double drag_wen_yu(double Re) {
  if (Re <= 1000) return 24.0 / (Re + kTollDrag) * (1.0 + 0.15 * pow(Re, 0.687));
  return 0.43;
}
ICE_DEV double drag_henderson_sub(double Re, double Ma, double G, double Tr) {
  const double t = kTollDrag;
  double sg = sqrt(0.5 * G);
  return 24.0 * (1.0 / (Re + Ma * sg * (4.33 + (3.65 - 1.53 * Tr) / (1.0 + 0.353 * Tr) * exp(-0.247 * Re / (Ma * sg + t))) + t)) +
         exp(-0.5 * Ma / (sqrt(Re) + t)) *
             ((4.5 + 0.38 * (0.03 * Re + 0.48 * sqrt(Re))) / (1.0 + 0.03 * Re + 0.48 * sqrt(Re)) +
              0.1 * (Ma * Ma) + 0.2 * pow(Ma, 8.0)) +
         0.6 * Ma * sg * (1.0 - exp(-Ma / (Re + t)));
}
ICE_DEV double drag_henderson_sup(double Re, double Ma, double G, double Tr) {
  const double t = kTollDrag;
  return (0.9 + 0.34 / ((Ma * Ma) + t) +
          1.86 * sqrt(Ma / (Re + t)) *
              (2.0 + 2.0 / (Ma * Ma * 0.5 * G + t) + (1.058 / (Ma * sqrt(0.5 * G) + t)) * sqrt(Tr)) -
          1.0 / (pow(Ma, 4.0) * (G * G) * 0.25 + t)) /
         (1.0 + 1.86 * sqrt(Ma / (Re + t)));
}
ICE_DEV double drag_cd(double Re, double Ma, double G, double Tr, int sel) {
  const double t = kTollDrag;
  switch (sel) {
    case 1: return 0.45;
    case 2: return 24.0 / (Re + t);
    case 3: return 24.0 / (Re + t) * (1.0 + 3.0 * Re / 16.0);
    case 4: return 24.0 / (Re + t) * (1.0 + 0.15 * pow(Re, 0.687));
    case 5: return drag_wen_yu(Re);
    case 6: return Re < 1000 ? 24.0 / (Re + t) * (1.0 + pow(Re, 2.0 / 3.0) / 6.0) : 0.4392;
    case 7: return 24.0 / (Re + t) * (1.0 + 0.15 * pow(Re, 0.687) + 0.0175 * Re / (1.0 + 4.25 * 1e4 * pow(Re, -1.16)));
    case 8: {
      double a1, a2, a3;
      if (Re <= 0.1) { a1 = 0; a2 = 24; a3 = 0; }
      else if (Re <= 1.0) { a1 = 3.69; a2 = 22.73; a3 = 0.0903; }
      else if (Re <= 10) { a1 = 1.222; a2 = 29.1667; a3 = -3.8889; }
      else if (Re <= 100) { a1 = 0.6167; a2 = 46.5; a3 = -116.67; }
      else if (Re <= 1000) { a1 = 0.3644; a2 = 98.33; a3 = -2778; }
      else if (Re <= 5000) { a1 = 0.357; a2 = 148.62; a3 = -4.75e+4; }
      else if (Re <= 10000) { a1 = 0.46; a2 = -490.546; a3 = 57.87e+4; }
      else { a1 = 0.5191; a2 = -1662.5; a3 = 5.4167e+6; }
      return a1 + a2 / (Re + t) + a3 / (Re * Re + t);
    }
    case 9: {
      double Cd0 = drag_wen_yu(Re);
      return Cd0 * (1.0 + exp(-0.427 / (pow(Ma, 4.63) + t) - 3.0 / (pow(Re, 0.88) + t))) /
             (1.0 + Ma / (Re + t) * (3.82 + 1.28 * exp(-1.25 * Re / (Ma + t))));
    }
    case 10: {
      if (Ma <= 1) return drag_henderson_sub(Re, Ma, G, Tr);
      if (Ma >= 1.75) return drag_henderson_sup(Re, Ma, G, Tr);
      double Cd1 = drag_henderson_sub(Re, 1.0, G, Tr), Cd2 = drag_henderson_sup(Re, 1.75, G, Tr);
      return Cd1 + 4.0 / 3.0 * (Ma - 1.0) * (Cd2 - Cd1);
    }
    case 11: {
      double g = 1.25 * (1.0 + tanh(0.77 * log10(Re + t) - 1.92));
      g = pow(10.0, g);
      double h = 2.3 + 1.7 * sqrt(Tr) - 2.3 * tanh(1.17 * log10(Ma + t));
      double Cd0 = drag_wen_yu(Re);
      return 2.0 + (Cd0 - 2.0) * exp(-3.07 * sqrt(G) * Ma * g / (Re + t)) +
             h / (Ma * sqrt(G) + t) * exp(-Re / (2 * Ma + t));
    }
    case 12: {
      double g = (1.0 + Re * (12.278 + 0.548 * Re)) / (1.0 + 11.278 * Re);
      double h = 5.6 / (1.0 + Ma) + 1.7 * sqrt(Tr);
      double Cd0 = drag_wen_yu(Re);
      return 2.0 + (Cd0 - 2.0) * exp(-3.07 * sqrt(G) * Ma * g / (Re + t)) +
             h / (Ma * sqrt(G) + t) * exp(-Re / (2 * Ma + t));
    }
    default: return 0.0;
  }
}

ICE_DEV double nusselt(double Re, double Pr, double Ma, int sel) {
  const double t = kTollHeat;
  switch (sel) {
    case 1: return 2.0;
    case 2: return 2.5 * pow(Re, 0.15) + 0.04 * Re;
    case 3: return 2.0 + 0.37 * pow(Re, 0.6) * pow(Pr, 1.0 / 3.0);
    case 4: return 2.0 + 0.459 * pow(Re, 0.55) * pow(Pr, 1.0 / 3.0);
    case 5: return 1.0 / (1.0 / (2.0 + 0.654 * sqrt(Re) * sqrt(Pr)) + 3.42 * Ma / (Re * Pr + t));
    case 6: return 2.0 + 0.6 * sqrt(Re) * pow(Pr, 1.0 / 3.0);
    case 7: {
      double Nu = 2.0 + 0.459 * pow(Re, 0.55) * pow(Pr, 0.33);
      return Nu / (1.0 + 3.42 * Ma / (Re * Pr + t) * Nu);
    }
    default: return 0.0;
  }
}

// RK_stage weights (Lib_RK.f90:9-25), SSP Shu-Osher
ICE_DEV double rk_weight(int nrk, int irk) {
  if (irk == 1) return 1.0;
  if (nrk == 2) return 0.50;
  return irk == 2 ? 0.25 : 2.0 / 3.0;
}

}  // namespace ice

namespace ice {

// ------------------------------------------------------------- streaming forms
// Component v of the conservative vector and of the physical flux, evaluated
// with exactly the operations of prim2cons / physflux, so a Riemann solver can
// emit one flux component at a time without holding the four vectors.
template <int M, class MT> ICE_DEV double cons_v(int v, const double* q, const MT& m) {
  if (v == 0) return q[0];
  if (v <= 3) return q[0] * q[v];
  if constexpr (M == 0) {
    if (v == 4) {
      double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
      return energy_times(m, q[0], q[4]) + 0.5 * q[0] * n2;
    }
    return q[5];
  } else if constexpr (M == 1) {
    double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    double c4 = q[0] * n2 + 3.0 * q[4];
    if (v == 4) return c4;
    if (v == 5) return energy_times(m, q[0], q[5]) + 0.5 * c4;
    return q[6];
  } else {
    switch (v) {
      case 4: return q[0] * q[1] * q[1] + q[4];
      case 5: return q[0] * q[1] * q[2] + q[5];
      case 6: return q[0] * q[1] * q[3] + q[6];
      case 7: return q[0] * q[2] * q[2] + q[7];
      case 8: return q[0] * q[2] * q[3] + q[8];
      case 9: return q[0] * q[3] * q[3] + q[9];
      case 10: {
        double c4 = q[0] * q[1] * q[1] + q[4], c7 = q[0] * q[2] * q[2] + q[7], c9 = q[0] * q[3] * q[3] + q[9];
        return energy_times(m, q[0], q[10]) + 0.5 * (c4 + c7 + c9);
      }
      default: return q[11];
    }
  }
}

// scalars of the physical flux of a state across a face of normal n
struct FluxScal {
  double n2, un, f0, p1n, p2n, p3n;
};
template <int M> ICE_DEV FluxScal flux_scal(const double* q, const double* n) {
  FluxScal s;
  s.n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  s.un = q[1] * n[0] + q[2] * n[1] + q[3] * n[2];
  s.f0 = q[0] * s.un;
  if constexpr (M == 2) {
    s.p1n = q[4] * n[0] + q[5] * n[1] + q[6] * n[2];
    s.p2n = q[5] * n[0] + q[7] * n[1] + q[8] * n[2];
    s.p3n = q[6] * n[0] + q[8] * n[1] + q[9] * n[2];
  } else {
    s.p1n = s.p2n = s.p3n = 0.0;
  }
  return s;
}
template <int M, class MT>
ICE_DEV double flux_v(int v, const double* q, const double* n, const FluxScal& s, const MT& m) {
  if (v == 0) return s.f0;
  if constexpr (M == 0) {
    if (v <= 3) return s.f0 * q[v];
    if (v == 4) return s.f0 * (0.5 * s.n2 + energy_of(m, q[4]));
    return q[5] * s.un;
  } else if constexpr (M == 1) {
    if (v <= 3) return s.f0 * q[v] + q[4] * n[v - 1];
    if (v == 4) return s.f0 * s.n2 + 5.0 * q[4] * s.un;
    if (v == 5) return s.f0 * (0.5 * s.n2 + energy_of(m, q[5])) + 2.5 * q[4] * s.un;
    return q[6] * s.un;
  } else {
    switch (v) {
      case 1: return s.f0 * q[1] + s.p1n;
      case 2: return s.f0 * q[2] + s.p2n;
      case 3: return s.f0 * q[3] + s.p3n;
      case 4: return s.f0 * q[1] * q[1] + s.un * q[4] + q[1] * s.p1n + q[1] * s.p1n;
      case 5: return s.f0 * q[1] * q[2] + s.un * q[5] + q[1] * s.p2n + q[2] * s.p1n;
      case 6: return s.f0 * q[1] * q[3] + s.un * q[6] + q[1] * s.p3n + q[3] * s.p1n;
      case 7: return s.f0 * q[2] * q[2] + s.un * q[7] + q[2] * s.p2n + q[2] * s.p2n;
      case 8: return s.f0 * q[2] * q[3] + s.un * q[8] + q[2] * s.p3n + q[3] * s.p2n;
      case 9: return s.f0 * q[3] * q[3] + s.un * q[9] + q[3] * s.p3n + q[3] * s.p3n;
      case 10:
        return s.f0 * (0.5 * s.n2 + 0.5 * (q[4] + q[7] + q[9]) / (q[0] + kEps) + energy_of(m, q[10])) + q[1] * s.p1n +
               q[2] * s.p2n + q[3] * s.p3n;
      default: return q[11] * s.un;
    }
  }
}

// The Riemann flux of states ql | qr times the face area, each component added
// to (MODE 0) or subtracted from (MODE 1) the accumulator r, or stored in r (MODE 2).
template <int M, int S, int MODE, class MT>
ICE_DEV void riemann_acc(const double* q1, const double* q4, const double* n, double area, double* r, const MT& m) {
  constexpr int NC = Closure<M>::NC;
  if constexpr (S == 0) {
    double f[NC];
    riemann<M, 0>(q1, q4, n, f, m);
#pragma unroll
    for (int v = 0; v < NC; ++v) r[v] = MODE == 2 ? f[v] * area : (MODE == 1 ? r[v] - f[v] * area : r[v] + f[v] * area);
  } else {
    double vn1 = q1[1] * n[0] + q1[2] * n[1] + q1[3] * n[2];
    double vn4 = q4[1] * n[0] + q4[2] * n[1] + q4[3] * n[2];
    double a1 = wavespeed<M>(q1, n), a4 = wavespeed<M>(q4, n);
    FluxScal s1 = flux_scal<M>(q1, n), s4 = flux_scal<M>(q4, n);
    if constexpr (S == 1) {
      double A = dmax2(dmax2(dmax2(fabs(vn1 - a1), fabs(vn4 - a4)), fabs(vn1 + a1)), fabs(vn4 + a4));
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        double fv = 0.5 * (flux_v<M>(v, q1, n, s1, m) + flux_v<M>(v, q4, n, s4, m) -
                           A * (cons_v<M>(v, q4, m) - cons_v<M>(v, q1, m)));
        r[v] = MODE == 2 ? fv * area : (MODE == 1 ? r[v] - fv * area : r[v] + fv * area);
      }
    } else {
      double sq1 = sqrt(q1[0]), sq4 = sqrt(q4[0]);
      const SharedDiv bys = div_prepare(sq1 + sq4);
      double uR = div_shared(sq1 * q1[1] + sq4 * q4[1], bys);
      double vR = div_shared(sq1 * q1[2] + sq4 * q4[2], bys);
      double wR = div_shared(sq1 * q1[3] + sq4 * q4[3], bys);
      double vnR = uR * n[0] + vR * n[1] + wR * n[2];
      double aR = div_shared(sq1 * a1 + sq4 * a4, bys);
      double S1 = dmin2(dmin2(0.0, vn1 - a1), vnR - aR);
      double S4 = dmax2(dmax2(0.0, vn4 + a4), vnR + aR);
      double k0 = -S1, k1 = S1 * S4, k2 = S4;
      int loc = 0;
      double mn = k0;
      if (k1 < mn) { loc = 1; mn = k1; }
      if (k2 < mn) loc = 2;
      if (loc != 1) {
        const double* qu = loc == 0 ? q1 : q4;
        FluxScal su = loc == 0 ? s1 : s4;
#pragma unroll
        for (int v = 0; v < NC; ++v) {
          double fv = flux_v<M>(v, qu, n, su, m);
          r[v] = MODE == 2 ? fv * area : (MODE == 1 ? r[v] - fv * area : r[v] + fv * area);
        }
      } else {
        const SharedDiv byw = div_prepare(S4 - S1);
#pragma unroll
        for (int v = 0; v < NC; ++v) {
          double fv = div_shared(S4 * flux_v<M>(v, q1, n, s1, m) - S1 * flux_v<M>(v, q4, n, s4, m) +
                                     S1 * S4 * (cons_v<M>(v, q4, m) - cons_v<M>(v, q1, m)),
                                 byw);
          r[v] = MODE == 2 ? fv * area : (MODE == 1 ? r[v] - fv * area : r[v] + fv * area);
        }
      }
    }
  }
}

// MUSCL reconstruction reading the stencil slot by slot from memory: p points
// at slot 0 of the cell before the face's lower cell, consecutive stencil
// cells are `stride` apart, slots N apart.
template <int M, int LIM = -1>
ICE_DEV bool reconstruct_mem(int lim, const double* __restrict__ p, long long stride, long long N, double dl0,
                             double dl1, double dl2, double dll, double dlr, double beta, double* ql, double* qr) {
  if constexpr (LIM >= 0) lim = LIM;
  constexpr int NC = Closure<M>::NC;
  double cxl, cyl, cxr, cyr;
  if (dll > 0.0) { cxl = dl1 / dll; cyl = dl0 / dll; } else { cxl = 2.0; cyl = 2.0; }
  if (dlr > 0.0) { cxr = dl2 / dlr; cyr = dl1 / dlr; } else { cxr = 2.0; cyr = 2.0; }
  double r0 = p[0], r1 = p[stride], r2 = p[2 * stride], r3 = p[3 * stride];
  bool first_l = r0 <= kRhoEmpty || r1 <= kRhoEmpty || r2 <= kRhoEmpty;
  bool first_r = r1 <= kRhoEmpty || r2 <= kRhoEmpty || r3 <= kRhoEmpty;
  if (first_l) {
#pragma unroll
    for (int v = 0; v < NC; ++v) ql[v] = p[v * N + stride];
  }
  if (first_r) {
#pragma unroll
    for (int v = 0; v < NC; ++v) qr[v] = p[v * N + 2 * stride];
  }
  bool ok_l = first_l, ok_r = first_r;
  double limval = beta;
  // the slopes of every slot divide by the same three spacings
  const SharedDiv by0 = div_prepare(dl0), by1 = div_prepare(dl1), by2 = div_prepare(dl2);
  while (!ok_l || !ok_r) {
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      const double* pv = p + v * N;
      double prev = pv[0], local = pv[stride], next = pv[2 * stride], next2 = pv[3 * stride];
      double s1 = div_shared(next - local, by1);
      if (!first_l) {
        double s0 = div_shared(local - prev, by0);
        ql[v] = local + limiter(lim, s1, s0, cxl, cyl) * limval * dll;
      }
      if (!first_r) {
        double s2 = div_shared(next2 - next, by2);
        qr[v] = next - limiter(lim, s2, s1, cxr, cyr) * limval * dlr;
      }
    }
    if (!first_l) ok_l = checkprim<M>(ql);
    if (!first_r) ok_r = checkprim<M>(qr);
    limval = 0.5 * limval;
    if (limval < 1.e-18 && !(ok_l && ok_r)) {
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        ql[v] = p[v * N + stride];
        qr[v] = p[v * N + 2 * stride];
      }
      return checkprim<M>(ql) && checkprim<M>(qr);
    }
  }
  return true;
}

}  // namespace ice
