// ORACLE -- test infrastructure only.  A CPU restatement of ICE's per-cell and
// per-face arithmetic, written op for op after the reference Fortran (every
// function cites the file:line it follows) so that the HIP path can be checked
// against it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg may load it; the product never does.
//
// Evaluation order follows Fortran's left-to-right rule for + - * / of equal
// precedence.  Literals that the reference writes as default (single) reals in
// double expressions are promoted from float: eps = 1e-25 (closures), toll =
// 1e-20 (Lib_Drag), + 1e-20 (Mod_Sources).  Fortran MIN/MAX/MINVAL/MAXVAL are
// taken as NaN-ignoring (gfortran's MINVAL/MAXVAL; fmin/fmax).
//
// Pinning (see tests/test_oracle_ref.py): every function here that has a
// counterpart in the reference's leaf modules is compared against the
// flang-built reference (oracle/_ref) on random inputs; the rest are checked
// against closed forms and invariants in tests/test_oracle_*.py.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

static const double PI = 3.14159265358979323846;          // Parameters_m.f90:13
static const double SIGMA_SB = 5.670374419e-8;            // Parameters_m.f90:14
static const double RHO_EMPTY = 1e-6;                     // Global_m.f90:10
static const double THETA_CAP = 1.0;                      // Global_m.f90:11
static const double EPS25 = (double)1e-25f;               // closures' eps, a default-real literal
static const double TOLL_DRAG = (double)1e-20f;           // Lib_Drag.f90:125..302
static const double TOLL_HEAT = 1.e-20;                   // Lib_Heat.f90:13
static const double TAU_TOLL = (double)1e-20f;            // Mod_Sources.f90:611

enum { MK = 0, IG = 1, AG = 2 };

static inline int ncond_of(int model) { return model == MK ? 6 : (model == IG ? 7 : 12); }

struct Mat {
  double rho, cs, lv, emiss;
  double ep[8];
  int evap, intf;
  // property table (Config_Types_m.f90:273-291): ICE_PROP_NCOL columns over T = tmin..tmax
  const double* tab = nullptr;
  int tmin = 0, tmax = 0, flags = 0;
};

// NaN-ignoring min / max of two (Fortran MIN/MAX as gfortran compiles them)
static inline double fmin2(double a, double b) { return (b < a || a != a) ? b : a; }
static inline double fmax2(double a, double b) { return (b > a || a != a) ? b : a; }

// ---------------------------------------------------------------- Lib_Properties.f90
// tab(T) of one column, T an integer kelvin in tmin..tmax
static inline double tab_at(const Mat& m, int col, int T) {
  return m.tab[(size_t)col * (size_t)(m.tmax - m.tmin + 1) + (size_t)(T - m.tmin)];
}
// table_value :233-250
static inline double table_value(const Mat& m, int col, double T) {
  if (T != T) return T;
  double Tc = std::min(std::max(T, (double)m.tmin), (double)m.tmax);
  if (Tc >= (double)m.tmax) return tab_at(m, col, m.tmax);
  int i = (int)Tc;
  return tab_at(m, col, i) + (tab_at(m, col, i + 1) - tab_at(m, col, i)) * (Tc - (double)i);
}
// mat_rho :18-29 (rho holds rho_tab(T_min) when a table is set)
static inline double mat_rho(const Mat& m, double T) {
  return (m.flags & ICE_PROP_RHO_VARIES) ? table_value(m, ICE_PROP_RHO, T) : m.rho;
}
// mat_lv :97-106, mat_emiss :123-132
static inline double mat_lv(const Mat& m, double T) {
  return (m.flags & ICE_PROP_USE_LV) ? table_value(m, ICE_PROP_LV, T) : m.lv;
}
static inline double mat_emiss(const Mat& m, double T) {
  return (m.flags & ICE_PROP_USE_EMISS) ? table_value(m, ICE_PROP_EMISS, T) : m.emiss;
}
// mat_e :163-184
static inline double mat_e(const Mat& m, double T) {
  const int lo = m.tmin, hi = m.tmax;
  if (T != T) return T;
  if (T < (double)lo) {
    if (lo > 0) return tab_at(m, ICE_PROP_E, lo) * (T / (double)lo);
    return (tab_at(m, ICE_PROP_E, lo + 1) - tab_at(m, ICE_PROP_E, lo)) * T;
  }
  if (T >= (double)hi)
    return tab_at(m, ICE_PROP_E, hi) + (tab_at(m, ICE_PROP_E, hi) - tab_at(m, ICE_PROP_E, hi - 1)) * (T - (double)hi);
  int i = (int)T;
  return tab_at(m, ICE_PROP_E, i) + (tab_at(m, ICE_PROP_E, i + 1) - tab_at(m, ICE_PROP_E, i)) * (T - (double)i);
}
// mat_T_from_e :188-218
static inline double mat_T_from_e(const Mat& m, double e) {
  const int lo = m.tmin, hi = m.tmax;
  if (e != e) return e;
  if (e < tab_at(m, ICE_PROP_E, lo)) {
    if (lo > 0) return (double)lo * (e / tab_at(m, ICE_PROP_E, lo));
    return e / (tab_at(m, ICE_PROP_E, lo + 1) - tab_at(m, ICE_PROP_E, lo));
  }
  if (e >= tab_at(m, ICE_PROP_E, hi))
    return (double)hi + (e - tab_at(m, ICE_PROP_E, hi)) / (tab_at(m, ICE_PROP_E, hi) - tab_at(m, ICE_PROP_E, hi - 1));
  int i = lo, j = hi;
  while (j - i > 1) {
    int k = (i + j) / 2;
    if (tab_at(m, ICE_PROP_E, k) <= e) i = k;
    else j = k;
  }
  return (double)i + (e - tab_at(m, ICE_PROP_E, i)) / (tab_at(m, ICE_PROP_E, i + 1) - tab_at(m, ICE_PROP_E, i));
}
// mat_T_from_rhoe :224-229
static inline double mat_T_from_rhoe(const Mat& m, double rhoe, double rho, double eps) {
  return mat_T_from_e(m, rhoe / (rho + eps / (tab_at(m, ICE_PROP_E, m.tmin + 1) - tab_at(m, ICE_PROP_E, m.tmin))));
}
static inline bool cs_varies(const Mat& m) { return (m.flags & ICE_PROP_CS_VARIES) != 0; }

// Finite_m.f90:15-23: NaN or infinity, by the exponent bits
static inline bool not_finite(double x) {
  uint64_t u; __builtin_memcpy(&u, &x, 8);
  return ((u >> 52) & 0x7FFu) == 0x7FFu;
}

// gfortran's NORM2 of a 3-vector: the scaled sum of squares
static inline double norm2_3(double a, double b, double c) {
  double scale = 0.0, ssq = 1.0;
  const double x[3] = {a, b, c};
  for (int m = 0; m < 3; ++m) {
    if (x[m] != 0.0) {
      double ax = std::fabs(x[m]);
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
  return scale * std::sqrt(ssq);
}

// ------------------------------------------------------------------ closures
// Lib_Model_MK.f90:11-33, Lib_Model_IG.f90:11-40, Lib_Model_AG.f90:11-40
static inline void p2c(int model, const double* q, double* c, const Mat& m) {
  double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  if (model == MK) {
    c[0] = q[0]; c[1] = q[0] * q[1]; c[2] = q[0] * q[2]; c[3] = q[0] * q[3];
    c[4] = (cs_varies(m) ? q[0] * mat_e(m, q[4]) : q[0] * m.cs * q[4]) + 0.5 * q[0] * n2;
    c[5] = q[5];
  } else if (model == IG) {
    c[0] = q[0]; c[1] = q[0] * q[1]; c[2] = q[0] * q[2]; c[3] = q[0] * q[3];
    c[4] = q[0] * n2 + 3.0 * q[4];
    c[5] = (cs_varies(m) ? q[0] * mat_e(m, q[5]) : q[0] * m.cs * q[5]) + 0.5 * c[4];
    c[6] = q[6];
  } else {
    c[0] = q[0]; c[1] = q[0] * q[1]; c[2] = q[0] * q[2]; c[3] = q[0] * q[3];
    c[4] = q[0] * q[1] * q[1] + q[4];
    c[5] = q[0] * q[1] * q[2] + q[5];
    c[6] = q[0] * q[1] * q[3] + q[6];
    c[7] = q[0] * q[2] * q[2] + q[7];
    c[8] = q[0] * q[2] * q[3] + q[8];
    c[9] = q[0] * q[3] * q[3] + q[9];
    c[10] = (cs_varies(m) ? q[0] * mat_e(m, q[10]) : q[0] * m.cs * q[10]) + 0.5 * (c[4] + c[7] + c[9]);
    c[11] = q[11];
  }
}

// Lib_Model_MK.f90:36-77, Lib_Model_IG.f90:43-89, Lib_Model_AG.f90:43-126
static inline void c2p(int model, const double* c, double* q, const Mat& m) {
  const double eps = EPS25;
  if (model == MK) {
    q[0] = c[0];
    q[1] = c[1] / (q[0] + eps); q[2] = c[2] / (q[0] + eps); q[3] = c[3] / (q[0] + eps);
    double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    q[4] = cs_varies(m) ? mat_T_from_rhoe(m, c[4] - 0.5 * q[0] * n2, q[0], eps)
                         : (c[4] - 0.5 * q[0] * n2) / (m.cs * q[0] + eps);
    q[5] = c[5];
    if (q[0] < eps) for (int v = 0; v < 6; ++v) q[v] = eps;
    if (q[4] < eps) q[4] = eps;
    if (q[5] < eps) q[5] = eps;
  } else if (model == IG) {
    q[0] = c[0];
    q[1] = c[1] / (q[0] + eps); q[2] = c[2] / (q[0] + eps); q[3] = c[3] / (q[0] + eps);
    double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    q[4] = 1.0 / 3.0 * (c[4] - q[0] * n2);
    if (q[0] <= RHO_EMPTY && q[4] > q[0] * THETA_CAP) q[4] = q[0] * THETA_CAP;
    q[5] = cs_varies(m) ? mat_T_from_rhoe(m, c[5] - 0.5 * c[4], q[0], eps) : (c[5] - 0.5 * c[4]) / (m.cs * q[0] + eps);
    q[6] = c[6];
    if (q[0] < eps) for (int v = 0; v < 7; ++v) q[v] = eps;
    if (q[4] < eps) q[4] = eps;
    if (q[5] < eps) q[5] = eps;
    if (q[6] < eps) q[6] = eps;
  } else {
    q[0] = c[0];
    q[1] = c[1] / (q[0] + eps); q[2] = c[2] / (q[0] + eps); q[3] = c[3] / (q[0] + eps);
    q[4] = c[4] - q[0] * q[1] * q[1];
    q[5] = c[5] - q[0] * q[1] * q[2];
    q[6] = c[6] - q[0] * q[1] * q[3];
    q[7] = c[7] - q[0] * q[2] * q[2];
    q[8] = c[8] - q[0] * q[2] * q[3];
    q[9] = c[9] - q[0] * q[3] * q[3];
    q[10] = cs_varies(m) ? mat_T_from_rhoe(m, c[10] - 0.5 * (c[4] + c[7] + c[9]), q[0], eps)
                          : (c[10] - 0.5 * (c[4] + c[7] + c[9])) / (m.cs * q[0] + eps);
    q[11] = c[11];
    if (q[0] <= RHO_EMPTY && q[4] + q[7] + q[9] > 3.0 * q[0] * THETA_CAP) {
      double f = 3.0 * q[0] * THETA_CAP / (q[4] + q[7] + q[9]);
      for (int v = 4; v < 10; ++v) q[v] = q[v] * f;
    }
    if (q[0] < eps) {
      for (int v = 0; v < 12; ++v) q[v] = eps;
      q[5] = 0.0; q[6] = 0.0; q[8] = 0.0;
    }
    if (q[4] < eps) q[4] = eps;
    if (std::fabs(q[5]) < eps) q[5] = 0.0;
    if (std::fabs(q[6]) < eps) q[6] = 0.0;
    if (q[7] < eps) q[7] = eps;
    if (std::fabs(q[8]) < eps) q[8] = 0.0;
    if (q[9] < eps) q[9] = eps;
    if (q[10] < eps) q[10] = eps;
    if (q[11] < eps) q[11] = eps;
  }
}

// check_prim: Lib_Model_MK.f90:80-93, Lib_Model_IG.f90:92-105, Lib_Model_AG.f90:129-142
static inline bool check_prim(int model, const double* q) {
  int nc = ncond_of(model);
  bool bad = false;
  for (int v = 0; v < nc; ++v) bad = bad || not_finite(q[v]);
  if (model == MK) return !(q[0] < 0.0 || bad);
  if (model == IG) return !(q[0] < 0.0 || q[4] < 0.0 || bad);
  return !(q[0] < 0.0 || q[4] < 0.0 || q[7] < 0.0 || q[9] < 0.0 || bad);
}

// wavespeed_make: MK :128-135 (0), IG :128-137, AG :165-177
static inline double wavespeed(int model, const double* q, const double* nrm) {
  const double eps = EPS25;
  if (model == MK) return 0.0;
  if (model == IG) {
    double pr = (q[4] + q[4] + q[4]) / 3.0;
    return std::sqrt(3.0 * pr / (q[0] + eps));
  }
  double pnn = nrm[0] * (q[4] * nrm[0] + q[5] * nrm[1] + q[6] * nrm[2]) +
               nrm[1] * (q[5] * nrm[0] + q[7] * nrm[1] + q[8] * nrm[2]) +
               nrm[2] * (q[6] * nrm[0] + q[8] * nrm[1] + q[9] * nrm[2]);
  return std::sqrt(3.0 * fmax2(pnn, 0.0) / (q[0] + eps));
}

// flux_make: MK :138-162, IG :140-169, AG :202-242
static inline void flux_make(int model, const double* q, const double* nrm, double* f, const Mat& m) {
  double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  double un = q[1] * nrm[0] + q[2] * nrm[1] + q[3] * nrm[2];
  if (model == MK) {
    f[0] = q[0] * un;
    f[1] = f[0] * q[1]; f[2] = f[0] * q[2]; f[3] = f[0] * q[3];
    f[4] = f[0] * (0.5 * n2 + (cs_varies(m) ? mat_e(m, q[4]) : m.cs * q[4]));
    f[5] = q[5] * un;
  } else if (model == IG) {
    f[0] = q[0] * un;
    f[1] = f[0] * q[1] + q[4] * nrm[0];
    f[2] = f[0] * q[2] + q[4] * nrm[1];
    f[3] = f[0] * q[3] + q[4] * nrm[2];
    f[4] = f[0] * n2 + 5.0 * q[4] * un;
    f[5] = f[0] * (0.5 * n2 + (cs_varies(m) ? mat_e(m, q[5]) : m.cs * q[5])) + 2.5 * q[4] * un;
    f[6] = q[6] * un;
  } else {
    const double eps = EPS25;
    double p1n = q[4] * nrm[0] + q[5] * nrm[1] + q[6] * nrm[2];
    double p2n = q[5] * nrm[0] + q[7] * nrm[1] + q[8] * nrm[2];
    double p3n = q[6] * nrm[0] + q[8] * nrm[1] + q[9] * nrm[2];
    f[0] = q[0] * un;
    f[1] = f[0] * q[1] + p1n;
    f[2] = f[0] * q[2] + p2n;
    f[3] = f[0] * q[3] + p3n;
    f[4] = f[0] * q[1] * q[1] + un * q[4] + q[1] * p1n + q[1] * p1n;
    f[5] = f[0] * q[1] * q[2] + un * q[5] + q[1] * p2n + q[2] * p1n;
    f[6] = f[0] * q[1] * q[3] + un * q[6] + q[1] * p3n + q[3] * p1n;
    f[7] = f[0] * q[2] * q[2] + un * q[7] + q[2] * p2n + q[2] * p2n;
    f[8] = f[0] * q[2] * q[3] + un * q[8] + q[2] * p3n + q[3] * p2n;
    f[9] = f[0] * q[3] * q[3] + un * q[9] + q[3] * p3n + q[3] * p3n;
    f[10] = f[0] * (0.5 * n2 + 0.5 * (q[4] + q[7] + q[9]) / (q[0] + eps) + (cs_varies(m) ? mat_e(m, q[10]) : m.cs * q[10])) +
            q[1] * p1n + q[2] * p2n + q[3] * p3n;
    f[11] = q[11] * un;
  }
}

// source_make: MK :165-185, IG :172-203, AG :245-297 (constant properties: lv = lv_al)
static inline void source_make(int model, const double* q, const double* F, double* s, const Mat& m) {
  const double eps = EPS25;
  if (model == MK) {
    double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    s[0] = -F[0];
    s[1] = -F[0] * q[1] + q[0] * F[1] / (F[5] + eps);
    s[2] = -F[0] * q[2] + q[0] * F[2] / (F[5] + eps);
    s[3] = -F[0] * q[3] + q[0] * F[3] / (F[5] + eps);
    s[4] = -(F[0] * (0.5 * n2 + (cs_varies(m) ? mat_e(m, q[4]) : m.cs * q[4]))) - F[0] * mat_lv(m, q[4]) + F[4] +
           q[0] * (F[1] * q[1] + F[2] * q[2] + F[3] * q[3]) / (F[5] + eps);
    s[5] = 0.0;
  } else if (model == IG) {
    double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    s[0] = -F[0];
    s[1] = -F[0] * q[1] + q[0] * F[1] / (F[5] + eps);
    s[2] = -F[0] * q[2] + q[0] * F[2] / (F[5] + eps);
    s[3] = -F[0] * q[3] + q[0] * F[3] / (F[5] + eps);
    s[4] = -(F[0] * (n2 + 3.0 * q[4] / q[0])) +
           2.0 * q[0] * (F[1] * q[1] + F[2] * q[2] + F[3] * q[3]) / (F[5] + eps) -
           6.0 * q[4] / (F[5] + eps);
    s[5] = -(cs_varies(m) ? F[0] * mat_e(m, q[5]) : F[0] * m.cs * q[5]) - F[0] * mat_lv(m, q[5]) + F[4] + 0.5 * s[4];
    s[6] = 0.0;
  } else {
    s[0] = -F[0];
    s[1] = -F[0] * q[1] + q[0] * F[1] / (F[5] + eps);
    s[2] = -F[0] * q[2] + q[0] * F[2] / (F[5] + eps);
    s[3] = -F[0] * q[3] + q[0] * F[3] / (F[5] + eps);
    s[4] = -(F[0] * (q[1] * q[1] + q[4] / q[0])) + q[0] * (F[1] * q[1] + F[1] * q[1]) / (F[5] + eps) -
           2.0 * q[4] / (F[5] + eps);
    s[5] = -(F[0] * (q[1] * q[2] + q[5] / q[0])) + q[0] * (F[2] * q[1] + F[1] * q[2]) / (F[5] + eps) -
           2.0 * q[5] / (F[5] + eps);
    s[6] = -(F[0] * (q[1] * q[3] + q[6] / q[0])) + q[0] * (F[3] * q[1] + F[1] * q[3]) / (F[5] + eps) -
           2.0 * q[6] / (F[5] + eps);
    s[7] = -(F[0] * (q[2] * q[2] + q[7] / q[0])) + q[0] * (F[2] * q[2] + F[2] * q[2]) / (F[5] + eps) -
           2.0 * q[7] / (F[5] + eps);
    s[8] = -(F[0] * (q[2] * q[3] + q[8] / q[0])) + q[0] * (F[3] * q[2] + F[2] * q[3]) / (F[5] + eps) -
           2.0 * q[8] / (F[5] + eps);
    s[9] = -(F[0] * (q[3] * q[3] + q[9] / q[0])) + q[0] * (F[3] * q[3] + F[3] * q[3]) / (F[5] + eps) -
           2.0 * q[9] / (F[5] + eps);
    s[10] = -(cs_varies(m) ? F[0] * mat_e(m, q[10]) : F[0] * m.cs * q[10]) - F[0] * mat_lv(m, q[10]) + F[4] + 0.5 * s[4] +
             0.5 * s[7] + 0.5 * s[9];
    s[11] = 0.0;
  }
}

// AG :181-199  H P H, H = I - 2 n n^T, on (xx,xy,xz,yy,yz,zz)
static inline void mirror_tensor(const double* p, const double* n, double* pm) {
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

// ------------------------------------------------------------------ Riemann
enum { SAUREL = 0, RUSANOV = 1, HLLE = 2 };

// Lib_Riemann.f90:61-109
static inline void riemann_saurel(const double* q1, const double* q4, const double* n, double* f, const Mat& m) {
  double vn1 = q1[1] * n[0] + q1[2] * n[1] + q1[3] * n[2];
  double vn4 = q4[1] * n[0] + q4[2] * n[1] + q4[3] * n[2];
  double vn = 0.5 * (vn1 + vn4);
  if (vn > 0.0) {
    double ut = q1[1] - vn1 * n[0], vt = q1[2] - vn1 * n[1], wt = q1[3] - vn1 * n[2];
    double uu = vn1 * n[0] + ut, vv = vn1 * n[1] + vt, ww = vn1 * n[2] + wt;
    f[0] = q1[0] * vn;
    f[1] = uu * f[0]; f[2] = vv * f[0]; f[3] = ww * f[0];
    f[4] = f[0] * (0.5 * (uu * uu + vv * vv + ww * ww) + (cs_varies(m) ? mat_e(m, q1[4]) : m.cs * q1[4]));
    f[5] = q1[5] * vn;
  } else if (vn < 0.0) {
    double ut = q4[1] - vn4 * n[0], vt = q4[2] - vn4 * n[1], wt = q4[3] - vn4 * n[2];
    double uu = vn4 * n[0] + ut, vv = vn4 * n[1] + vt, ww = vn4 * n[2] + wt;
    f[0] = q4[0] * vn;
    f[1] = uu * f[0]; f[2] = vv * f[0]; f[3] = ww * f[0];
    f[4] = f[0] * (0.5 * (uu * uu + vv * vv + ww * ww) + (cs_varies(m) ? mat_e(m, q4[4]) : m.cs * q4[4]));
    f[5] = q4[5] * vn;
  } else {
    for (int v = 0; v < 6; ++v) f[v] = 0.0;
  }
}

// Lib_Riemann.f90:153-186
static inline void riemann_rusanov(int model, const double* q1, const double* q4, const double* n, double* f,
                                   const Mat& m) {
  int nc = ncond_of(model);
  double vn1 = q1[1] * n[0] + q1[2] * n[1] + q1[3] * n[2];
  double vn4 = q4[1] * n[0] + q4[2] * n[1] + q4[3] * n[2];
  double a1 = wavespeed(model, q1, n), a4 = wavespeed(model, q4, n);
  double c1[12], c4[12], f1[12], f4[12];
  p2c(model, q1, c1, m); p2c(model, q4, c4, m);
  flux_make(model, q1, n, f1, m); flux_make(model, q4, n, f4, m);
  double A = fmax2(fmax2(fmax2(std::fabs(vn1 - a1), std::fabs(vn4 - a4)), std::fabs(vn1 + a1)), std::fabs(vn4 + a4));
  for (int v = 0; v < nc; ++v) f[v] = 0.5 * (f1[v] + f4[v] - A * (c4[v] - c1[v]));
}

// Lib_Riemann.f90:190-241
static inline void riemann_hlle(int model, const double* q1, const double* q4, const double* n, double* f,
                                const Mat& m) {
  int nc = ncond_of(model);
  double vn1 = q1[1] * n[0] + q1[2] * n[1] + q1[3] * n[2];
  double vn4 = q4[1] * n[0] + q4[2] * n[1] + q4[3] * n[2];
  double a1 = wavespeed(model, q1, n), a4 = wavespeed(model, q4, n);
  double s1 = std::sqrt(q1[0]), s4 = std::sqrt(q4[0]);
  double uR = (s1 * q1[1] + s4 * q4[1]) / (s1 + s4);
  double vR = (s1 * q1[2] + s4 * q4[2]) / (s1 + s4);
  double wR = (s1 * q1[3] + s4 * q4[3]) / (s1 + s4);
  double vnR = uR * n[0] + vR * n[1] + wR * n[2];
  double aR = (s1 * a1 + s4 * a4) / (s1 + s4);
  double S1 = fmin2(fmin2(0.0, vn1 - a1), vnR - aR);
  double S4 = fmax2(fmax2(0.0, vn4 + a4), vnR + aR);
  // minloc([-S1, S1*S4, S4]): the first minimum
  double k0 = -S1, k1 = S1 * S4, k2 = S4;
  int loc = 0; double mn = k0;
  if (k1 < mn) { loc = 1; mn = k1; }
  if (k2 < mn) { loc = 2; mn = k2; }
  if (loc == 0) {
    flux_make(model, q1, n, f, m);
  } else if (loc == 2) {
    flux_make(model, q4, n, f, m);
  } else {
    double f1[12], f4[12], c1[12], c4[12];
    flux_make(model, q1, n, f1, m); flux_make(model, q4, n, f4, m);
    p2c(model, q1, c1, m); p2c(model, q4, c4, m);
    for (int v = 0; v < nc; ++v) f[v] = (S4 * f1[v] - S1 * f4[v] + S1 * S4 * (c4[v] - c1[v])) / (S4 - S1);
  }
}

static inline void riemann(int solver, int model, const double* q1, const double* q4, const double* n, double* f,
                           const Mat& m) {
  if (solver == SAUREL) riemann_saurel(q1, q4, n, f, m);
  else if (solver == RUSANOV) riemann_rusanov(model, q1, q4, n, f, m);
  else riemann_hlle(model, q1, q4, n, f, m);
}

// ------------------------------------------------------------------ limiters
// Lib_Limiters.f90:116-295; the bounds min(phi, cx, cy*r) for r > 0
static inline double limiter(int id, double x, double y, double cx, double cy) {
  if (id == 0) return 0.0;  // IORD :116-126
  double ru = y / (x + 1e-30);
  double phi;
  switch (id) {
    case 1:  // minmod :130-143
      phi = fmax2(0.0, fmin2(1.0, ru));
      return phi * x;
    case 2:  // van Albada :146-160
      phi = (ru * ru + ru) / (ru * ru + 1.0);
      break;
    case 3:  // van Leer :163-177
      phi = (ru + std::fabs(ru)) / (1.0 + std::fabs(ru));
      break;
    case 4:  // OSPRE :180-194
      phi = (1.5 * (ru * ru + ru)) / (ru * ru + ru + 1.0);
      break;
    case 5:  // UMIST :197-211
      phi = fmax2(0.0, fmin2(fmin2(fmin2(2.0 * ru, 0.25 + 0.75 * ru), 0.75 + 0.25 * ru), 2.0));
      break;
    case 6:  // Osher :214-228
      phi = fmax2(0.0, fmin2(ru, 1.5));
      break;
    case 7:  // Sweby :231-245
      phi = fmax2(fmax2(0.0, fmin2(1.5 * ru, 1.0)), fmin2(ru, 1.5));
      break;
    case 8:  // MC :248-262
      phi = fmax2(0.0, fmin2(fmin2(2.0 * ru, 0.5 * (1.0 + ru)), 2.0));
      break;
    case 9:  // Koren :265-279
      phi = fmax2(0.0, fmin2(2.0 * ru, fmin2((1.0 + 2.0 * ru) / 3.0, 2.0)));
      break;
    default:  // superbee :282-295
      phi = fmax2(fmax2(0.0, fmin2(2.0 * ru, 1.0)), fmin2(ru, 2.0));
      break;
  }
  if (ru > 0.0) phi = fmin2(fmin2(phi, cx), cy * ru);
  return phi * x;
}

// Lib_Reconstruction.f90:84-99
static inline void slope_bounds(double dl0, double dl1, double dl2, double dll, double dlr, double& cxl,
                                double& cyl, double& cxr, double& cyr) {
  if (dll > 0.0) { cxl = dl1 / dll; cyl = dl0 / dll; } else { cxl = 2.0; cyl = 2.0; }
  if (dlr > 0.0) { cxr = dl2 / dlr; cyr = dl1 / dlr; } else { cxr = 2.0; cyr = 2.0; }
}

// Lib_Reconstruction.f90:17-78.  Returns false where the stencil is unphysical
// even at first order (report_stencil).
static inline bool reconstruct(int model, int lim, const double* prev, const double* local, const double* next,
                               const double* next2, double dl0, double dl1, double dl2, double dll, double dlr,
                               double beta, double* ql, double* qr) {
  int nc = ncond_of(model);
  double limval = beta;
  double cxl, cyl, cxr, cyr;
  slope_bounds(dl0, dl1, dl2, dll, dlr, cxl, cyl, cxr, cyr);
  bool first_l = prev[0] <= RHO_EMPTY || local[0] <= RHO_EMPTY || next[0] <= RHO_EMPTY;
  bool first_r = local[0] <= RHO_EMPTY || next[0] <= RHO_EMPTY || next2[0] <= RHO_EMPTY;
  if (first_l) for (int v = 0; v < nc; ++v) ql[v] = local[v];
  if (first_r) for (int v = 0; v < nc; ++v) qr[v] = next[v];
  bool ok_l = first_l, ok_r = first_r;
  while (!ok_l || !ok_r) {
    for (int v = 0; v < nc; ++v) {
      double s1 = (next[v] - local[v]) / dl1;
      if (!first_l) {
        double s0 = (local[v] - prev[v]) / dl0;
        double sl = limiter(lim, s1, s0, cxl, cyl) * limval;
        ql[v] = local[v] + sl * dll;
      }
      if (!first_r) {
        double s2 = (next2[v] - next[v]) / dl2;
        double sr = limiter(lim, s2, s1, cxr, cyr) * limval;
        qr[v] = next[v] - sr * dlr;
      }
    }
    if (!first_l) ok_l = check_prim(model, ql);
    if (!first_r) ok_r = check_prim(model, qr);
    limval = 0.5 * limval;
    if (limval < 1.e-18 && !(ok_l && ok_r)) {
      for (int v = 0; v < nc; ++v) { ql[v] = local[v]; qr[v] = next[v]; }
      return check_prim(model, ql) && check_prim(model, qr);
    }
  }
  return true;
}

// Mod_Fluxes.f90:655-685: lengths -> dl0..dlr, reconstruction, Riemann, times the area
static inline bool compute_flux_(int model, int solver, int lim, const double* const* st, const double* len,
                                 const double* nrm, double area, double beta, double* flux, const Mat& m) {
  int nc = ncond_of(model);
  double dl0 = 0.5 * (len[0] + len[1]);
  double dl1 = 0.5 * (len[1] + len[2]);
  double dl2 = 0.5 * (len[2] + len[3]);
  double dll = 0.5 * len[1];
  double dlr = 0.5 * len[2];
  double ql[12], qr[12];
  bool ok = reconstruct(model, lim, st[0], st[1], st[2], st[3], dl0, dl1, dl2, dll, dlr, beta, ql, qr);
  riemann(solver, model, ql, qr, nrm, flux, m);
  for (int v = 0; v < nc; ++v) flux[v] = flux[v] * area;
  return ok;
}

// Lib_Shock_Detector.f90:17-44 (r0 centre, r1/r2 i-1/i+1, r3/r4 j-1/j+1, r5/r6 k-1/k+1)
static inline double sd_rho(double r, double r1, double r2, double r3, double r4, double r5, double r6) {
  double kp1 = std::fabs((r2 - 2.0 * r + r1) / (r2 + 2.0 * r + r1 + 1e-30));
  double kp2 = std::fabs((r4 - 2.0 * r + r3) / (r4 + 2.0 * r + r3 + 1e-30));
  double kp3 = std::fabs((r6 - 2.0 * r + r5) / (r6 + 2.0 * r + r5 + 1e-30));
  double sw = fmax2(fmax2(kp1, kp2), kp3);
  if (sw < 1.0 / 20.0) {
    double phi = fmax2(sw * 20.0, 0.0);
    return 1.0 - std::tanh(10.0 * phi * phi * phi);
  }
  return 0.0;
}

// ------------------------------------------------------------------ RK
// Lib_RK.f90:9-25
static inline double rkc(int nrk, int irk) {
  static const double C[3][3] = {{1.0, 0.0, 0.0}, {1.0, 0.50, 0.0}, {1.0, 0.25, 2.0 / 3.0}};
  return C[nrk - 1][irk - 1];
}

// ------------------------------------------------------------------ drag
// pow through a pointer the compiler cannot see through: g++ folds pow(10, x)
// into exp10(x), which rounds differently from the libm pow flang calls
static double (*volatile pow_libm)(double, double) = std::pow;

// Lib_Drag.f90:83-324
static inline double drag_wen_yu(double Re) {
  if (Re <= 1000) return 24.0 / (Re + TOLL_DRAG) * (1.0 + 0.15 * std::pow(Re, 0.687));
  return 0.43;
}
static inline double henderson1(double Re, double Ma, double G, double Tr) {
  const double t = TOLL_DRAG;
  return 24.0 * (1.0 / (Re + Ma * std::sqrt(0.5 * G) *
                             (4.33 + (3.65 - 1.53 * Tr) / (1.0 + 0.353 * Tr) *
                                         std::exp(-0.247 * Re / (Ma * std::sqrt(0.5 * G) + t))) +
                         t)) +
         std::exp(-0.5 * Ma / (std::sqrt(Re) + t)) *
             ((4.5 + 0.38 * (0.03 * Re + 0.48 * std::sqrt(Re))) / (1.0 + 0.03 * Re + 0.48 * std::sqrt(Re)) +
              0.1 * (Ma * Ma) + 0.2 * std::pow(Ma, 8.0)) +
         0.6 * Ma * std::sqrt(0.5 * G) * (1.0 - std::exp(-Ma / (Re + t)));
}
static inline double henderson2(double Re, double Ma, double G, double Tr) {
  const double t = TOLL_DRAG;
  return (0.9 + 0.34 / ((Ma * Ma) + t) +
          1.86 * std::sqrt(Ma / (Re + t)) *
              (2.0 + 2.0 / (Ma * Ma * 0.5 * G + t) + (1.058 / (Ma * std::sqrt(0.5 * G) + t)) * std::sqrt(Tr)) -
          1.0 / (std::pow(Ma, 4.0) * (G * G) * 0.25 + t)) /
         (1.0 + 1.86 * std::sqrt(Ma / (Re + t)));
}
static inline double drag(double Re, double Ma, double G, double Tr, int sel) {
  const double t = TOLL_DRAG;
  switch (sel) {
    case 1: return 0.45;
    case 2: return 24.0 / (Re + t);
    case 3: return 24.0 / (Re + t) * (1.0 + 3.0 * Re / 16.0);
    case 4: return 24.0 / (Re + t) * (1.0 + 0.15 * std::pow(Re, 0.687));
    case 5: return drag_wen_yu(Re);
    case 6:
      if (Re < 1000) return 24.0 / (Re + t) * (1.0 + std::pow(Re, 2.0 / 3.0) / 6.0);
      return 0.4392;
    case 7:
      return 24.0 / (Re + t) *
             (1.0 + 0.15 * std::pow(Re, 0.687) + 0.0175 * Re / (1.0 + 4.25 * 1e4 * std::pow(Re, -1.16)));
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
      return Cd0 * (1.0 + std::exp(-0.427 / (std::pow(Ma, 4.63) + t) - 3.0 / (std::pow(Re, 0.88) + t))) /
             (1.0 + Ma / (Re + t) * (3.82 + 1.28 * std::exp(-1.25 * Re / (Ma + t))));
    }
    case 10: {
      if (Ma <= 1) return henderson1(Re, Ma, G, Tr);
      if (Ma >= 1.75) return henderson2(Re, Ma, G, Tr);
      double Cd1 = henderson1(Re, 1.0, G, Tr), Cd2 = henderson2(Re, 1.75, G, Tr);
      return Cd1 + 4.0 / 3.0 * (Ma - 1.0) * (Cd2 - Cd1);
    }
    case 11: {
      double g = 1.25 * (1.0 + std::tanh(0.77 * std::log10(Re + t) - 1.92));
      g = pow_libm(10.0, g);
      double h = 2.3 + 1.7 * std::sqrt(Tr) - 2.3 * std::tanh(1.17 * std::log10(Ma + t));
      double Cd0 = drag_wen_yu(Re);
      return 2.0 + (Cd0 - 2.0) * std::exp(-3.07 * std::sqrt(G) * Ma * g / (Re + t)) +
             h / (Ma * std::sqrt(G) + t) * std::exp(-Re / (2 * Ma + t));
    }
    case 12: {
      double g = (1.0 + Re * (12.278 + 0.548 * Re)) / (1.0 + 11.278 * Re);
      double h = 5.6 / (1.0 + Ma) + 1.7 * std::sqrt(Tr);
      double Cd0 = drag_wen_yu(Re);
      return 2.0 + (Cd0 - 2.0) * std::exp(-3.07 * std::sqrt(G) * Ma * g / (Re + t)) +
             h / (Ma * std::sqrt(G) + t) * std::exp(-Re / (2 * Ma + t));
    }
    default: return 0.0;  // 13 NoDrag
  }
}

// ------------------------------------------------------------------ heat
// Lib_Heat.f90:68-160
static inline double heat(double Re, double Pr, double Ma, int sel) {
  const double t = TOLL_HEAT;
  switch (sel) {
    case 1: return 2.0;
    case 2: return 2.5 * std::pow(Re, 0.15) + 0.04 * Re;
    case 3: return 2.0 + 0.37 * std::pow(Re, 0.6) * std::pow(Pr, 1.0 / 3.0);
    case 4: return 2.0 + 0.459 * std::pow(Re, 0.55) * std::pow(Pr, 1.0 / 3.0);
    case 5:
      return 1.0 / (1.0 / (2.0 + 0.654 * std::sqrt(Re) * std::sqrt(Pr)) + 3.42 * Ma / (Re * Pr + t));
    case 6: return 2.0 + 0.6 * std::sqrt(Re) * std::pow(Pr, 1.0 / 3.0);
    case 7: {
      double Nu = 2.0 + 0.459 * std::pow(Re, 0.55) * std::pow(Pr, 0.33);
      return Nu / (1.0 + 3.42 * Ma / (Re * Pr + t) * Nu);
    }
    default: return 0.0;  // 8 NoHeat
  }
}

}  // namespace orc
