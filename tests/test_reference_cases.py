"""The reference's invariant-checked cases, restated as tests: test/Axis/{IG,MK,AG}
(a uniform or radially layered cloud in a one-degree wedge about x, the 200 side
faces and the 300 axis row), test/Reflect/AG (a sheared AG blob along a 200
plane) and test/Thermal/{IG,AG} (a pressurised blob in a closed box).  The
fixtures (tests/golden/refcase_*.npz, make_reference_cases.py) hold the
reference's input files; each test holds the run to the PASS criterion of the
case's verify.py:

* Axis IG, AG (verify.py:43-60): after the 500 steady RK2 iterations, over every
  cell max(|u|,|v|,|w|)/c, max|rho - 1| and max|P - 1| (AG: P11, P22, P33) below
  1e-10, c = sqrt(3P/rho);
* Axis MK (verify.py:49-53): after 200 steady RK3 iterations the axis row j = 1
  keeps its initial density, max|rho - rho_ic| < 1e-9;
* Reflect AG (verify.py:30-41): RK3 to t = 0.1, the x-momentum sum(rho u V)
  changes by less than 1e-9 relative;
* Thermal IG, AG (verify.py:35-50): RK3 to t = 0.1, max|T - 1| < 1e-9 over the
  cells with rho > 1e-3, and the blob has expanded (max|u| >= 0.1);
* Riemann AG (verify.py:70-100): an anisotropic Sod tube, explicit Euler at first
  order with Rusanov, against the exact gamma = 3 Riemann solution (Toro ch. 4,
  restated here and checked on Toro's Sod values): min(P11) > 1e3, L1 density
  error <= 0.05, P22/rho kept to 1e-9, total variation of rho and P11 at most
  1.05 times the exact one;
* Berthon SCS, RCS, RCR, SCS-hlle (berthon.py:87-110): AG Riemann problems to
  t = 0.125 against the exact profiles the reference tabulates, the L1 error of
  rho, u, v, P11, P22 and det(P) over the span of each exact profile within
  2.5e-2 (SCS) or 1.5e-2 (RCS, RCR).

The wedge is the curvilinear mesh of these cases: its metrics are also held to
their closed forms (a prism per cell), and the volume formula to the divergence
theorem on a mesh with twisted faces."""
import json
import os

import numpy as np
import pytest

from ice_amd import cases, geometry, native

from helpers import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name, mode=0):
    d = np.load(os.path.join(HERE, "golden", f"refcase_{name}.npz"))
    ini = json.loads(str(d["ini"]))
    case = cases.case_from_inputs(name, ini, [d["nodes"]], [d["ic"]], d["reci"], d["recd"], d["donor"], d["weight"],
                                  float(d["rho_mat"]), float(d["cs"]), mode=mode)
    return case, ini, d


def run(lib, prefix, case, ini):
    """Steps the case as ICE_solve would: iter-threshold iterations, or time-accurate
    steps until the time passes time-threshold.  Returns the runner and status."""
    r = cases.Runner(lib, case, prefix)
    par = ini["ICE-Parameters"]
    if "iter-threshold" in par:
        st = r.step(int(par["iter-threshold"]))
    else:
        t_end = float(par["time-threshold"])
        st = r.step(1)
        n = 1
        while st.time < t_end:
            st = r.step(1)
            n += 1
            assert n < 100000
    assert st.bad_state == 0 and st.bad_recon == 0
    return r, st


def interior(r):
    """(nc, ni, nj) of the one block's interior, family 1 (primitive slots)"""
    return r.state(1, 1)[:, :, :, 0]


def check_axis_rest(q, closure):
    rho = q[0]
    vel = np.abs(q[1:4]).max(axis=0)
    pcols = {native.IG: [4], native.AG: [4, 7, 9]}[closure]
    P = [q[c] for c in pcols]
    c = np.sqrt(3.0 * sum(P) / len(P) / rho)
    mach, drho = float(np.max(vel / c)), float(np.abs(rho - 1.0).max())
    dP = max(float(np.abs(p - 1.0).max()) for p in P)
    print(f"max |u|/c {mach:.3e}, max|rho - 1| {drho:.3e}, max|P - 1| {dP:.3e}")
    assert np.isfinite(q).all()
    assert mach < 1e-10 and drho < 1e-10 and dP < 1e-10


def check_axis_mk(q, ic):
    d_axis = float(np.abs(q[0][:, 0] - ic[0][:, 0, 0]).max())
    print(f"axis row max|rho - rho_ic| {d_axis:.3e}")
    assert np.isfinite(q).all() and d_axis < 1e-9


def check_reflect(q, ic):
    m0, m1 = float((ic[0] * ic[1]).sum()), float((q[0] * q[1]).sum())  # uniform cells: V cancels
    drift = abs(m1 - m0) / abs(m0)
    print(f"x-momentum relative change {drift:.3e}")
    assert np.isfinite(q).all() and drift < 1e-9


def check_thermal(q, closure):
    tcol = {native.IG: 5, native.AG: 10}[closure]
    live = q[0] > 1e-3
    dT = float(np.abs(q[tcol][live] - 1.0).max())
    umax = float(np.sqrt(q[1] ** 2 + q[2] ** 2 + q[3] ** 2).max())
    print(f"max|T - 1| {dT:.3e} over {int(live.sum())} cells, max|u| {umax:.3e}")
    assert np.isfinite(q).all() and umax >= 0.1 and dT < 1e-9


def exact_riemann(WL, WR, gam):
    """The exact Riemann solution of the Euler equations, W = (rho, u, p) (Toro,
    Riemann Solvers, ch. 4): Newton on f_L(p) + f_R(p) + u_R - u_L = 0, then the
    wave pattern sampled at S = (x - x0)/t.  Returns (p*, u*, sample)."""
    rL, uL, pL = WL
    rR, uR, pR = WR
    cL, cR = np.sqrt(gam * pL / rL), np.sqrt(gam * pR / rR)

    def f(p, r, pk, c):
        if p > pk:  # shock
            A, B = 2.0 / ((gam + 1.0) * r), (gam - 1.0) / (gam + 1.0) * pk
            q = np.sqrt(A / (p + B))
            return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + B))
        e = (gam - 1.0) / (2.0 * gam)  # rarefaction
        return 2.0 * c / (gam - 1.0) * ((p / pk) ** e - 1.0), (p / pk) ** (-(gam + 1.0) / (2.0 * gam)) / (r * c)

    p = 0.5 * (pL + pR)
    for _ in range(100):
        fl, dl = f(p, rL, pL, cL)
        fr, dr = f(p, rR, pR, cR)
        pn = max(p - (fl + fr + uR - uL) / (dl + dr), 1e-12 * p)
        if abs(pn - p) <= 1e-14 * p:
            p = pn
            break
        p = pn
    us = 0.5 * (uL + uR) + 0.5 * (f(p, rR, pR, cR)[0] - f(p, rL, pL, cL)[0])
    g = gam

    def side(S, r, u, pk, c, sgn):
        """the state at S on the side of sign sgn (-1 left, +1 right) of the contact"""
        if p > pk:  # shock
            Ssh = u + sgn * c * np.sqrt((g + 1) / (2 * g) * p / pk + (g - 1) / (2 * g))
            if sgn * (S - Ssh) >= 0:
                return r, u, pk
            q = p / pk
            return r * (q + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * q + 1), us, p
        cs = c * (p / pk) ** ((g - 1) / (2 * g))
        head, tail = u + sgn * c, us + sgn * cs
        if sgn * (S - head) >= 0:
            return r, u, pk
        if sgn * (S - tail) <= 0:
            return r * (p / pk) ** (1 / g), us, p
        b = 2 / (g + 1) - sgn * (g - 1) / ((g + 1) * c) * (u - S)
        return r * b ** (2 / (g - 1)), 2 / (g + 1) * (-sgn * c + (g - 1) / 2 * u + S), pk * b ** (2 * g / (g - 1))

    def sample(S):
        return side(S, rL, uL, pL, cL, -1) if S <= us else side(S, rR, uR, pR, cR, 1)

    return p, us, sample


def test_exact_riemann_solver_on_toro_sod():
    """Toro's Sod problem (Table 4.3, gamma = 1.4): p* = 0.30313, u* = 0.92745;
    and with equal states the solution is that state everywhere."""
    ps, us, sample = exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    assert abs(ps - 0.30313) < 1e-5 and abs(us - 0.92745) < 1e-5
    assert np.allclose(sample(-10.0), (1.0, 0.0, 1.0)) and np.allclose(sample(10.0), (0.125, 0.0, 0.1))
    ps, us, sample = exact_riemann((2.0, 0.5, 3.0), (2.0, 0.5, 3.0), 3.0)
    assert abs(ps - 3.0) < 1e-12 and abs(us - 0.5) < 1e-12 and np.allclose(sample(0.3), (2.0, 0.5, 3.0))


def cell_x(d):
    x = d["nodes"][0][:, 0, 0]
    return 0.5 * (x[:-1] + x[1:])


def check_riemann_ag(q, d, t):
    ic = d["ic"][:, :, :, 0]
    xc = cell_x(d)
    jump = int(np.argmax(np.abs(np.diff(ic[0][:, 0]))))
    x0 = d["nodes"][0][jump + 1, 0, 0]
    WL, WR = (ic[0][0, 0], 0.0, ic[4][0, 0]), (ic[0][-1, 0], 0.0, ic[4][-1, 0])
    ratio = float(ic[7][0, 0] / ic[0][0, 0])
    _, _, sample = exact_riemann(WL, WR, 3.0)
    ex = np.array([sample((x - x0) / t) for x in xc])
    rho, P11, P22 = q[0], q[4], q[7]
    l1 = float(np.abs(rho.mean(axis=1) - ex[:, 0]).sum() / ex[:, 0].sum())
    tv = lambda a: float(np.abs(np.diff(a)).sum())  # noqa: E731
    tv_rho, tv_p = tv(rho.mean(axis=1)) / tv(ex[:, 0]), tv(P11.mean(axis=1)) / tv(ex[:, 2])
    dpt = float(np.abs(P22 / rho - ratio).max())
    print(f"min P11 {P11.min():.3e}, L1(rho) {l1:.3e}, max|P22/rho - {ratio}| {dpt:.3e}, TV {tv_rho:.4f} {tv_p:.4f}")
    assert np.isfinite(q).all() and P11.min() > 1e3 and l1 <= 0.05 and dpt < 1e-9
    assert tv_rho <= 1.05 and tv_p <= 1.05


BERTHON_TOL = {"Berthon_SCS": 2.5e-2, "Berthon_RCS": 1.5e-2, "Berthon_RCR": 1.5e-2, "Berthon_SCS-hlle": 2.5e-2}
BERTHON_Q = {"Density": lambda q: q[0], "X Velocity": lambda q: q[1], "Y Velocity": lambda q: q[2],
             "XX Pressure": lambda q: q[4], "YY Pressure": lambda q: q[7],
             "Det Pressure": lambda q: q[4] * q[7] - q[5] * q[5]}


def sample_profile(xe, ve, x):
    """A piecewise-linear profile at x; repeated abscissae are jumps."""
    if x <= xe[0]:
        return ve[0]
    if x >= xe[-1]:
        return ve[-1]
    for i in range(len(xe) - 1):
        if xe[i] <= x <= xe[i + 1] and xe[i + 1] > xe[i]:
            return ve[i] + (ve[i + 1] - ve[i]) * (x - xe[i]) / (xe[i + 1] - xe[i])
    return ve[-1]


def check_berthon(name, q, d):
    xc = cell_x(d)
    worst = 0.0
    for zone, get in BERTHON_Q.items():
        xe, ve = d[f"exact_x_{zone}"], d[f"exact_v_{zone}"]
        ref = np.array([sample_profile(xe, ve, x) for x in xc])
        num = get(q[:, :, 0])
        err = float(np.abs(num - ref).sum() / (len(xc) * (ve.max() - ve.min())))
        print(f"  {zone}: L1 {err:.3e}")
        worst = max(worst, err)
    assert np.isfinite(q).all() and worst <= BERTHON_TOL[name]


def verify(name, case, ini, d, r, st):
    q = interior(r)
    closure = case.fams[0].closure
    t_end = float(ini["ICE-Parameters"].get("time-threshold", "0"))
    assert st.time >= t_end * (1.0 - 1e-12)
    if name in ("Axis_IG", "Axis_AG"):
        check_axis_rest(q, closure)
    elif name == "Axis_MK":
        check_axis_mk(q, d["ic"])
    elif name == "Reflect_AG":
        check_reflect(q, d["ic"][:, :, :, 0])
    elif name == "Riemann_AG":
        check_riemann_ag(q, d, st.time)
    elif name.startswith("Berthon"):
        check_berthon(name, q, d)
    else:
        check_thermal(q, closure)


NAMES = ["Axis_IG", "Axis_MK", "Axis_AG", "Reflect_AG", "Thermal_IG", "Thermal_AG", "Riemann_AG", "Berthon_SCS",
         "Berthon_RCS", "Berthon_RCR", "Berthon_SCS-hlle"]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_passes_reference_verify(oracle, name):
    case, ini, d = load(name)
    r, st = run(oracle, "orc_", case, ini)
    print(f"{name}: {st.iter} steps, t = {st.time:.6e}")
    verify(name, case, ini, d, r, st)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("name", NAMES)
def test_gpu_passes_reference_verify(gpu, oracle, name, mode):
    print(f"case {name} mode {mode}", flush=True)
    case, ini, d = load(name, mode=mode)
    rg, sg = run(gpu, "ice_gpu_", case, ini)
    verify(name, case, ini, d, rg, sg)
    ro, so = run(oracle, "orc_", load(name)[0], ini)
    assert sg.iter == so.iter
    a, b = rg.state(1, 1), ro.state(1, 1)
    err = max(rel_l2(a[v], b[v]) for v in range(a.shape[0]))
    print(f"{name}: GPU against the oracle, max relative L2 {err:.3e}")
    assert err <= 1e-12


# ---------------------------------------------------------------- the wedge's metrics
def wedge_cells(nodes):
    """Per cell of the wedge: x range, inner and outer radius, angle range (from its nodes)."""
    x, y, z = nodes
    r = np.hypot(y, z)
    th = np.arctan2(z, y)
    return x, r, th


def test_wedge_metrics_closed_forms():
    """Each wedge cell is a prism: a trapezoid of the (y, z) plane, (r2^2 - r1^2)
    sin(dth)/2, times its length along x.  Its i faces are those trapezoids, its
    j faces chords 2 r sin(dth/2) long, its k faces (r2 - r1) wide, with the
    normals of the planes they lie in."""
    d = np.load(os.path.join(HERE, "golden", "refcase_Axis_IG.npz"))
    nodes = d["nodes"]
    g = geometry.metrics(nodes)
    ni, nj, nk = g["dims"]
    x, r, th = wedge_cells(nodes)
    x1, x2 = x[:-1, :-1, 0], x[1:, :-1, 0]
    r1, r2 = r[:-1, :-1, 0], r[:-1, 1:, 0]
    dth = th[0, 0, 1] - th[0, 0, 0]
    vol = 0.5 * (r2 ** 2 - r1 ** 2) * np.sin(dth) * (x2 - x1)
    V = g["vol"].reshape((ni, nj, nk), order="F")[:, :, 0]
    assert np.max(np.abs(V - vol) / vol) < 1e-12
    # i faces (0..ni, 1..nj, 1..nk): {N(3), A}
    fi = g["face_i"].reshape((4, ni + 1, nj, nk), order="F")[:, :, :, 0]
    area_i = 0.5 * (r[:, 1:, 0] ** 2 - r[:, :-1, 0] ** 2) * np.sin(dth)
    assert np.max(np.abs(fi[3] - area_i) / area_i) < 1e-12
    assert np.allclose(fi[0], 1.0, rtol=0, atol=1e-15) and np.allclose(fi[1:3], 0.0, rtol=0, atol=1e-15)
    # j faces (1..ni, 0..nj, 1..nk): chords at radius r
    fj = g["face_j"].reshape((4, ni, nj + 1, nk), order="F")[:, :, :, 0]
    area_j = 2.0 * r[:-1, :, 0] * np.sin(dth / 2) * (x[1:, :, 0] - x[:-1, :, 0])
    assert np.max(np.abs(fj[3] - area_j) / area_j) < 1e-12
    assert np.allclose(fj[1], 1.0, rtol=0, atol=1e-15) and np.allclose(fj[[0, 2]], 0.0, rtol=0, atol=1e-15)
    # k faces (1..ni, 1..nj, 0..nk): the two side planes at angles th0 and th1
    fk = g["face_k"].reshape((4, ni, nj, nk + 1), order="F")
    for kk in range(nk + 1):
        t = th[0, 0, kk]
        area_k = (r2 - r1) * (x2 - x1)
        assert np.max(np.abs(fk[3, :, :, kk] - area_k) / area_k) < 1e-12
        assert np.allclose(fk[0, :, :, kk], 0.0, atol=1e-15)
        assert np.allclose(fk[1, :, :, kk], -np.sin(t), rtol=0, atol=1e-15)
        assert np.allclose(fk[2, :, :, kk], np.cos(t), rtol=0, atol=1e-15)


def twisted_nodes(n=(5, 4, 3), seed=7, amp=0.3):
    """A block whose nodes are a unit lattice moved at random by up to amp, so that
    its faces are twisted (not planar)."""
    rng = np.random.default_rng(seed)
    X = geometry.cartesian_nodes(*[np.arange(m + 1, dtype=float) for m in n])
    return X + amp * rng.uniform(-1.0, 1.0, X.shape)


def patch_flux(p00, p10, p01, p11):
    """Integral of x . (x_u x x_v) over the bilinear patch x(u, v) of its four
    corners, by 2 x 2 Gauss points (the integrand has degree 2 in u and in v)."""
    g = 0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)
    tot = 0.0
    for u in g:
        for v in g:
            x = p00 * (1 - u) * (1 - v) + p10 * u * (1 - v) + p01 * (1 - u) * v + p11 * u * v
            xu = (p10 - p00) * (1 - v) + (p11 - p01) * v
            xv = (p01 - p00) * (1 - u) + (p11 - p10) * u
            tot = tot + 0.25 * np.einsum("i...,i...->...", x, np.cross(xu, xv, axis=0))
    return tot


def test_volumes_by_divergence_theorem_on_twisted_cells():
    """Compute_metrics' volume (Mod_Metrics.f90:622-660) is the integral of det J
    over the trilinear cell, exact with twisted faces: each cell's volume equals
    one third of the flux of x through its six bilinear faces (divergence
    theorem), computed here by quadrature, so the cells' volumes sum to the
    volume the block's boundary faces enclose; and each closed cell's area
    vectors sum to zero."""
    X = twisted_nodes()
    g = geometry.metrics(X)
    ni, nj, nk = g["dims"]
    V = g["vol"].reshape((ni, nj, nk), order="F")
    c = lambda a, b, e: X[:, a:ni + a, b:nj + b, e:nk + e]  # noqa: E731  corner (i-1+a, j-1+b, k-1+e)
    F = (patch_flux(c(1, 0, 0), c(1, 1, 0), c(1, 0, 1), c(1, 1, 1)) - patch_flux(c(0, 0, 0), c(0, 1, 0), c(0, 0, 1), c(0, 1, 1))
         + patch_flux(c(0, 1, 0), c(0, 1, 1), c(1, 1, 0), c(1, 1, 1)) - patch_flux(c(0, 0, 0), c(0, 0, 1), c(1, 0, 0), c(1, 0, 1))
         + patch_flux(c(0, 0, 1), c(1, 0, 1), c(0, 1, 1), c(1, 1, 1)) - patch_flux(c(0, 0, 0), c(1, 0, 0), c(0, 1, 0), c(1, 1, 0)))
    Vd = F / 3.0
    assert np.max(np.abs(V - Vd) / Vd) < 1e-13
    # closed cells: the six area vectors (A N, signed along +d) sum to zero
    fi = g["face_i"].reshape((4, ni + 1, nj, nk), order="F")
    fj = g["face_j"].reshape((4, ni, nj + 1, nk), order="F")
    fk = g["face_k"].reshape((4, ni, nj, nk + 1), order="F")
    s = (fi[:3, 1:] * fi[3, 1:] - fi[:3, :-1] * fi[3, :-1] + fj[:3, :, 1:] * fj[3, :, 1:] - fj[:3, :, :-1] * fj[3, :, :-1]
         + fk[:3, :, :, 1:] * fk[3, :, :, 1:] - fk[:3, :, :, :-1] * fk[3, :, :, :-1])
    assert np.max(np.abs(s)) < 1e-14 * np.max(fi[3])
