"""Closed-form verification cases of ICE (test/verification/*/run.py), run on the
oracle and on the device from the same inputs.

* B, thermal relaxation (B-thermal-relaxation/run.py:1-12): a cloud moving with
  the gas (no slip, so no drag), hotter than the gas, Stokes heat (Nu = 2),
  no radiation.  T_p(t) = T_g + (T_p0 - T_g) exp(-t / tau_T), tau_T =
  rho_al cs dp^2 / (12 k_g); velocity, density and number density do not move.
  The case's bar: 1e-5 of the initial gap (RK2 at dt = 1e-4 s, tau_T = 0.029 s).
* J, the time step bounded by the relaxation time (J-tau-limit/run.py:1-25): the
  first step is dt-max, the later ones tau_p, and the slip follows the SSP-RK3
  amplification of each step (below).
* A, Stokes drag relaxation, C, periodic sine advection, G, a cloud carried by a
  frozen gas on a periodic line, H, a cloud in a linear strain, and I, a cloud in
  a 2-D vortex: the closed forms and observed orders of those cases (below, each
  with its bars)."""
import math

import numpy as np
import pytest

from ice_amd import cases
from ice_amd.native import MK

RHO_AL, CS, DP, ALPHA = 1000.0, 900.0, 1.0e-4, 1.0e-3
TG, TP0, UG, KG = 300.0, 400.0, 10.0, 0.026
TAU_T = RHO_AL * CS * DP ** 2 / (12.0 * KG)


def thermal_case(steps, mode=0):
    case = cases.case_duct400(MK, u=UG, gas_u=UG, steps=steps, mode=mode)
    rp = 0.5 * DP
    rho_p = ALPHA * RHO_AL
    n = rho_p / (RHO_AL * 4.0 / 3.0 * math.pi * rp ** 3)
    prim = case.blocks[0].prim[0].reshape((6, 20, 5, 5), order="F")
    prim[0, 2:18], prim[1, 2:18], prim[4, 2:18], prim[5, 2:18] = rho_p, UG, TP0, n
    case.blocks[0].prim[0] = np.ascontiguousarray(prim.ravel(order="F"))
    case.cfg.heat_select = 1    # Stokes: Nu = 2
    case.cfg.drag_select = 2    # Stokes
    case.cfg.dt_max = 1.0e-4
    return case, rho_p, n


def check_thermal(runner, st, rho_p, n):
    q = runner.state(1, 1)
    exact = TG + (TP0 - TG) * math.exp(-st.time / TAU_T)
    T = q[4]
    assert np.max(T) - np.min(T) <= 1e-10 * TP0
    assert np.max(np.abs(T - exact)) <= 1e-5 * (TP0 - TG), (T[0, 0, 0], exact)
    assert np.max(np.abs(q[1] - UG)) <= 1e-12 * UG
    assert np.max(np.abs(q[0] - rho_p)) <= 1e-12 * rho_p
    assert np.max(np.abs(q[5] - n)) <= 1e-9 * n


@pytest.mark.parametrize("ratio", [0.5, 2.0])
def test_thermal_relaxation_oracle(oracle, ratio):
    steps = int(round(ratio * TAU_T / 1.0e-4))
    case, rho_p, n = thermal_case(steps)
    r = cases.Runner(oracle, case, "orc_")
    st = r.step(steps)
    assert abs(st.time - steps * 1.0e-4) <= 1e-12
    check_thermal(r, st, rho_p, n)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_thermal_relaxation_gpu(gpu, mode):
    steps = int(round(1.0 * TAU_T / 1.0e-4))
    case, rho_p, n = thermal_case(steps, mode)
    r = cases.Runner(gpu, case, "ice_gpu_")
    st = r.step(steps)
    check_thermal(r, st, rho_p, n)


# J, the time step bounded by the relaxation time (J-tau-limit/run.py:1-25): a cloud at
# rest in a 10 m/s gas, Stokes drag, tau_p = rho_al dp^2 / (18 mu) = 2.2e-5 s = dt-max / 4.5.
# The first step, before tau is known, is dt-max and overshoots: the SSP-RK3 amplification
# R(-4.5) = -8.5625 puts u_p at 10 (1 - R) = 95.625.  Every later step is tau_p (tau-factor
# 1), so (t - dt-max) / tau_p is an integer and the slip shrinks by R(-1) = 1/3 a step.
RHO_ALJ, DPJ, MUJ, RHOGJ = 2000.0, 2.0e-6, 2.0e-5, 1.0
TAU_P = RHO_ALJ * DPJ ** 2 / (18.0 * MUJ)


def tau_case(mode=0):
    case = cases.case_duct400(MK, u=0.0, gas_u=UG, steps=5, mode=mode)
    rp = 0.5 * DPJ
    rho_p = 1.0e-3
    n = rho_p / (RHO_ALJ * 4.0 / 3.0 * math.pi * rp ** 3)
    prim = case.blocks[0].prim[0].reshape((6, 20, 5, 5), order="F")
    prim[0, 2:18], prim[1, 2:18], prim[4, 2:18], prim[5, 2:18] = rho_p, 0.0, 300.0, n
    case.blocks[0].prim[0] = np.ascontiguousarray(prim.ravel(order="F"))
    g5 = np.zeros((5,) + case.blocks[0].dims)
    g5[0], g5[1], g5[4] = RHOGJ, UG, 300.0
    case.blocks[0].gas = cases.gas_arrays(g5, 287.05, 1.4, 0.025, MUJ)
    case.fams[0].rho_mat = RHO_ALJ
    case.fams[0].cs = 1000.0
    case.cfg.nrk = 3
    case.cfg.drag_select = 2
    case.cfg.heat_select = 1
    case.cfg.dt_max = 1.0e-4
    case.cfg.tau_factor = 1.0
    return case


def check_tau(runner, st, steps):
    R = lambda z: 1.0 + z + z * z / 2.0 + z ** 3 / 6.0  # noqa: E731
    u1 = UG * (1.0 - R(-1.0e-4 / TAU_P))
    assert abs(u1 - 95.625) < 0.01
    k = (st.time - 1.0e-4) / TAU_P
    assert abs(k - (steps - 1)) <= 1e-6, k
    want = UG - (UG - u1) * R(-1.0) ** (steps - 1)
    u = runner.state(1, 1)[1]
    assert np.max(np.abs(u - want)) <= 1e-9 * abs(want), (u[0, 0, 0], want)


def test_tau_limit_oracle(oracle):
    r = cases.Runner(oracle, tau_case(), "orc_")
    st = r.step(1)
    assert abs(r.state(1, 1)[1][0, 0, 0] - 95.625) <= 1e-9 * 95.625
    st = r.step(4)
    check_tau(r, st, 5)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_tau_limit_gpu(gpu, mode):
    r = cases.Runner(gpu, tau_case(mode), "ice_gpu_")
    st = r.step(5)
    check_tau(r, st, 5)


# C, a sine wave of density carried at a uniform speed on a periodic mesh, no gas
# (C-advection-sine/run.py:1-90): the exact cell averages of rho_0 + A sin(k (x - U t)),
# mass kept to 1e-12, the velocity uniform to 1e-10, and the observed order of the
# space error at dt = dt-max = 1e-4 s: >= 1.5 for MUSCL van Leer, >= 0.85 first order.
L_C, SPEED, RHO0, AMP = 1.0, 1.0, 10.0, 5.0
K_C = 2.0 * math.pi / L_C


def sine_average(xl, xr, t):
    return RHO0 + AMP * (np.cos(K_C * (xl - SPEED * t)) - np.cos(K_C * (xr - SPEED * t))) / (K_C * (xr - xl))


def periodic_line(nx, averages, u0, limiter=3, steps=1, dt_max=1.0e-4, cfl=0.8, gas_u=None, mode=0, periodic=True,
                  dp=DP):
    """A line of nx cells on [0, L_C] whose faces 1 and 2 are joined (201, periodic;
    with periodic False, extrapolation 400), faces 3-4 extrapolation and 5-6 type 0:
    MK particles of diameter dp at the cell averages `averages(xl, xr)` of density,
    moving at u0 (a number or a function of the cell centre), at T_g; with `gas_u`
    (likewise), a frozen gas (1.2 kg/m3, 300 K) at that speed drags them (Stokes)."""
    from ice_amd import geometry, native
    dims = (nx, 1, 1)
    dx = L_C / nx
    nodes = geometry.cartesian_nodes(np.linspace(0.0, L_C, nx + 1), [0.0, dx], [0.0, dx])
    geo = geometry.metrics(nodes)
    vol = RHO_AL * 4.0 / 3.0 * math.pi * (0.5 * dp) ** 3
    at = lambda f, x: f(x) if callable(f) else f + 0.0 * x  # noqa: E731

    def fill(I, J, K):
        rho = averages((I - 1) * dx, I * dx)
        z = np.zeros(I.shape)
        return np.stack([rho, at(u0, (I - 0.5) * dx), z, z, z + TG, rho / vol])

    prim = cases.padded_state(dims, 6, fill)
    tb = cases.TableBuilder(1)

    def connect(ri, fijk):
        f = fijk[:, 0]
        per = f <= 2
        ri[per, native.R_BS] = 1
        ri[per, native.R_FS] = np.where(f[per] == 1, 2, 1)
        ri[per, native.R_IS] = np.where(f[per] == 1, nx, 1)
        ri[per, native.R_JS] = fijk[per, 2]
        ri[per, native.R_KS] = fijk[per, 3]

    side = 201 if periodic else 400
    tb.add_block(1, dims, lambda f, i, j, k: np.where(f <= 2, side, np.where(f <= 4, 400, 0)),
                 connect if periodic else None)
    reci, recd, donor, weight = tb.arrays()
    gas = None
    if gas_u is not None:
        g5 = np.zeros((5,) + dims)
        g5[0], g5[1], g5[4] = 1.2, at(gas_u, (np.arange(nx)[:, None, None] + 0.5) * dx), TG
        gas = cases.gas_arrays(g5, 287.05, 1.4, KG, 1.8e-5)
    cfg = cases.make_config(ngroups=1, nrk=2, time_accurate=True, ndir=1, limiter=limiter, coupled=gas is not None,
                            drag_select=2, heat_select=1, cfl=cfl, dt_max=dt_max, mode=mode)
    fam = cases.make_family(MK, rho_mat=RHO_AL, cs=CS)
    return cases.Case("line", cfg, [fam], [cases.Block(dims, geo, [prim], gas)], reci, recd, donor, weight, steps)


def sine_case(nx, limiter, steps, mode=0):
    return periodic_line(nx, lambda xl, xr: sine_average(xl, xr, 0.0), SPEED, limiter, steps, mode=mode)


SINE_GRIDS = (25, 50, 100, 200)
ICE_LIM_IORD, ICE_LIM_VANLEER = 0, 3  # first order, van Leer (Lib_Limiters.f90:35-84)


def sine_errors(lib, prefix, limiter, steps, mode=0):
    errs = []
    for nx in SINE_GRIDS:
        r = cases.Runner(lib, sine_case(nx, limiter, steps, mode), prefix)
        st = r.step(steps)
        q = r.state(1, 1)
        dx = L_C / nx
        x = np.arange(nx) * dx
        rho = q[0][:, 0, 0]
        ref = sine_average(x, x + dx, st.time)
        errs.append(np.abs(rho - ref).sum() / nx / AMP)
        mass0 = sine_average(x, x + dx, 0.0).sum() * dx
        assert abs(rho.sum() * dx - mass0) <= 1e-12 * mass0, nx
        assert np.max(np.abs(q[1] - SPEED)) <= 1e-10 * SPEED, nx
        assert abs(st.time - steps * 1.0e-4) <= 1e-12
    return np.array(errs)


def check_sine_orders(muscl, first):
    orders = lambda e: np.log(e[:-1] / e[1:]) / np.log(2.0)  # noqa: E731
    assert orders(muscl).min() >= 1.5, (muscl, orders(muscl))
    assert orders(first).min() >= 0.85, (first, orders(first))
    assert np.all(muscl < first)


def test_sine_advection_oracle(oracle):
    steps = 2500  # t = 0.25 s, a quarter of a period
    check_sine_orders(sine_errors(oracle, "orc_", ICE_LIM_VANLEER, steps), sine_errors(oracle, "orc_", ICE_LIM_IORD, steps))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_sine_advection_gpu(gpu, mode):
    steps = 2500
    check_sine_orders(sine_errors(gpu, "ice_gpu_", ICE_LIM_VANLEER, steps, mode),
                      sine_errors(gpu, "ice_gpu_", ICE_LIM_IORD, steps, mode))


# A, Stokes drag relaxation (A-drag-relaxation/run.py:1-16, 44-135): a cloud at rest in
# a 10 m/s gas, at the gas temperature.  u_p(t) = u_g (1 - exp(-t / tau_p)), tau_p =
# rho_al dp^2 / (18 mu); density, number density and temperature do not move (the drag
# work cancels the kinetic energy it makes).  Bars: 1e-5 of u_g at dt = 1e-4 s, and the
# observed order of the time error over t = 0.02 s as dt halves: RK2 >= 1.65, RK3 >= 2.65.
MU_A = 1.8e-5
TAU_A = RHO_AL * DP ** 2 / (18.0 * MU_A)


def drag_case(steps, dt=1.0e-4, nrk=2, mode=0):
    case, rho_p, n = thermal_case(steps, mode)
    prim = case.blocks[0].prim[0].reshape((6, 20, 5, 5), order="F")
    prim[1, 2:18], prim[4, 2:18] = 0.0, TG
    case.blocks[0].prim[0] = np.ascontiguousarray(prim.ravel(order="F"))
    case.cfg.dt_max = dt
    case.cfg.nrk = nrk
    return case, rho_p, n


def drag_run(lib, prefix, steps, dt=1.0e-4, nrk=2, mode=0):
    case, rho_p, n = drag_case(steps, dt, nrk, mode)
    r = cases.Runner(lib, case, prefix)
    st = r.step(steps)
    assert abs(st.time - steps * dt) <= 1e-12
    q = r.state(1, 1)
    exact = UG * (1.0 - math.exp(-st.time / TAU_A))
    u = q[1]
    assert np.max(u) - np.min(u) <= 1e-12 * UG
    assert np.max(np.abs(q[0] - rho_p)) <= 1e-12 * rho_p
    assert np.max(np.abs(q[5] - n)) <= 1e-9 * n
    assert np.max(np.abs(q[4] - TG)) <= 1e-9 * TG
    return abs(u[0, 0, 0] - exact) / UG


def drag_orders(lib, prefix, nrk, mode=0):
    errs = np.array([drag_run(lib, prefix, int(round(0.02 / dt)), dt, nrk, mode) for dt in (1e-4, 5e-5, 2.5e-5)])
    assert np.all(errs > 1e-12), errs
    return np.log(errs[:-1] / errs[1:]) / np.log(2.0)


@pytest.mark.parametrize("ratio", [0.5, 2.0])
def test_drag_relaxation_oracle(oracle, ratio):
    assert drag_run(oracle, "orc_", int(round(ratio * TAU_A / 1.0e-4))) <= 1e-5


def test_drag_time_order_oracle(oracle):
    assert drag_orders(oracle, "orc_", 2).min() >= 1.65
    assert drag_orders(oracle, "orc_", 3).min() >= 2.65


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_drag_relaxation_gpu(gpu, mode):
    assert drag_run(gpu, "ice_gpu_", int(round(TAU_A / 1.0e-4)), mode=mode) <= 1e-5
    assert drag_orders(gpu, "ice_gpu_", 2, mode).min() >= 1.65
    assert drag_orders(gpu, "ice_gpu_", 3, mode).min() >= 2.65


# G, a cloud released at rest into a uniform frozen gas on a periodic line
# (G-cloud-translation/run.py:1-36, 96-182): n proportional to rho_p keeps the diameter,
# hence tau_p, uniform, so u_p(t) = u_g (1 - exp(-t / tau_p)) everywhere and the cloud
# translates rigidly by X(t) = u_g t - u_g tau_p (1 - exp(-t / tau_p)).  To t = 3 tau_p:
# the mass-weighted u_p within 1e-6 u_g, mass to 1e-12, the centroid within 0.05 dx;
# a wrapped Gaussian converging at order >= 1.5 in space (dt = 1e-5 s); the spread of
# u_p vanishing at order >= 1.8 in dt, below 1e-6 u_g at 5e-6 s; a top hat's
# invariants.
XC, SIG, HALF, FLOOR, RHO_G0 = 0.25, 0.06, 0.10, 1.0e-6, 10.0


def wrapped_gaussian(x):
    return sum(np.exp(-(x - XC - m * L_C) ** 2 / (2.0 * SIG ** 2)) for m in (-1, 0, 1))


def top_hat(x):
    d = np.mod(x - XC + 0.5 * L_C, L_C) - 0.5 * L_C
    return np.where(np.abs(d) <= HALF, 1.0, 0.0)


def cloud_average(shape, xl, xr, shift, m=32):
    """the midpoint average over each cell of the profile displaced by shift"""
    h = (xr - xl) / m
    pts = xl[:, None] + (np.arange(m)[None, :] + 0.5) * h[:, None]
    return (RHO_G0 * (shape(np.mod(pts - shift, L_C)) + FLOOR)).mean(axis=1)


def cloud_run(lib, prefix, nx, shape, dt, mode=0):
    t_end = 3.0 * TAU_A
    steps = int(math.ceil(t_end / dt - 1e-9))

    def averages(xl, xr):
        flat = cloud_average(shape, xl.ravel(), xr.ravel(), 0.0)
        return flat.reshape(xl.shape)

    r = cases.Runner(lib, periodic_line(nx, averages, 0.0, 3, steps, dt, 0.4, UG, mode), prefix)
    st = r.step(steps)
    q = r.state(1, 1)
    rho, up = q[0][:, 0, 0], q[1][:, 0, 0]
    dx = L_C / nx
    x = np.arange(nx) * dx
    t = st.time
    assert abs(float((rho * up).sum() / rho.sum()) - UG * (1.0 - math.exp(-t / TAU_A))) <= 1e-6 * UG
    mass0 = cloud_average(shape, x, x + dx, 0.0).sum() * dx
    assert abs(rho.sum() * dx / mass0 - 1.0) <= 1e-12
    X = UG * t - UG * TAU_A * (1.0 - math.exp(-t / TAU_A))
    ang = (rho * np.exp(2j * math.pi * (x + 0.5 * dx) / L_C)).sum()
    drift = abs((math.atan2(ang.imag, ang.real) * L_C / (2 * math.pi)) % L_C - (XC + X) % L_C)
    assert min(drift, L_C - drift) <= 0.05 * dx, (nx, drift / dx)
    ref = cloud_average(shape, x, x + dx, X)
    return np.abs(rho - ref).sum() / ref.sum(), (up.max() - up.min()) / UG


def check_cloud(lib, prefix, mode=0):
    errs = np.array([cloud_run(lib, prefix, nx, wrapped_gaussian, 1.0e-5, mode)[0] for nx in (50, 100, 200)])
    assert (np.log(errs[:-1] / errs[1:]) / np.log(2.0)).min() >= 1.5, errs
    spreads = np.array([cloud_run(lib, prefix, 100, wrapped_gaussian, dt, mode)[1] for dt in (2.0e-5, 1.0e-5, 5.0e-6)])
    assert (np.log(spreads[:-1] / spreads[1:]) / np.log(2.0)).min() >= 1.8, spreads
    assert spreads[-1] <= 1e-6
    cloud_run(lib, prefix, 100, top_hat, 1.0e-5, mode)


def test_cloud_translation_oracle(oracle):
    check_cloud(oracle, "orc_")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_cloud_translation_gpu(gpu, mode):
    check_cloud(gpu, "ice_gpu_", mode)


# H, a cloud in a linearly straining frozen gas u_g = U0 + a x (H-linear-strain/run.py:
# 1-36, 86-176), seeded at the local gas velocity with n proportional to rho_p: with
# xi = x + U0/a every particle is displaced by the same factor C(t), C = A e^{l+ t} +
# B e^{l- t}, l± = (-1 ± sqrt(1 + 4 a tau)) / (2 tau), A + B = 1, A l+ + B l- = a, so
# rho_p(x, t) = rho_p0(xi / C - U0/a) / C and u_p = (C'/C) xi.  To t = 0.1 s at dt =
# 2e-5 s, extrapolation ends: the centroid within 0.05 dx, du_p/dxi within 2e-4 of
# C'/C, the cloud's mass within 2e-3 of the exact stretch, the profile converging at
# order >= 1.5 (nx = 100, 200, 400); then, at nx = 200, the slip coefficient a - C'/C
# over a^2 tau (the u_g - tau Du_g/Dt asymptote) rising monotonically to 1 as St =
# a tau falls (0.2 ... 0.025), 1 - ratio within [1.3, 2.2] St, and within 5e-2 of 1
# at St = 0.025.
U0_H, A_H, X0_H, SIG_H = 2.0, 6.0, 0.20, 0.035
T_END_H, DT_H = 0.10, 2.0e-5


def strain_profile(x):
    return RHO_G0 * (np.exp(-(x - X0_H) ** 2 / (2.0 * SIG_H ** 2)) + FLOOR)


def stretch(tau, t):
    root = math.sqrt(1.0 + 4.0 * A_H * tau)
    lp, lm = (-1.0 + root) / (2.0 * tau), (-1.0 - root) / (2.0 * tau)
    a = (A_H - lm) / (lp - lm)
    b = (lp - A_H) / (lp - lm)
    return a * math.exp(lp * t) + b * math.exp(lm * t), a * lp * math.exp(lp * t) + b * lm * math.exp(lm * t)


def strain_run(lib, prefix, nx, dp, mode=0):
    steps = int(math.ceil(T_END_H / DT_H - 1e-9))
    dx = L_C / nx
    gas = lambda x: U0_H + A_H * x  # noqa: E731
    case = periodic_line(nx, lambda xl, xr: strain_profile(0.5 * (xl + xr)), gas, 3, steps, DT_H, 0.4, gas, mode,
                         periodic=False, dp=dp)
    r = cases.Runner(lib, case, prefix)
    st = r.step(steps)
    q = r.state(1, 1)
    rho, up = q[0][:, 0, 0], q[1][:, 0, 0]
    xc = (np.arange(nx) + 0.5) * dx
    xi = xc + U0_H / A_H
    slope = (rho * up * xi).sum() / (rho * xi * xi).sum()
    return st.time, rho, slope, xc, dx


def check_strain(lib, prefix, mode=0):
    tau = RHO_AL * DP ** 2 / (18.0 * MU_A)
    errs = []
    for nx in (100, 200, 400):
        t, rho, slope, xc, dx = strain_run(lib, prefix, nx, DP, mode)
        C, Cd = stretch(tau, t)
        h = dx / 32
        pts = (xc - 0.5 * dx)[:, None] + (np.arange(32)[None, :] + 0.5) * h
        ref = (strain_profile((pts + U0_H / A_H) / C - U0_H / A_H) / C).mean(axis=1)
        errs.append(np.abs(rho - ref).sum() / ref.sum())
        xbar = (rho * xc).sum() / rho.sum()
        assert abs(xbar - (C * (X0_H + U0_H / A_H) - U0_H / A_H)) <= 0.05 * dx, nx
        assert abs(slope / (Cd / C) - 1.0) <= 2e-4, (nx, slope, Cd / C)
        assert abs(rho.sum() / ref.sum() - 1.0) <= 2e-3, nx
    errs = np.array(errs)
    assert (np.log(errs[:-1] / errs[1:]) / np.log(2.0)).min() >= 1.5, errs
    stokes = (0.20, 0.10, 0.05, 0.025)
    ratios = []
    for St in stokes:
        tau_s = St / A_H
        dp_s = math.sqrt(18.0 * MU_A * tau_s / RHO_AL)
        t, rho, slope, xc, dx = strain_run(lib, prefix, 200, dp_s, mode)
        C, Cd = stretch(tau_s, t)
        assert abs(slope / (Cd / C) - 1.0) <= 5e-4, (St, slope, Cd / C)
        ratios.append((A_H - slope) / (A_H ** 2 * tau_s))
    assert all(a < b for a, b in zip(ratios[:-1], ratios[1:])), ratios
    assert all(1.3 * s <= 1.0 - r <= 2.2 * s for r, s in zip(ratios, stokes)), ratios
    assert abs(ratios[-1] - 1.0) <= 0.05, ratios


def test_linear_strain_oracle(oracle):
    check_strain(oracle, "orc_")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_linear_strain_gpu(gpu, mode):
    check_strain(gpu, "ice_gpu_", mode)


# I, a cloud in a frozen gas in solid-body rotation u_g = -Omega y, v_g = Omega x
# (I-vortex-cloud/run.py:1-35, 97-200), seeded at the local gas velocity: in z = x + i y
# every particle is displaced by the same complex factor C(t), C = A e^{l+ t} + B e^{l- t},
# l± = (-1 ± sqrt(1 + 4 i Omega tau)) / (2 tau), A + B = 1, A l+ + B l- = i Omega, so
# rho_p(z, t) = rho_p0(z / C) / |C|^2 and u_p + i v_p = (C'/C) z.  Over a quarter turn on
# [-2, 2]^2 (extrapolation faces, dt-max 1e-3 s): at St = 0.1 the profile and the
# centroid converge at order >= 1.5 (n = 48, 96, 192), the centroid within 0.1 h; over
# St = 0.01 ... 10 at n = 96, |centroid| / x0 within 5e-3 of |C|, the lag within 5e-3 rad
# of Omega t - arg C, the complex velocity slope within 1e-2 of C'/C, L1 <= 0.15, the
# mass within 1e-3, the lag and the ejection growing with St, a tracer at St = 0.01 and
# a lag >= 0.4 rad at St = 10.
OMEGA, HALF_I, X0_I, SIG_I, FLOOR_I = 1.0, 2.0, 0.6, 0.12, 1.0e-8
QUARTER = 0.5 * math.pi / OMEGA


def vortex_stretch(tau, t):
    import cmath
    root = cmath.sqrt(1.0 + 4.0j * OMEGA * tau)
    lp, lm = (-1.0 + root) / (2.0 * tau), (-1.0 - root) / (2.0 * tau)
    a = (1.0j * OMEGA - lm) / (lp - lm)
    b = (lp - 1.0j * OMEGA) / (lp - lm)
    return a * cmath.exp(lp * t) + b * cmath.exp(lm * t), a * lp * cmath.exp(lp * t) + b * lm * cmath.exp(lm * t)


def vortex_profile(x, y):
    return RHO_G0 * (np.exp(-((x - X0_I) ** 2 + y ** 2) / (2.0 * SIG_I ** 2)) + FLOOR_I)


def vortex_case(n, tau, mode=0):
    """the square of n x n cells, the cloud and the gas at the cell centres"""
    from ice_amd import geometry
    h = 2.0 * HALF_I / n
    x = np.linspace(-HALF_I, HALF_I, n + 1)
    nodes = geometry.cartesian_nodes(x, x, [0.0, h])
    geo = geometry.metrics(nodes)
    dims = (n, n, 1)
    dp = math.sqrt(18.0 * MU_A * tau / RHO_AL)
    vol = RHO_AL * 4.0 / 3.0 * math.pi * (0.5 * dp) ** 3

    def fill(I, J, K):
        xc, yc = -HALF_I + (I - 0.5) * h, -HALF_I + (J - 0.5) * h
        rho = vortex_profile(xc, yc)
        z = np.zeros(I.shape)
        return np.stack([rho, -OMEGA * yc, OMEGA * xc, z, z + TG, rho / vol])

    prim = cases.padded_state(dims, 6, fill)
    xc = -HALF_I + (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    g5 = np.zeros((5,) + dims)
    g5[0], g5[1], g5[2], g5[4] = 1.2, -OMEGA * Y[:, :, None], OMEGA * X[:, :, None], TG
    gas = cases.gas_arrays(g5, 287.05, 1.4, KG, MU_A)
    tb = cases.TableBuilder(1)
    tb.add_block(1, dims, lambda f, i, j, k: np.where(f <= 4, 400, 0))
    reci, recd, donor, weight = tb.arrays()
    steps = int(math.ceil(QUARTER / 1.0e-3 - 1e-9))
    cfg = cases.make_config(ngroups=1, nrk=2, time_accurate=True, ndir=2, limiter=3, coupled=True, drag_select=2,
                            heat_select=1, cfl=0.4, dt_max=1.0e-3, mode=mode)
    fam = cases.make_family(MK, rho_mat=RHO_AL, cs=CS)
    case = cases.Case("vortex", cfg, [fam], [cases.Block(dims, geo, [prim], gas)], reci, recd, donor, weight, steps)
    return case, X + 1j * Y, h


def vortex_run(lib, prefix, n, tau, mode=0):
    """to a quarter turn (the last step passes it, as the reference's runs end at or
    after t_end); the cloud's mass, centroid, complex velocity slope and L1 error"""
    case, Z, h = vortex_case(n, tau, mode)
    r = cases.Runner(lib, case, prefix)
    t = r.step(case.steps).time
    assert QUARTER <= t < QUARTER + 1.0e-3, t  # dt = dt-max on every step
    q = r.state(1, 1)
    rho, w = q[0][:, :, 0], q[1][:, :, 0] + 1j * q[2][:, :, 0]
    C, Cd = vortex_stretch(tau, t)
    mass = rho.sum()
    centre = (rho * Z).sum() / mass
    slope = (rho * w * np.conj(Z)).sum() / (rho * np.abs(Z) ** 2).sum()
    z0 = Z / C
    ref = vortex_profile(z0.real, z0.imag) / abs(C) ** 2
    return t, C, Cd, mass, centre, slope, np.abs(rho - ref).sum() / ref.sum(), ref.sum(), h


def check_vortex(lib, prefix, grids=(48, 96, 192), n_sweep=96, mode=0):
    import cmath
    errs, drifts = [], []
    for n in grids:
        t, C, Cd, mass, centre, slope, err, mref, h = vortex_run(lib, prefix, n, 0.1 / OMEGA, mode)
        errs.append(err)
        drifts.append(abs(centre - C * X0_I))
        assert drifts[-1] <= 0.1 * h, (n, drifts[-1] / h)
    order = lambda e: np.log(np.array(e[:-1]) / np.array(e[1:])) / np.log(2.0)  # noqa: E731
    assert order(errs).min() >= 1.5, errs
    assert order(drifts).min() >= 1.5, drifts
    if n_sweep is None:
        return
    lags, radii = [], []
    for St in (0.01, 0.1, 1.0, 10.0):
        t, C, Cd, mass, centre, slope, err, mref, h = vortex_run(lib, prefix, n_sweep, St / OMEGA, mode)
        scale, lag = abs(centre) / X0_I, OMEGA * t - cmath.phase(centre)
        assert abs(scale / abs(C) - 1.0) <= 5e-3, (St, scale, abs(C))
        assert abs(lag - (OMEGA * t - cmath.phase(C))) <= 5e-3, (St, lag)
        assert abs(slope / (Cd / C) - 1.0) <= 1e-2, (St, slope, Cd / C)
        assert err <= 0.15, (St, err)
        assert abs(mass / mref - 1.0) <= 1e-3, (St, mass / mref)
        lags.append(lag)
        radii.append(scale)
    assert all(a < b for a, b in zip(lags[:-1], lags[1:])), lags
    assert all(a < b for a, b in zip(radii[:-1], radii[1:])), radii
    assert lags[0] <= 5e-3 and abs(radii[0] - 1.0) <= 0.02
    assert lags[-1] >= 0.4


def test_vortex_cloud_oracle(oracle):
    """the refinement leg on the two coarser grids (the full case takes the oracle minutes)"""
    check_vortex(oracle, "orc_", grids=(48, 96), n_sweep=None)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_vortex_cloud_gpu(gpu, mode):
    check_vortex(gpu, "ice_gpu_", mode=mode)


# L1, each family on its own material (L-materials/run.py:8-17, 68-72, 156-194): MK
# families with the same (rho_p, n) at rest and 300 K in a gas at rest at 400 K, Stokes
# heat (Nu = 2), no drag.  Each family recovers Rp from its own material's density and
# relaxes with tau = rho_p cp / (4 pi k Rp n); materials A (cp 1000, rho 2000) and B
# (cp 2000, rho 1000): tau_B / tau_A = 2 (1/2)^(1/3).  To t = tau_A at dt = tau_min / 1000,
# every family within 1e-6 of the gap (one heating as the other material misses by 0.165).
MAT_A, MAT_B = (1000.0, 2000.0), (2000.0, 1000.0)   # (cp, rho)


def materials_case(mats, mode=0):
    from ice_amd import geometry
    nf = len(mats)
    nx = 8
    dims = (nx, 1, 1)
    nodes = geometry.cartesian_nodes(np.linspace(0.0, 1.0, nx + 1), [0.0, 1.0 / nx], [0.0, 1.0 / nx])
    geo = geometry.metrics(nodes)
    rho_p = ALPHA * 2000.0
    n = rho_p / (2000.0 * 4.0 / 3.0 * math.pi * (0.5 * DP) ** 3)
    taus = []
    for cp, rho_m in mats:
        Rp = (0.75 * rho_p / (n * math.pi * rho_m)) ** (1.0 / 3.0)
        taus.append(rho_p * cp / (4.0 * math.pi * KG * Rp * n))
    dt = min(taus) / 1000.0
    steps = int(math.ceil(taus[0] / dt - 1e-9))

    def fill(I, J, K):
        z = np.zeros(I.shape)
        return np.stack([z + rho_p, z, z, z, z + TG, z + n])

    prims = [cases.padded_state(dims, 6, fill) for _ in range(nf)]
    g5 = np.zeros((5,) + dims)
    g5[0], g5[4] = 1.2, 400.0
    gas = cases.gas_arrays(g5, 287.05, 1.4, KG, MU_A)
    tb = cases.TableBuilder(nf)
    tb.add_block(1, dims, lambda f, i, j, k: np.where(f <= 4, 400, 0))
    reci, recd, donor, weight = tb.arrays()
    cfg = cases.make_config(ngroups=nf, nrk=2, time_accurate=True, ndir=1, limiter=3, coupled=True, drag_select=13,
                            heat_select=1, cfl=0.8, dt_max=dt, mode=mode)
    fams = [cases.make_family(MK, rho_mat=rho_m, cs=cp) for cp, rho_m in mats]
    case = cases.Case("materials", cfg, fams, [cases.Block(dims, geo, prims, gas)], reci, recd, donor, weight, steps)
    return case, taus


def check_materials(lib, prefix, mats, mode=0):
    case, taus = materials_case(mats, mode)
    r = cases.Runner(lib, case, prefix)
    t = r.step(case.steps).time
    for p in range(len(mats)):
        T = r.state(1, p + 1)[4]
        ref = 400.0 + (TG - 400.0) * math.exp(-t / taus[p])
        assert np.max(np.abs(T - ref)) <= 1e-6 * (400.0 - TG), (p, float(T[0, 0, 0]), ref)
    other = 400.0 + (TG - 400.0) * math.exp(-t / taus[-1])
    assert abs(r.state(1, 1)[4][0, 0, 0] - other) >= 0.1 * (400.0 - TG) or mats[0] == mats[-1]


@pytest.mark.parametrize("mats", [(MAT_A, MAT_B), (MAT_A, MAT_A, MAT_B)], ids=["A1-B1", "A2-B1"])
def test_materials_oracle(oracle, mats):
    check_materials(oracle, "orc_", mats)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("mats", [(MAT_A, MAT_B), (MAT_A, MAT_A, MAT_B)], ids=["A1-B1", "A2-B1"])
def test_materials_gpu(gpu, mats, mode):
    check_materials(gpu, "ice_gpu_", mats, mode)
