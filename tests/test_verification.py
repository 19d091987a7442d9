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
* A, Stokes drag relaxation, C, periodic sine advection, and G, a cloud carried
  by a frozen gas on a periodic line: the closed forms and observed orders of
  those cases (below, each with its bars)."""
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


def periodic_line(nx, averages, u0, limiter=3, steps=1, dt_max=1.0e-4, cfl=0.8, gas_u=None, mode=0):
    """A line of nx cells on [0, L_C] whose faces 1 and 2 are joined (201, periodic), the
    others extrapolation (400) and type 0: MK particles of dp = 1e-4 m at the cell
    averages `averages(xl, xr)` of density, moving at u0, at T_g; with `gas_u`, a
    frozen gas (1.2 kg/m3, 300 K) at that speed drags them (Stokes)."""
    from ice_amd import geometry, native
    dims = (nx, 1, 1)
    dx = L_C / nx
    nodes = geometry.cartesian_nodes(np.linspace(0.0, L_C, nx + 1), [0.0, dx], [0.0, dx])
    geo = geometry.metrics(nodes)
    vol = RHO_AL * 4.0 / 3.0 * math.pi * (0.5 * DP) ** 3

    def fill(I, J, K):
        rho = averages((I - 1) * dx, I * dx)
        z = np.zeros(I.shape)
        return np.stack([rho, z + u0, z, z, z + TG, rho / vol])

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

    tb.add_block(1, dims, lambda f, i, j, k: np.where(f <= 2, 201, np.where(f <= 4, 400, 0)), connect)
    reci, recd, donor, weight = tb.arrays()
    gas = None
    if gas_u is not None:
        g5 = np.zeros((5,) + dims)
        g5[0], g5[1], g5[4] = 1.2, gas_u, TG
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
