"""Closed-form verification cases of ICE (test/verification/*/run.py), run on the
oracle and on the device from the same inputs.

* B, thermal relaxation (B-thermal-relaxation/run.py:1-12): a cloud moving with
  the gas (no slip, so no drag), hotter than the gas, Stokes heat (Nu = 2),
  no radiation.  T_p(t) = T_g + (T_p0 - T_g) exp(-t / tau_T), tau_T =
  rho_al cs dp^2 / (12 k_g); velocity, density and number density do not move.
  The case's bar: 1e-5 of the initial gap (RK2 at dt = 1e-4 s, tau_T = 0.029 s).
* J, the time step bounded by the relaxation time (J-tau-limit/run.py:1-25): the
  first step is dt-max, the later ones tau_p, and the slip follows the SSP-RK3
  amplification of each step (below)."""
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
