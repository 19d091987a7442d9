"""Evaporating drops against published closed forms: parts 1, 3, 4 and 5 of ICE's
verification case F (test/verification/F-evaporation/run.py:1-40, 392-530), on the
oracle and on the device.

Water drops (d0 = 1e-4 m, 1000 kg/m3, alpha = 1e-3) at rest at 300 K in air at rest
at 800 K and one atmosphere: Re = 0, no gradients, every cell alike.  The diameter
comes back from the state as d = 2 (0.75 rho_p / (n pi rho_l))^(1/3).

Part 1, the temperature frozen (cs = 1e12): mdot is proportional to d for these
models, so d^2 = d0^2 - K t with, from the Clausius-Clapeyron surface fraction
X_s = psat / p, Y_s = X_s Mv / (X_s Mv + (1 - X_s) M_gas) and B_M = Y_s / (1 - Y_s),
  d2-law  K = 8 k ln(1 + cp (T_g - T_p) / Lv) / (rho_l cp)    (Godsave, Spalding)
  CEM     K = 8 k ln(1 + B_M) / (cp Le rho_l)                  (Spalding's mass transfer)
  CEM-B   the same on a film at T_f = T_p + (T_g - T_p)/3 (Hubbard's 1/3 rule), with
          k_f = k (T_f / T_g)^0.7
  ASM     CEM's K: at Re = 0, Sh* = Sh0 = 2 (Abramzon-Sirignano).
Bars: K within 1e-6, rho_p uniform to 1e-12, n unchanged to 1e-9, T_p within 1e-3 K;
evaporation none leaves rho_p to 1e-14.

Part 3, nothing frozen (cs = 4180): with the d2-law, Nu = 2 and the Miller-Harstad-
Bellan blowing factor, b = ln(1 + B_T) identically, so the blown convective heat
equals the latent sink term by term and the drop holds 300 K (to 1e-8 K) while its
d^2 falls at the d2-law's K (1e-6); without the blowing factor it heats by over 1 K.

Part 4, the same slopes at 380 K, above the boiling point: psat > p puts X_s on the
cap 1 - 1e-12 (B_M about 6e11), to 1e-5 (one ulp of Y_s at the cap moves the rate
by 2.5e-6), over the time CEM takes to lose 10 % of the mass.

Part 5, psat from a property table's column instead of the curve, the drop held on
the 300 K row (cp = 1e16): CEM's slope with the column's value, 1.2 times the curve
(a rate 1.207 times higher) and the curve itself (the control), to 1e-6.

And ICE's case L2: two families side by side, one evaporating (CEM) and one not, each
on its own model: the first follows CEM's slope, the second keeps its state bit for bit.

The TC model's slope needs its film equation solved and is left to the bit-for-bit
tests against the reference's Lib_Evaporation."""
import math

import numpy as np
import pytest

from ice_amd import cases, geometry, native

RU, PATM = 8314.46, 101325.0
TG, RGAS, GAM, MU, KG = 800.0, 287.05, 1.4, 1.8e-5, 0.026
RHO_G = PATM / (RGAS * TG)
CPG = GAM * RGAS / (GAM - 1.0)
MGAS = RU / RGAS
RHO_L, CS_L, LV, MV, TBOIL, LE = 1000.0, 4180.0, 2.26e6, 18.015, 373.15, 1.0
D0, TP0, ALPHA = 1.0e-4, 300.0, 1.0e-3
RHO_P0 = ALPHA * RHO_L
NDENS = RHO_P0 / (RHO_L * 4.0 / 3.0 * math.pi * (0.5 * D0) ** 3)
MODELS = {"d2-law": 1, "CEM": 2, "CEM-B": 3, "ASM": 4}


def psat_cc(Tp):
    return PATM * math.exp(-(LV * MV / RU) * (1.0 / Tp - 1.0 / TBOIL))


def slope(model, Tp=TP0, factor=1.0):
    psat = factor * psat_cc(Tp)
    Xs = min(psat / (RHO_G * RGAS * TG), 1.0 - 1.0e-12)
    Ys = Xs * MV / (Xs * MV + (1.0 - Xs) * MGAS)
    BM = Ys / (1.0 - Ys)
    if model == "d2-law":
        return 8.0 * KG * math.log(1.0 + CPG * (TG - Tp) / LV) / (RHO_L * CPG)
    if model in ("CEM", "ASM"):
        return 8.0 * KG * math.log(1.0 + BM) / (CPG * LE * RHO_L)
    if model == "CEM-B":
        Tf = Tp + (TG - Tp) / 3.0
        return 8.0 * KG * (Tf / TG) ** 0.7 * math.log(1.0 + BM) / (CPG * LE * RHO_L)
    raise ValueError(model)


def diameter(rho_p):
    return 2.0 * (0.75 * rho_p / (NDENS * math.pi * RHO_L)) ** (1.0 / 3.0)


def drop_case(evap, cs, t_end, dt, blow=0, mode=0, tp0=TP0, table=None):
    """a line of 8 cells (extrapolation faces) holding the drops, RK2 at dt, Stokes drag and heat"""
    nx = 8
    dims = (nx, 1, 1)
    geo = geometry.metrics(geometry.cartesian_nodes(np.linspace(0.0, 1.0, nx + 1), [0.0, 0.125], [0.0, 0.125]))
    prim = cases.padded_state(dims, 6, lambda I, J, K: np.stack(
        [np.full(I.shape, RHO_P0), 0 * I, 0 * I, 0 * I, np.full(I.shape, tp0), np.full(I.shape, NDENS)]))
    g5 = np.zeros((5,) + dims)
    g5[0], g5[4] = RHO_G, TG
    gas = cases.gas_arrays(g5, RGAS, GAM, KG, MU)
    tb = cases.TableBuilder(1)
    tb.add_block(1, dims, lambda f, i, j, k: np.where(f <= 4, 400, 0))
    reci, recd, donor, weight = tb.arrays()
    steps = int(math.ceil(t_end / dt - 1e-9))
    cfg = cases.make_config(ngroups=1, nrk=2, time_accurate=True, ndir=1, limiter=3, coupled=True, drag_select=2,
                            heat_select=1, blow_select=blow, cfl=0.8, dt_max=dt, mode=mode)
    fam = cases.make_family(native.MK, rho_mat=RHO_L, cs=cs, lv=LV, evap=evap, intf=0,
                            ep=cases.water_ep(yinf=0.0, alphae=1.0))
    return cases.Case("drops", cfg, [fam], [cases.Block(dims, geo, [prim], gas)], reci, recd, donor, weight, steps,
                      meta={"tables": {1: table}} if table is not None else {})


def run(lib, prefix, evap, cs, t_end, dt, blow=0, mode=0, tp0=TP0, table=None):
    case = drop_case(evap, cs, t_end, dt, blow, mode, tp0, table)
    r = cases.Runner(lib, case, prefix)
    t = r.step(case.steps).time
    q = r.state(1, 1)
    assert np.max(q[0]) - np.min(q[0]) <= 1e-12 * RHO_P0
    assert np.max(np.abs(q[5] - NDENS)) <= 1e-9 * NDENS
    return t, q


def frozen_slopes(lib, prefix, mode=0):
    bad = []
    for model, evap in MODELS.items():
        t, q = run(lib, prefix, evap, 1.0e12, 0.04, 2.0e-5, mode=mode)
        K = (D0 ** 2 - diameter(q[0][0, 0, 0]) ** 2) / t
        err = abs(K - slope(model)) / slope(model)
        if err > 1e-6 or np.max(np.abs(q[4] - TP0)) > 1e-3:
            bad.append((model, K, slope(model), err))
    assert not bad, bad
    t, q = run(lib, prefix, 0, 1.0e12, 0.04, 2.0e-5, mode=mode)
    assert np.max(np.abs(q[0] - RHO_P0)) <= 1e-14 * RHO_P0


def boiling_slopes(lib, prefix, mode=0):
    """part 4: at 380 K psat exceeds p, so X_s sits on the cap 1 - 1e-12; the run lasts
    until CEM has lost 10 % of its mass, at 1000 steps"""
    tb = 380.0
    t_end = (1.0 - 0.9 ** (2.0 / 3.0)) * D0 ** 2 / slope("CEM", tb)
    bad = []
    for model, evap in MODELS.items():
        t, q = run(lib, prefix, evap, 1.0e12, t_end, t_end / 1000.0, mode=mode, tp0=tb)
        K = (D0 ** 2 - diameter(q[0][0, 0, 0]) ** 2) / t
        err = abs(K - slope(model, tb)) / slope(model, tb)
        if err > 1e-5:
            bad.append((model, K, slope(model, tb), err))
    assert not bad, bad


def psat_column(lib, prefix, mode=0):
    """part 5: psat from the table's column (250..450 K) in place of the curve, the drop
    held on the 300 K row by cp = 1e16; CEM's slope follows the column"""
    from ice_amd import properties
    T = np.arange(250.0, 451.0)
    cp = np.full(T.size, 1.0e16)
    Ks = []
    for factor in (1.2, 1.0):
        tab = properties.material_table(T, cp, np.full(T.size, RHO_L), cp * T,
                                        psat=factor * np.array([psat_cc(x) for x in T]), evaporates=True)
        t, q = run(lib, prefix, MODELS["CEM"], 1.0e16, 0.04, 2.0e-5, mode=mode, table=tab)
        K = (D0 ** 2 - diameter(q[0][0, 0, 0]) ** 2) / t
        assert abs(K - slope("CEM", factor=factor)) <= 1e-6 * slope("CEM", factor=factor), (factor, K)
        Ks.append(K)
    assert abs(Ks[0] / Ks[1] - slope("CEM", factor=1.2) / slope("CEM")) <= 1e-6
    assert Ks[0] / Ks[1] > 1.2


def per_family_models(lib, prefix, mode=0):
    """ICE's case L2 (L-materials/run.py:18-24): two families side by side, A with CEM and B
    with no evaporation: A follows CEM's slope, B keeps rho_p and n to the last bit"""
    case = drop_case(MODELS["CEM"], 1.0e12, 0.04, 2.0e-5, mode=mode)
    fam_b = cases.make_family(native.MK, rho_mat=RHO_L, cs=1.0e12, lv=LV, evap=0, intf=0,
                              ep=cases.water_ep(yinf=0.0, alphae=1.0))
    case.fams.append(fam_b)
    case.cfg.ngroups = 2
    B = case.blocks[0]
    B.prim = [B.prim[0], B.prim[0].copy()]
    tb = cases.TableBuilder(2)
    tb.add_block(1, B.dims, lambda f, i, j, k: np.where(f <= 4, 400, 0))
    case.reci, case.recd, case.donor, case.weight = tb.arrays()
    r = cases.Runner(lib, case, prefix)
    t = r.step(case.steps).time
    qa, qb = r.state(1, 1), r.state(1, 2)
    K = (D0 ** 2 - diameter(qa[0][0, 0, 0]) ** 2) / t
    assert abs(K - slope("CEM")) <= 1e-6 * slope("CEM"), K
    assert np.all(qb[0] == RHO_P0) and np.all(qb[5] == NDENS)


def isothermal(lib, prefix, mode=0):
    t, q = run(lib, prefix, 1, CS_L, 0.05, 5.0e-6, blow=1, mode=mode)
    assert np.max(np.abs(q[4] - TP0)) <= 1e-8, np.max(np.abs(q[4] - TP0))
    K = (D0 ** 2 - diameter(q[0][0, 0, 0]) ** 2) / t
    assert abs(K - slope("d2-law")) <= 1e-6 * slope("d2-law"), (K, slope("d2-law"))
    t, q = run(lib, prefix, 1, CS_L, 0.05, 5.0e-6, blow=0, mode=mode)
    assert q[4][0, 0, 0] - TP0 > 1.0


def test_slopes_published():
    """the slopes themselves: ASM equals CEM at Re = 0, the film lowers CEM-B's below CEM's
    (k_f < k at T_f < T_g), all positive at 300 K"""
    assert slope("ASM") == slope("CEM")
    assert 0.0 < slope("CEM-B") < slope("CEM")
    assert slope("d2-law") > 0.0


def test_frozen_slopes_oracle(oracle):
    frozen_slopes(oracle, "orc_")


def test_boiling_slopes_oracle(oracle):
    boiling_slopes(oracle, "orc_")


def test_psat_column_oracle(oracle):
    psat_column(oracle, "orc_")


def test_per_family_models_oracle(oracle):
    per_family_models(oracle, "orc_")


def test_isothermal_blowing_oracle(oracle):
    isothermal(oracle, "orc_")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_frozen_slopes_gpu(gpu, mode):
    frozen_slopes(gpu, "ice_gpu_", mode)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_boiling_slopes_gpu(gpu, mode):
    boiling_slopes(gpu, "ice_gpu_", mode)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_psat_column_gpu(gpu, mode):
    psat_column(gpu, "ice_gpu_", mode)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_per_family_models_gpu(gpu, mode):
    per_family_models(gpu, "ice_gpu_", mode)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_isothermal_blowing_gpu(gpu, mode):
    isothermal(gpu, "ice_gpu_", mode)
