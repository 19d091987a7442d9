"""The oracle's cell source with evaporation, pinned by what the mathematics
fixes (no GPU).

* d2-law drops frozen at their temperature in still air: d^2 = d0^2 - K t with
  K = 8 k ln(1 + B_T) / (rho_l cp), so the bulk density follows
  rho_p(t) = n rho_l pi/6 (d0^2 - K t)^(3/2) (Godsave/Spalding; the setting of
  test/verification/F-evaporation/run.py, part 1).  SSP-RK3 integrates it to
  its truncation error; a wrong sign, slot or radius formula fails by orders.
* A far-field vapour fraction given per cell, equal everywhere to ep(iYinf),
  gives the same bits as the material constant (the per-cell form replaces
  only that value)."""
import ctypes
import math
import os

import numpy as np
import pytest

from ice_amd import cases, geometry, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def orc():
    return native.bind(ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so")), "orc_")


def one_cell_case(evap, intf=0, tp=300.0, d0=1e-4, yinf=0.0, var_yinf=False, heat=8, drag=13, blow=0,
                  dt=1e-4, steps=1, cs=1e12, emiss=0.0, tg=800.0, alphae=1.0):
    dims = (1, 1, 1)
    geo = geometry.metrics(geometry.cartesian_nodes(*[np.array([0.0, 1e-3])] * 3))
    R, gam, mu, kg = 287.05, 1.4, 1.8e-5, 0.026
    g5 = np.array([101325.0 / (R * tg), 0.0, 0.0, 0.0, tg]).reshape(5, 1, 1, 1)
    gas = cases.gas_arrays(g5, R, gam, kg, mu)
    rho_l = 1000.0
    alpha = 1e-3
    n = alpha / (math.pi / 6 * d0 ** 3)
    prim = cases.padded_state(dims, 6, lambda I, J, K: np.array([alpha * rho_l, 0, 0, 0, tp, n]).reshape(6, 1, 1, 1))
    blk = cases.Block(dims, geo, [prim], gas, yinf=np.full(1, yinf) if var_yinf else None)
    tb = cases.TableBuilder(1)
    tb.add_block(1, dims, lambda f, i, j, k: np.full(len(f), 400, np.int32))
    reci, recd, donor, weight = tb.arrays()
    cfg = cases.make_config(ngroups=1, nrk=3, time_accurate=True, ndir=3, limiter=1, coupled=True,
                            drag_select=drag, heat_select=heat, blow_select=blow, var_yinf=var_yinf, cfl=0.5,
                            dt_max=dt, tau_factor=0.0)
    ep = cases.water_ep(yinf=0.0 if var_yinf else yinf, alphae=alphae)
    fam = cases.make_family(native.MK, native.SAUREL, rho_mat=rho_l, cs=cs, lv=2.26e6, emiss=emiss, evap=evap,
                            intf=intf, ep=ep)
    return cases.Case("cell", cfg, [fam], [blk], reci, recd, donor, weight, steps), n


def test_d2_law_closed_form(orc):
    dt, steps = 2e-4, 200
    case, n = one_cell_case(evap=1, dt=dt)
    r = cases.Runner(orc, case, "orc_")
    r.step(steps)
    rho = r.state(1, 1)[0, 0, 0, 0]
    tg, tp, R, gam, kg, lv, rho_l = 800.0, 300.0, 287.05, 1.4, 0.026, 2.26e6, 1000.0
    cp = gam * R / (gam - 1)
    K = 8 * kg * math.log(1 + cp * (tg - tp) / lv) / (rho_l * cp)
    d2 = 1e-8 - K * dt * steps
    exact = n * rho_l * math.pi / 6 * d2 ** 1.5
    assert d2 > 0.3e-8  # a sizeable part of the drop went
    assert abs(rho - exact) / exact < 1e-9, (rho, exact)


@pytest.mark.parametrize("evap,intf", [(2, 0), (2, 1), (4, 1), (5, 0)])
def test_per_cell_yinf_is_the_constant(orc, evap, intf):
    a, _ = one_cell_case(evap=evap, intf=intf, tp=345.0, yinf=0.05, heat=6, steps=3, alphae=0.04, tg=1200.0)
    b, _ = one_cell_case(evap=evap, intf=intf, tp=345.0, yinf=0.05, var_yinf=True, heat=6, steps=3, alphae=0.04,
                         tg=1200.0)
    ra, rb = cases.Runner(orc, a, "orc_"), cases.Runner(orc, b, "orc_")
    ra.step(3)
    rb.step(3)
    sa, sb = ra.state(1, 1), rb.state(1, 1)
    assert np.array_equal(sa, sb)
    assert sa[0, 0, 0, 0] < 1.0  # the drops lost mass
