"""Known answers of the grey-body radiation term (Mod_Sources.f90:559), after the
reference's own unit test test/unit/test_radiation.f90: a cloud of MK drops of
d = 30 um at rest at T_p = 2400 K in a gas at rest at T_g = 600 K, NoDrag, NoHeat,
no evaporation, so radiation is the only energy term.

  R1 emissivity 0.8: the energy slot of the source is 0.8 sigma pi d^2 n (T_g^4 - T_p^4),
     to 1e-14 relative (the solver's radius is a cube root of the state)
  R2 emissivity 0: the energy slot is 0 and every other slot is R1's
  R3 a material whose emissivity is left at its default (0) radiates nothing
  R4 an emissivity column T/3000 (a property table) radiates with its value at T_p, 0.8

The oracle is held to the closed form on the CPU; the device to the same closed
form and to the oracle, on the plain and the batched path."""
import numpy as np
import pytest

from ice_amd import cases, geometry, native, properties

SIGMA = 5.670374419e-8
PI = np.pi
D, RHO_P, TP, TG = 30e-6, 1.0, 2400.0, 600.0
RG, GAM, KG, MU = 287.05, 1.4, 0.05, 3e-5
RHO_L, CP_L = 2000.0, 1000.0


def radiation_case(emiss: float, table: bool = False, mode: int = 0, n: int = 4):
    """n^3 cells of the test's cloud in the test's gas, 400 faces, one Euler step
    (the source of its only stage is the source at the initial state)."""
    dims = (n, n, n)
    geo = geometry.uniform_metrics(dims, 1e-3)
    num = RHO_P / (RHO_L * PI / 6.0 * D ** 3)

    def fill(I, J, K):
        z = np.zeros(I.shape)
        return np.stack([z + RHO_P, z, z, z, z + TP, z + num])

    prim = cases.padded_state(dims, 6, fill)
    g5 = np.zeros((5,) + dims)
    g5[0] = 101325.0 / (RG * TG)
    g5[4] = TG
    gas = cases.gas_arrays(g5, RG, GAM, KG, MU)
    tb = cases.TableBuilder(1)
    tb.add_block(1, dims, lambda f, i, j, k: np.full(len(f), 400, np.int32))
    reci, recd, donor, weight = tb.arrays()
    cfg = cases.make_config(ngroups=1, nrk=1, time_accurate=True, ndir=3, limiter=3, coupled=True, drag_select=13,
                            heat_select=8, cfl=0.5, dt_max=1e-6, mode=mode)
    fam = cases.make_family(native.MK, native.SAUREL, rho_mat=RHO_L, cs=CP_L, lv=2.26e6, emiss=emiss)
    meta = {}
    if table:
        T = np.arange(1, 3001, dtype=float)
        meta["tables"] = {1: properties.material_table(T, np.full(T.size, CP_L), np.full(T.size, RHO_L), CP_L * T,
                                                       emiss=T / 3000.0)}
    return cases.Case("radiation", cfg, [fam], [cases.Block(dims, geo, [prim], gas)], reci, recd, donor, weight, 1,
                      meta=meta)


def closed_form(eps: float) -> float:
    num = RHO_P / (RHO_L * PI / 6.0 * D ** 3)
    return eps * SIGMA * PI * D * D * num * (TG * TG * TG * TG - TP * TP * TP * TP)


def sources(lib, prefix, case):
    r = cases.Runner(lib, case, prefix)
    assert getattr(lib, prefix + "set_keep_source")(r.ctx, 1) == 0
    r.step(1)
    s, tau = r.source(1, 1)
    r.close()
    return s, tau


def legs(lib, prefix, mode=0):
    s8, t8 = sources(lib, prefix, radiation_case(0.8, mode=mode))
    s0, t0 = sources(lib, prefix, radiation_case(0.0, mode=mode))
    st, tt = sources(lib, prefix, radiation_case(0.0, table=True, mode=mode))
    return (s8, t8), (s0, t0), (st, tt)


def check_legs(s8, t8, s0, t0, st):
    q = closed_form(0.8)
    # R1
    rel = np.abs(s8[4] - q) / abs(q)
    assert rel.max() <= 1e-14, rel.max()
    # R2 (and R3: the default emissivity is 0)
    assert np.all(s0[4] == 0.0)
    assert np.array_equal(s0[:4], s8[:4]) and np.array_equal(s0[5], s8[5]) and np.array_equal(t0, t8)
    # R4
    rel = np.abs(st[4] - q) / abs(q)
    assert rel.max() <= 1e-14, rel.max()
    # no drag: the relaxation time is infinite
    assert np.all(np.isinf(t8))


def test_radiation_oracle(oracle):
    (s8, t8), (s0, t0), (st, tt) = legs(oracle, "orc_")
    check_legs(s8, t8, s0, t0, st)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_radiation_device(gpu, oracle, mode):
    (s8, t8), (s0, t0), (st, tt) = legs(gpu, "ice_gpu_", mode)
    check_legs(s8, t8, s0, t0, st)
    (o8, _), (o0, _), (ot, _) = legs(oracle, "orc_", mode)
    for a, b in ((s8, o8), (s0, o0), (st, ot)):
        assert np.allclose(a, b, rtol=1e-14, atol=0.0)
