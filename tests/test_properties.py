"""Tabulated material properties (Lib_Properties.f90:18-250, loaded by
Load_Table.f90): density, cp and its energy e(T) with the bisection inverse,
Psat, Lv, Cpv and emissivity, and every place the path takes them (closures,
cell source, evaporation, inlet ghosts).

Pinned three ways:
* bit for bit against the reference's own Lib_Properties, closures and
  evaporation (flang-built, oracle/_ref), on random tables and inputs;
* against the closed forms of ICE's verification case K-properties-table
  (test/verification/K-properties-table/run.py:1-60): the density read at
  300.4 K between a 2000 and a 1000 row is 1600 and sets the drag time of a cloud
  (case A's exponential, here its RK3 polynomial); below the table it saturates;
  a uniform cloud with a varying cp keeps its temperature to round-off over 200
  RK2 steps (each step a T -> e -> T round trip);
* the GPU against the oracle, on those cases and on an evaporation case whose
  every property comes from the table."""
import ctypes as C
import math

import numpy as np
import pytest

from ice_amd import cases, native, properties

from helpers import compare_states, run_both
from test_oracle_ref import NC, orc, random_prim, ref, same, unit  # noqa: F401  (fixtures)

D = C.c_double
P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
EPS = float(np.float32(1e-25))


def bind_tables(ref, orc):
    ref.ref_setup_table.argtypes = [C.c_int, C.c_int, P, C.c_int]
    ref.ref_mat.argtypes = [C.c_int, D]
    ref.ref_mat.restype = D
    for n in ("ref_mat_T_from_e",):
        getattr(ref, n).argtypes = [D]
        getattr(ref, n).restype = D
    ref.ref_mat_T_from_rhoe.argtypes = [D, D, D]
    ref.ref_mat_T_from_rhoe.restype = D
    ref.ref_evaporation_psat.argtypes = [D] * 9 + [C.c_int, C.c_int, P, D, C.POINTER(D), C.POINTER(D),
                                                   C.POINTER(C.c_int)]
    orc.orc_fn_setup_table.argtypes = [C.c_int, C.c_int, P, C.c_int]
    orc.orc_fn_table_value.argtypes = [C.c_int, D]
    for n in ("orc_fn_mat_e", "orc_fn_mat_T_from_e", "orc_fn_mat_rho", "orc_fn_table_value", "orc_fn_mat_T_from_rhoe"):
        getattr(orc, n).restype = D
    orc.orc_fn_mat_e.argtypes = [D]
    orc.orc_fn_mat_T_from_e.argtypes = [D]
    orc.orc_fn_mat_rho.argtypes = [D]
    orc.orc_fn_mat_T_from_rhoe.argtypes = [D, D, D]
    orc.orc_fn_evaporation_psat.argtypes = [D] * 9 + [C.c_int, C.c_int, P, D, C.POINTER(D), C.POINTER(D),
                                                      C.POINTER(C.c_int)]


def random_table(rng, t_min, n, cs_varies=True):
    """A table that passes Load_Table's checks: cp linear in T with its exact
    enthalpy, every optional column positive and varying."""
    T = np.arange(t_min, t_min + n, dtype=float)
    a, b = rng.uniform(500.0, 4000.0), (rng.uniform(0.05, 2.0) if cs_varies else 0.0)
    cp = a + b * T
    h = a * T + 0.5 * b * T * T
    if not cs_varies:
        h = cp * T
    rho = rng.uniform(800.0, 3000.0) + rng.uniform(-1.0, 1.0) * (T - T[0])
    psat = 1e3 * np.exp(0.02 * (T - T[0]))
    return properties.material_table(T, cp, rho, h, psat=psat, lv=2e6 - 100.0 * (T - T[0]),
                                      cpv=1500.0 + 0.5 * T, emiss=0.5 + 0.4 * (T - T[0]) / n, evaporates=True,
                                      relative=cs_varies)


def set_both(ref, orc, tab):
    ref.ref_setup_table(tab.t_min, tab.n, tab.cols, tab.flags)
    orc.orc_fn_setup_table(tab.t_min, tab.n, tab.cols, tab.flags)


@pytest.mark.parametrize("t_min,n", [(0, 40), (1, 300), (280, 121)])
def test_table_functions_bit_for_bit(ref, orc, t_min, n):
    bind_tables(ref, orc)
    rng = np.random.default_rng(7 + t_min)
    ref.ref_setup(1, 1000.0, 1000.0, 2e6, 0.5)
    orc.orc_fn_setup(1, 1000.0, 1000.0, 2e6, 0.5)
    tab = random_table(rng, t_min, n)
    set_both(ref, orc, tab)
    hi = t_min + n - 1
    Ts = np.concatenate([rng.uniform(t_min - 30.0, hi + 30.0, 400), np.arange(t_min, hi + 1, dtype=float),
                         [float(t_min), float(hi), t_min - 1e-9, hi + 1e-9, np.nan]])
    for T in Ts:
        assert same(ref.ref_mat(1, T), orc.orc_fn_mat_rho(T))
        assert same(ref.ref_mat(3, T), orc.orc_fn_mat_e(T))
        for col in (properties.PSAT, properties.LV, properties.CPV, properties.EMISS):
            assert same(ref.ref_mat(col + 1, T), orc.orc_fn_table_value(col, T)), (col, T)
    es = [orc.orc_fn_mat_e(T) for T in Ts] + list(rng.uniform(-1e5, 2e7, 200))
    for e in es:
        assert same(ref.ref_mat_T_from_e(e), orc.orc_fn_mat_T_from_e(e)), e
        rho = 10 ** rng.uniform(-9, 1)
        assert same(ref.ref_mat_T_from_rhoe(e * rho, rho, EPS), orc.orc_fn_mat_T_from_rhoe(e * rho, rho, EPS))


@pytest.mark.parametrize("model", [0, 1, 2])
def test_closures_with_table_bit_for_bit(ref, orc, model):
    """cs_varies: the closures' energy is rho e(T), its inverse the bisection"""
    bind_tables(ref, orc)
    rng = np.random.default_rng(300 + model)
    nc = NC[model]
    ref.ref_setup(model, 1000.0, 4186.0, 2.26e6, 0.9)
    orc.orc_fn_setup(model, 1000.0, 4186.0, 2.26e6, 0.9)
    set_both(ref, orc, random_table(rng, 200, 300))
    for t in range(300):
        q = random_prim(rng, model, empty=(t % 7 == 0))
        n = unit(rng)
        a, b = np.zeros(nc), np.zeros(nc)
        ref.ref_p2c(nc, q, a)
        orc.orc_fn_p2c(q, b)
        assert same(a, b)
        ref.ref_c2p(nc, a.copy(), a)
        orc.orc_fn_c2p(b.copy(), b)
        assert same(a, b), (q, a, b)
        ref.ref_flux(nc, q, n, a)
        orc.orc_fn_flux(q, n, b)
        assert same(a, b)
        F = np.array([rng.normal() * 1e-3, *rng.normal(0, 5, 3), rng.normal() * 1e2, 10 ** rng.uniform(-5, -2)])
        ref.ref_source_make(nc, q, F, a)
        orc.orc_fn_source_make(q, F, b)
        assert same(a, b)
        q4 = random_prim(rng, model)
        for solver in ((0, 1, 2) if model == 0 else (1, 2)):
            ref.ref_riemann(solver, nc, q, q4, n, a)
            orc.orc_fn_riemann(solver, q, q4, n, b)
            assert same(a, b), solver


@pytest.mark.parametrize("evap,intf", [(1, 0), (2, 0), (2, 1), (3, 1), (4, 0), (4, 1), (5, 1)])
def test_evaporation_with_table_psat_bit_for_bit(ref, orc, evap, intf):
    bind_tables(ref, orc)
    rng = np.random.default_rng(40 + 10 * evap + intf)
    ep = np.asarray(cases.water_ep(yinf=0.05, cpv=1900.0), float)
    for _ in range(200):
        Tg, Tp = rng.uniform(300, 1500), rng.uniform(280, 370)
        rhog = rng.uniform(1e5, 5e5) / (287.05 * Tg)
        args = (rhog, Tg, 1.4, 287.05, 1.8e-5, 0.026, Tp, 10 ** rng.uniform(-5.5, -4), rng.uniform(0, 50))
        psat = 101325.0 * math.exp(-ep[5] * (1.0 / Tp - ep[6])) * rng.uniform(0.5, 1.5)
        m1, q1, o1 = D(), D(), C.c_int()
        m2, q2, o2 = D(), D(), C.c_int()
        ref.ref_evaporation_psat(*args, evap, intf, ep, psat, C.byref(m1), C.byref(q1), C.byref(o1))
        orc.orc_fn_evaporation_psat(*args, evap, intf, ep, psat, C.byref(m2), C.byref(q2), C.byref(o2))
        assert same([m1.value, q1.value, o1.value], [m2.value, q2.value, o2.value])


# ------------------------------------------------------------------ closed forms (K-properties-table)
def step_table(t_lo=1, t_hi=5000):
    """K1: cp constant, rho 2000 up to 300 K and 1000 from 301 K"""
    T = np.arange(t_lo, t_hi + 1, dtype=float)
    return properties.material_table(T, np.full(T.size, 1598.0), np.where(T <= 300.0, 2000.0, 1000.0), 1598.0 * T)


def test_density_between_rows_and_outside(orc):
    orc.orc_fn_setup_table.argtypes = [C.c_int, C.c_int, P, C.c_int]
    orc.orc_fn_mat_rho.argtypes = [D]
    orc.orc_fn_mat_rho.restype = D
    tab = step_table()
    orc.orc_fn_setup(1, 1.0, 1.0, 0.0, 0.0)
    orc.orc_fn_setup_table(tab.t_min, tab.n, tab.cols, tab.flags)
    assert abs(orc.orc_fn_mat_rho(300.4) - 1600.0) <= 1e-9
    T = np.arange(280, 401, dtype=float)      # K7: rho = 1500 + 5 (T - 280) on 280..400 K
    k7 = properties.material_table(T, np.full(T.size, 2000.0), 1500.0 + 5.0 * (T - 280.0), 2000.0 * T)
    orc.orc_fn_setup_table(k7.t_min, k7.n, k7.cols, k7.flags)
    assert orc.orc_fn_mat_rho(250.0) == 1500.0 and orc.orc_fn_mat_rho(1e4) == 2100.0


def drag_case(T0=300.4, mode=0):
    """test_oracle_setup's RK3 drag relaxation with K1's table at T_p = 300.4 K"""
    case = cases.case_riemann_1d(n=16, steps=1, limiter=1, mode=mode)
    B = case.blocks[0]
    prim = B.prim[0].reshape((7, 20, 5, 5), order="F")
    prim[0, 2:18], prim[1, 2:18], prim[4, 2:18], prim[5, 2:18], prim[6, 2:18] = 1e-3, 5.0, 1e-3 / 1.5, T0, 1e9
    B.prim[0] = np.ascontiguousarray(prim.ravel(order="F"))
    case.cfg.drag_select = 2
    case.reci[:, native.R_TYPE] = np.where(case.reci[:, native.R_F] <= 2, 300, 0)
    case.meta["tables"] = {1: step_table()}
    return case


def test_drag_time_from_table_density(oracle):
    """u(t) = u_g + (u0 - u_g) P(h) with h = dt/tau, tau = rho_mat dp^2 / (18 mu) at
    rho_mat = 1600, P the RK3 polynomial; the nearest row (2000) would put tau 7.7 % off"""
    r = cases.Runner(oracle, drag_case(), "orc_")
    st = r.step(1)
    q = r.state(1, 1)
    rho_mat, mu = 1600.0, 1.8e-5
    Rp = (0.75 * 1e-3 / (1e9 * np.pi * rho_mat)) ** (1 / 3)
    h = st.dtglobal / ((2 * Rp) ** 2 * rho_mat / (18 * mu))
    assert h > 0.05
    assert np.allclose(q[1, 4:12, 0, 0], 5.0 * (1 - h + h * h / 2 - h ** 3 / 6), rtol=1e-9)


def stationary_case(cs_table=True, mode=0, steps=200):
    """K2: a uniform IG cloud moving at 10 m/s, no gas, T_p = 500.37 K between two rows
    of cp = 1000 + 2 T on 1..5000 K, 200 RK2 steps through 400 faces"""
    case = cases.case_duct400(cases.IG, u=10.0, steps=steps, mode=mode)
    prim = case.blocks[0].prim[0].reshape((7, 20, 5, 5), order="F")
    prim[5, 2:18] = 500.37
    case.blocks[0].prim[0] = np.ascontiguousarray(prim.ravel(order="F"))
    T = np.arange(1, 5001, dtype=float)
    cp = 1000.0 + 2.0 * T if cs_table else np.full(T.size, 2000.0)
    h = 1000.0 * T + T * T if cs_table else 2000.0 * T
    case.meta["tables"] = {1: properties.material_table(T, cp, np.full(T.size, 1000.0), h)}
    return case


def test_varying_cp_keeps_temperature(oracle):
    r = cases.Runner(oracle, stationary_case(), "orc_")
    r.step(200)
    q = r.state(1, 1)
    assert np.max(np.abs(q[5] - 500.37)) <= 1e-9
    assert np.max(np.abs(q[1] - 10.0)) <= 1e-9


def test_loader_refuses_bad_tables():
    T = np.arange(300, 310, dtype=float)
    with pytest.raises(properties.FormatError):
        properties.material_table(T + 0.5, np.full(10, 1.0), np.full(10, 1.0), T)
    with pytest.raises(properties.FormatError):
        properties.material_table(T, np.full(10, 1.0), np.full(10, -1.0), T)
    with pytest.raises(properties.FormatError):
        properties.material_table(T, np.full(10, 1.0), np.full(10, 1.0), T[::-1].copy())
    t = properties.material_table(T, np.full(10, 1.0), np.full(10, 1.0), T, emiss=np.full(10, 1.5))
    assert t.flags == properties.USE_EMISS and np.all(t.col(properties.EMISS) == 1.0)


# ------------------------------------------------------------------ device
@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_drag_time_from_table(gpu, oracle, mode):
    case = drag_case(mode=mode)
    rg, ro, sg, so = run_both(gpu, oracle, case, 1)
    assert sg.dtglobal == so.dtglobal
    assert compare_states(rg, ro, case) == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_varying_cp_keeps_temperature(gpu, oracle, mode):
    case = stationary_case(mode=mode)
    rg, ro, sg, so = run_both(gpu, oracle, case, 200)
    assert compare_states(rg, ro, case) <= 1e-14
    assert np.max(np.abs(rg.state(1, 1)[5] - 500.37)) <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_evaporation_with_full_table(gpu, oracle, mode):
    """MK drops evaporating (CEM, LK interface, blowing) with density, cp, Psat, Lv,
    Cpv and emissivity from the table"""
    case = cases.case_evaporation(n=12, steps=5, mode=mode, table=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 5)
    assert compare_states(rg, ro, case) <= 1e-12


@pytest.mark.gpu
def test_gpu_table_changes_the_result(gpu):
    """non-vacuity: the table's properties move the evaporation result"""
    a = cases.Runner(gpu, cases.case_evaporation(n=12, steps=3, table=True), "ice_gpu_")
    b = cases.Runner(gpu, cases.case_evaporation(n=12, steps=3, table=False), "ice_gpu_")
    a.step(3)
    b.step(3)
    d = np.abs(a.state(1, 1) - b.state(1, 1)).max(axis=(1, 2, 3))
    assert d[0] > 0 and d[4] > 0


def write_stationary_inputs(d, mode="plain"):
    """K2's IG duct in ICE's own input files, with the varying-cp table written as
    ICE's part-properties.dat (Temperature, Cp, Density, Enthalpy)."""
    import os

    from ice_amd import formats
    case = stationary_case(steps=20)
    os.makedirs(os.path.join(d, "INPUT"), exist_ok=True)
    formats.write_ini(os.path.join(d, "input.ini"), {
        "ICE-Parameters": {"iter-threshold": "20"},
        "ICE-Numerics": {"time-scheme": "RK2", "cfl": "0.5", "time-accurate": ".true.",
                         "space-reconstruction": "MUSCL", "flux-limiter": "vanleer"},
        "ICE-Family1": {"closure": "IG"},
        "ICE-IO": {"shell-diter": "20"},
        "ICE-Device": {"mode": mode}})
    x = np.linspace(0.0, 0.16, 17)
    y = z = np.array([0.0, 0.01])
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    prim = case.blocks[0].prim[0].reshape((7, 20, 5, 5), order="F")[:, 2:18, 2:3, 2:3]
    formats.write_tecplot(os.path.join(d, "INPUT", "part-ic.tec"), ["x", "y", "z"] + [f"v{m}" for m in range(7)],
                          [formats.Zone("B1", (17, 2, 2), np.stack([X, Y, Z]), prim, {})])
    formats.write_bc(os.path.join(d, "INPUT", "part-bc.txt"), case.reci, case.recd, case.donor, case.weight)
    T = np.arange(1, 5001, dtype=float)
    with open(os.path.join(d, "INPUT", "part-properties.dat"), "w") as fh:
        fh.write('TITLE = "Mass Thermodynamic Properties"\n')
        fh.write('VARIABLES = "Temperature", "Density", "Enthalpy", "Cp"\nZONE T="A"\nI=5000, F=POINT\n')
        for t in T:
            fh.write("%.1f 1000.0 %.17g %.17g\n" % (t, 1000.0 * t + t * t, 1000.0 + 2.0 * t))
    return case


def test_fortran_host_reads_the_table(tmp_path, oracle):
    """The Fortran host reads the table (columns in any order after Temperature) and
    its run equals the Python host's over the same C-ABI, bit for bit"""
    import os

    from ice_amd import formats
    from test_fortran_host import ABI, run_driver
    d = str(tmp_path / "k2")
    case = write_stationary_inputs(d)
    p = run_driver(d, {"LD_LIBRARY_PATH": ABI, "OMP_NUM_THREADS": "2"})
    assert p.returncode == 0, p.stdout[-800:] + p.stderr[-800:]
    _, zones = formats.read_tecplot(os.path.join(d, "OUTPUT", "part-field.tec"))
    r = cases.Runner(oracle, case, "orc_")
    r.step(20)
    want = r.state(1, 1)
    got = zones[0].cells.reshape(want.shape)
    assert np.array_equal(got, want)
    assert np.max(np.abs(got[5] - 500.37)) <= 1e-9
