"""Pin the CPU oracle against the reference's own Fortran (flang-compiled from the
reference sources into oracle/_ref/libiceref.so by oracle/build_ref.sh) on
random inputs: closures, Riemann solvers, limiters, MUSCL reconstruction, the
shock sensor, the drag and heat laws and the RK stage.  Bit-for-bit: both sides
use IEEE double arithmetic in the same order and the same libm."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libiceref.so")
D = C.c_double
P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

NC = {0: 6, 1: 7, 2: 12}


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref/libiceref.so not built (the reference sources are not mounted here)")
    lib = C.CDLL(REF)
    lib.ref_setup.argtypes = [C.c_int, D, D, D, D]
    for n in ("ref_p2c", "ref_c2p"):
        getattr(lib, n).argtypes = [C.c_int, P, P]
    lib.ref_check.argtypes = [C.c_int, P]
    lib.ref_wavespeed.argtypes = [C.c_int, P, P]
    lib.ref_wavespeed.restype = D
    lib.ref_flux.argtypes = [C.c_int, P, P, P]
    lib.ref_source_make.argtypes = [C.c_int, P, P, P]
    lib.ref_riemann.argtypes = [C.c_int, C.c_int, P, P, P, P]
    lib.ref_limiter.argtypes = [C.c_int, D, D, D, D]
    lib.ref_limiter.restype = D
    lib.ref_reconstruct.argtypes = [C.c_int, C.c_int, P, P, P, P, D, D, D, D, D, D, P, P]
    lib.ref_sd.argtypes = [P]
    lib.ref_sd.restype = D
    lib.ref_drag.argtypes = [D, D, D, D, C.c_int]
    lib.ref_drag.restype = D
    lib.ref_heat.argtypes = [D, D, D, C.c_int]
    lib.ref_heat.restype = D
    lib.ref_rk_stage.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, P]
    return lib


@pytest.fixture(scope="module")
def orc(oracle):
    lib = oracle
    lib.orc_fn_setup.argtypes = [C.c_int, D, D, D, D]
    lib.orc_fn_p2c.argtypes = [P, P]
    lib.orc_fn_c2p.argtypes = [P, P]
    lib.orc_fn_check.argtypes = [P]
    lib.orc_fn_wavespeed.argtypes = [P, P]
    lib.orc_fn_wavespeed.restype = D
    lib.orc_fn_flux.argtypes = [P, P, P]
    lib.orc_fn_source_make.argtypes = [P, P, P]
    lib.orc_fn_riemann.argtypes = [C.c_int, P, P, P, P]
    lib.orc_fn_limiter.argtypes = [C.c_int, D, D, D, D]
    lib.orc_fn_limiter.restype = D
    lib.orc_fn_reconstruct.argtypes = [C.c_int, P, P, P, P, D, D, D, D, D, D, P, P]
    lib.orc_fn_sd.argtypes = [P]
    lib.orc_fn_sd.restype = D
    lib.orc_fn_drag.argtypes = [D, D, D, D, C.c_int]
    lib.orc_fn_drag.restype = D
    lib.orc_fn_heat.argtypes = [D, D, D, C.c_int]
    lib.orc_fn_heat.restype = D
    lib.orc_fn_rk_stage.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, P]
    return lib


def random_prim(rng, model, empty=False):
    """a physical primitive state of the closure, occasionally near-empty"""
    nc = NC[model]
    q = np.zeros(nc)
    q[0] = 10 ** rng.uniform(-9 if empty else -5, -1)
    q[1:4] = rng.normal(0, 10, 3)
    if model == 0:
        q[4] = rng.uniform(250, 400)
        q[5] = 10 ** rng.uniform(6, 12)
    elif model == 1:
        q[4] = q[0] * 10 ** rng.uniform(-3, 1)
        q[5] = rng.uniform(250, 400)
        q[6] = 10 ** rng.uniform(6, 12)
    else:
        A = rng.normal(0, 1, (3, 3))
        S = A @ A.T * q[0] * 10 ** rng.uniform(-2, 1)
        q[4:10] = [S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]]
        q[10] = rng.uniform(250, 400)
        q[11] = 10 ** rng.uniform(6, 12)
    return q


def unit(rng):
    n = rng.normal(size=3)
    return n / np.linalg.norm(n)


def same(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_closure_functions(ref, orc, model):
    rng = np.random.default_rng(100 + model)
    nc = NC[model]
    ref.ref_setup(model, 1000.0, 4186.0, 2.26e6, 0.9)
    orc.orc_fn_setup(model, 1000.0, 4186.0, 2.26e6, 0.9)
    for t in range(400):
        q = random_prim(rng, model, empty=(t % 7 == 0))
        n = unit(rng)
        a, b = np.zeros(nc), np.zeros(nc)
        ref.ref_p2c(nc, q, a)
        orc.orc_fn_p2c(q, b)
        assert same(a, b)
        cons = a.copy()
        if t % 5 == 0:
            cons[0] = -cons[0] * 1e-30  # floors
        ref.ref_c2p(nc, cons, a)
        orc.orc_fn_c2p(cons, b)
        assert same(a, b), (q, a, b)
        assert ref.ref_check(nc, q) == orc.orc_fn_check(q)
        assert ref.ref_wavespeed(nc, q, n) == orc.orc_fn_wavespeed(q, n)
        ref.ref_flux(nc, q, n, a)
        orc.orc_fn_flux(q, n, b)
        assert same(a, b)
        F = np.array([rng.normal() * 1e-3, *rng.normal(0, 5, 3), rng.normal() * 1e2, 10 ** rng.uniform(-5, -2)])
        ref.ref_source_make(nc, q, F, a)
        orc.orc_fn_source_make(q, F, b)
        assert same(a, b)


@pytest.mark.parametrize("model,solver", [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 1), (2, 2)])
def test_riemann(ref, orc, model, solver):
    rng = np.random.default_rng(200 + 10 * model + solver)
    nc = NC[model]
    ref.ref_setup(model, 2700.0, 1598.0, 10.8e6, 0.0)
    orc.orc_fn_setup(model, 2700.0, 1598.0, 10.8e6, 0.0)
    for t in range(400):
        q1, q4 = random_prim(rng, model), random_prim(rng, model)
        if t % 9 == 0:
            q4 = q1.copy()
        n = unit(rng)
        a, b = np.zeros(nc), np.zeros(nc)
        ref.ref_riemann(solver, nc, q1, q4, n, a)
        orc.orc_fn_riemann(solver, q1, q4, n, b)
        assert same(a, b), (t, q1, q4, a, b)


def test_limiters(ref, orc):
    rng = np.random.default_rng(300)
    for lim in range(11):
        for t in range(2000):
            x, y = rng.normal(size=2) * 10 ** rng.uniform(-3, 3, 2)
            if t % 11 == 0:
                y = x * rng.uniform(-3, 3)
            if t % 13 == 0:
                x = 0.0
            cx, cy = 2.0 if t % 2 else rng.uniform(1, 4), 2.0 if t % 2 else rng.uniform(1, 4)
            assert ref.ref_limiter(lim, x, y, cx, cy) == orc.orc_fn_limiter(lim, x, y, cx, cy), (lim, x, y, cx, cy)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_reconstruction(ref, orc, model):
    rng = np.random.default_rng(400 + model)
    nc = NC[model]
    ref.ref_setup(model, 2700.0, 1598.0, 10.8e6, 0.0)
    orc.orc_fn_setup(model, 2700.0, 1598.0, 10.8e6, 0.0)
    for t in range(600):
        st = [random_prim(rng, model, empty=(t % 17 == 0)) for _ in range(4)]
        if model == 2 and t % 3 == 0:
            st[1][4:10] *= -1.0  # drives the halving loop to first order
        if t % 5 == 0:
            st[2] = st[1] * (1 + 1e-3)
        L = 10 ** rng.uniform(-4, -2, 4)
        if t % 4 == 0:
            L[0] = 0.0
        dl0, dl1, dl2 = 0.5 * (L[0] + L[1]), 0.5 * (L[1] + L[2]), 0.5 * (L[2] + L[3])
        dll, dlr = 0.5 * L[1], 0.5 * L[2]
        beta = 1.0 if t % 3 else rng.uniform(0, 1)
        lim = t % 11
        a1, a2, b1, b2 = np.zeros(nc), np.zeros(nc), np.zeros(nc), np.zeros(nc)
        ok_r = ref.ref_reconstruct(lim, nc, *st, dl0, dl1, dl2, dll, dlr, beta, a1, a2)
        ok_o = orc.orc_fn_reconstruct(lim, *st, dl0, dl1, dl2, dll, dlr, beta, b1, b2)
        assert ok_r == ok_o
        assert same(a1, b1) and same(a2, b2), (t, lim)


def test_shock_sensor(ref, orc):
    rng = np.random.default_rng(500)
    for t in range(3000):
        r = 10 ** rng.uniform(-6, -2, 27)
        if t % 3 == 0:
            r = 1e-3 * (1 + 1e-3 * rng.normal(size=27))
        r = np.ascontiguousarray(r)
        assert ref.ref_sd(r) == orc.orc_fn_sd(r)


def test_drag_heat(ref, orc):
    rng = np.random.default_rng(600)
    for t in range(3000):
        Re = 10 ** rng.uniform(-4, 5)
        Ma = 10 ** rng.uniform(-4, 0.5)
        G = rng.uniform(1.1, 1.7)
        Tr = rng.uniform(0.3, 3)
        Pr = rng.uniform(0.5, 1.0)
        if t % 50 == 0:
            Re, Ma = 0.0, 0.0
        for sel in range(1, 14):
            a, b = ref.ref_drag(Re, Ma, G, Tr, sel), orc.orc_fn_drag(Re, Ma, G, Tr, sel)
            assert a == b or (np.isnan(a) and np.isnan(b)), (sel, Re, Ma, a, b)
        for sel in range(1, 9):
            a, b = ref.ref_heat(Re, Pr, Ma, sel), orc.orc_fn_heat(Re, Pr, Ma, sel)
            assert a == b or (np.isnan(a) and np.isnan(b)), (sel, Re, Pr, Ma)


def test_rk_stage(ref, orc):
    rng = np.random.default_rng(700)
    for nrk in (1, 2, 3):
        for irk in range(1, nrk + 1):
            for _ in range(200):
                cu, co, r = rng.normal(size=(3, 12))
                a, b = np.zeros(12), np.zeros(12)
                ref.ref_rk_stage(irk, nrk, 12, cu, co, r, a)
                orc.orc_fn_rk_stage(irk, nrk, 12, cu, co, r, b)
                assert same(a, b)


def water_ep(yinf=0.0, alphae=1.0, cpv=0.0):
    """ep(1:8) of water vapour as Setup_Materials packs it: Mv, Lv, cpv, Le, Yinf,
    Lv Mv/Ru, 1/Tboil, alphaE (values of test/verification/F-evaporation/run.py)"""
    Mv, Lv, Tb = 18.015, 2.26e6, 373.15
    return np.array([Mv, Lv, cpv, 1.0, yinf, Lv * Mv / 8314.46, 1.0 / Tb, alphae])


def test_evaporation_models(ref, orc):
    """Every gas-side model with and without the Langmuir-Knudsen interface, over
    the evaporation-heavy ranges (Tg 300-2000 K, 1-5 bar, Yinf 0-0.3, Tp 280-370 K,
    d 2-200 um, Re 0-300), bit for bit against the flang-built module."""
    ref.ref_evaporation.argtypes = [D] * 9 + [C.c_int, C.c_int, P, C.POINTER(D), C.POINTER(D),
                                              C.POINTER(C.c_int)]
    orc.orc_fn_evaporation.argtypes = [D] * 9 + [C.c_int, C.c_int, P, C.POINTER(D), C.POINTER(D),
                                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]
    rng = np.random.default_rng(900)
    lk_iters = []
    for t in range(2500):
        Tg, p = rng.uniform(300, 2000), rng.uniform(1e5, 5e5)
        R, gam, mu, k = 287.05, 1.4, 1.8e-5 * (Tg / 300) ** 0.7, 0.026 * (Tg / 300) ** 0.7
        rhog = p / (R * Tg)
        Tp, dp = rng.uniform(280, 370), 10 ** rng.uniform(np.log10(2e-6), np.log10(2e-4))
        Re = 0.0 if t % 10 == 0 else 10 ** rng.uniform(-3, 2.5)
        ep = water_ep(yinf=rng.uniform(0, 0.3), alphae=10 ** rng.uniform(-2, 0),
                      cpv=0.0 if t % 2 else 1800.0)
        for evap in range(0, 6):
            for intf in (0, 1):
                m1, q1, o1 = D(), D(), C.c_int()
                m2, q2, o2, it = D(), D(), C.c_int(), C.c_int()
                ref.ref_evaporation(rhog, Tg, gam, R, mu, k, Tp, dp, Re, evap, intf, ep, C.byref(m1), C.byref(q1),
                                    C.byref(o1))
                orc.orc_fn_evaporation(rhog, Tg, gam, R, mu, k, Tp, dp, Re, evap, intf, ep, C.byref(m2),
                                       C.byref(q2), C.byref(o2), C.byref(it))
                assert (m1.value, q1.value, o1.value) == (m2.value, q2.value, o2.value), (t, evap, intf)
                if intf == 1 and m1.value < 0:
                    lk_iters.append(it.value)
    assert lk_iters and max(lk_iters) > 2  # the Picard loop is exercised


def test_blowing_factor(ref, orc):
    ref.ref_blowing.argtypes = [D] * 8
    ref.ref_blowing.restype = D
    orc.orc_fn_blowing.argtypes = [D] * 8
    orc.orc_fn_blowing.restype = D
    rng = np.random.default_rng(901)
    for t in range(5000):
        dp = 10 ** rng.uniform(-6, -3.7)
        mp = 1000.0 * np.pi / 6 * dp ** 3
        mdot = -(10 ** rng.uniform(-14, -5)) if t % 7 else rng.uniform(0, 1e-9)
        args = (rng.uniform(1.1, 1.7), 287.05, 1.8e-5, 0.026, 1000.0, dp, mp, mdot)
        assert ref.ref_blowing(*args) == orc.orc_fn_blowing(*args), args
