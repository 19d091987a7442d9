"""The coupler's side of a step and the steady stop (Mod_Explicit.f90:150-300).

* The step's source (keep_source 2): each stage's source weighted as the update
  weighs the stage's residual (add_stage_source, RK_accumulate,
  Lib_RK.f90:27-44), what a two-way coupled run hands the gas.  Where no flux
  acts (a uniform cloud in a uniform gas, drag only), the update integrates the
  source alone, so the step's change of a conservative variable over dt is the
  step's source: U_new - U_0 = dt S_step (the invariant RK_accumulate's comment
  states), which the last stage's source does not satisfy.
* res-threshold: a steady run ends after the step whose density residual is at
  most res-threshold times the first non-zero one; the stop iteration follows
  from the residual history of a run stepped one step at a time."""
import numpy as np
import pytest

from ice_amd import cases

from helpers import compare_states


def uniform_drag_case(nrk=3, mode=0):
    case = cases.case_riemann_1d(n=16, steps=1, limiter=1, mode=mode)
    B = case.blocks[0]
    prim = B.prim[0].reshape((7, 20, 5, 5), order="F")
    prim[0, 2:18], prim[1, 2:18], prim[4, 2:18], prim[6, 2:18] = 1e-3, 5.0, 1e-3 / 1.5, 1e9
    B.prim[0] = np.ascontiguousarray(prim.ravel(order="F"))
    case.cfg.drag_select = 2
    case.cfg.nrk = nrk
    return case


@pytest.mark.parametrize("nrk", [2, 3])
def test_step_source_is_what_the_update_integrated(oracle, nrk):
    case = uniform_drag_case(nrk)
    r = cases.Runner(oracle, case, "orc_")
    r.check(r.f("set_keep_source")(r.ctx, 2), "set_keep_source")
    q0 = r.state(1, 1)
    st = r.step(1)
    q1 = r.state(1, 1)
    S, _ = r.source(1, 1)
    mid = slice(4, 12)
    du = (q1[0, mid] * q1[1, mid] - q0[0, mid] * q0[1, mid]) / st.dtglobal
    assert np.allclose(du, S[1, mid], rtol=1e-9, atol=0.0)
    r.check(r.f("set_keep_source")(r.ctx, 1), "set_keep_source")
    q0 = q1
    st = r.step(1)
    S1, _ = r.source(1, 1)
    du = (r.state(1, 1)[0, mid] * r.state(1, 1)[1, mid] - q0[0, mid] * q0[1, mid]) / st.dtglobal
    assert not np.allclose(du, S1[1, mid], rtol=1e-6, atol=0.0)   # the last stage's source alone is not


def steady_case(mode=0):
    case = cases.case_riemann_1d(n=64, steps=1, limiter=1, mode=mode)
    case.cfg.time_accurate = 0
    return case


def stop_iteration(oracle, threshold, nmax):
    r = cases.Runner(oracle, steady_case(), "orc_")
    ref = 0.0
    for it in range(1, nmax + 1):
        r.step(1)
        tot, _ = r.residual_norm()
        if ref == 0.0:
            ref = tot[0]
        if ref > 0.0 and tot[0] <= threshold * ref:
            return it
    return None


def test_res_threshold_stops_where_the_history_says(oracle):
    want = stop_iteration(oracle, 0.2, 300)
    assert want is not None and want > 2
    r = cases.Runner(oracle, steady_case(), "orc_")
    r.set_res_threshold(0.2)
    st = r.step(500)
    assert st.iter == want and r.res_stop()[0]
    r2 = cases.Runner(oracle, steady_case(), "orc_")
    r2.set_res_threshold(0.0)
    assert r2.step(want + 3).iter == want + 3 and not r2.res_stop()[0]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("nrk", [2, 3])
def test_gpu_step_source_equals_oracle(gpu, oracle, nrk, mode):
    case = uniform_drag_case(nrk, mode)
    case.blocks[0].prim[0] = cases.case_riemann_1d(n=16, steps=1, limiter=1).blocks[0].prim[0]
    rs = []
    for lib, pre in ((gpu, "ice_gpu_"), (oracle, "orc_")):
        r = cases.Runner(lib, case, pre)
        r.check(r.f("set_keep_source")(r.ctx, 2), "set_keep_source")
        r.step(3)
        rs.append(r)
    Sg, tg = rs[0].source(1, 1)
    So, to = rs[1].source(1, 1)
    # coupled: the device's pow and exp round differently from the host's libm
    assert np.max(np.abs(Sg - So)) <= 1e-12 * np.max(np.abs(So))
    assert np.max(np.abs(tg - to)) <= 1e-12 * np.max(np.abs(to))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_res_threshold_stops_with_the_oracle(gpu, oracle, mode):
    want = stop_iteration(oracle, 0.2, 300)
    case = steady_case(mode)
    rg = cases.Runner(gpu, case, "ice_gpu_")
    rg.set_res_threshold(0.2)
    st = rg.step(500)
    assert st.iter == want and rg.res_stop()[0]
    ro = cases.Runner(oracle, case, "orc_")
    ro.step(want)
    assert compare_states(rg, ro, case) <= 1e-12
