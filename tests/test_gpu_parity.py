"""GPU path against the CPU oracle on the same seeded inputs (FP64, 1e-12 relative
L2 per variable after N steps)."""
import numpy as np
import pytest

from ice_amd import cases

from helpers import compare_states, run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("limiter", [1, 3, 8])
def test_riemann_1d(gpu, oracle, limiter):
    case = cases.case_riemann_1d(n=400, steps=40, limiter=limiter)
    print(f"case riemann1d n=400 limiter={limiter}", flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 40)
    assert sg.iter == so.iter == 40
    assert sg.dtglobal == so.dtglobal
    assert abs(sg.time - so.time) <= 1e-15 * abs(so.time)
    err = compare_states(rg, ro, case)
    print("rel L2", err)
    assert err <= 1e-12


def test_crossing_jets_small(gpu, oracle):
    case = cases.case_crossing_jets(n=64, steps=30)
    print("case crossing_jets n=64", flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 30)
    err = compare_states(rg, ro, case)
    print("rel L2", err, "dt", sg.dtglobal, so.dtglobal)
    assert err <= 1e-12
