"""Chimera (overset) ghosts of the oracle (Lib_Ghost.f90:439-466), pinned without
the reference: weights sum to one per layer; where the overset blocks abut
(shift 0, every ghost one donor of weight 1) the chimera fill is the
face-matching connection by another route (conservative round trip), so the
two runs agree to round-off (the reference's single-block equivalence check,
docs/vv/multiblock-mpi.md); and a layer-1/layer-2 swap of the donors, planted
here in the case, breaks that agreement."""
import ctypes
import os

import numpy as np
import pytest

from ice_amd import cases, native
from ice_amd.native import R_NCHI1, R_NCHI2, R_DONOR0, R_TYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def orc():
    return native.bind(ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so")), "orc_")


def per_layer_weight_sums(case):
    out = []
    for r in case.reci[case.reci[:, R_TYPE] == 102]:
        d0, n1, n2 = r[R_DONOR0], r[R_NCHI1], r[R_NCHI2]
        out += [case.weight[d0:d0 + n1].sum(), case.weight[d0 + n1:d0 + n1 + n2].sum()]
    return np.array(out)


def run(orc, case, steps):
    r = cases.Runner(orc, case, "orc_")
    r.step(steps)
    return [r.state(b, 1) for b in (1, 2)]


def test_weights_sum_to_one():
    for shift in (3.7, 0.0, 2.25):
        s = per_layer_weight_sums(cases.case_chimera(shift=shift))
        assert s.size and np.allclose(s, 1.0, rtol=0, atol=1e-15)


def test_abutting_chimera_is_the_connection(orc):
    a = run(orc, cases.case_chimera(shift=0.0), 8)
    b = run(orc, cases.case_chimera(shift=0.0, connect=True), 8)
    err = max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip(a, b))
    assert err < 1e-13, err


def test_swapped_layers_break_it(orc):
    case = cases.case_chimera(shift=0.0)
    for r in case.reci[case.reci[:, R_TYPE] == 102]:
        d0 = r[R_DONOR0]
        case.donor[[d0, d0 + 1]] = case.donor[[d0 + 1, d0]]
    a = run(orc, case, 8)
    b = run(orc, cases.case_chimera(shift=0.0, connect=True), 8)
    err = max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip(a, b))
    assert err > 1e-6
