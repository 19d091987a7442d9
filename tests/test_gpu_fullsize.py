"""Parity at the full sizes of the workloads bench.py measures: the 1024² crossing
jets, the 64-block three-family tree (6.26 M cells) and the 256³ evaporation
block.  The oracle finishes one or two steps of each in seconds on the host's
cores; the properties checked do not depend on size: the device state equals the
oracle's to the path's bar after the same steps, and the batched path gives the
plain path's bits."""
import numpy as np
import pytest

from ice_amd import cases

from helpers import compare_states

pytestmark = pytest.mark.gpu

FULL = {
    "jets2d": lambda mode: cases.case_crossing_jets(n=1024, steps=2, mode=mode),
    "multiblock3d": lambda mode: cases.case_multiblock(mode=mode),
    "evap3d": lambda mode: cases.case_evaporation(n=256, steps=1, mode=mode),
}
STEPS = {"jets2d": 2, "multiblock3d": 1, "evap3d": 1}


@pytest.mark.parametrize("name", sorted(FULL))
def test_full_size_against_oracle(gpu, oracle, name):
    case = FULL[name](0)
    print("full size", name, case.cells, "cells", flush=True)
    rg = cases.Runner(gpu, case, "ice_gpu_")
    sg = rg.step(STEPS[name])
    ro = cases.Runner(oracle, case, "orc_")
    so = ro.step(STEPS[name])
    assert sg.iter == so.iter and sg.bad_state == 0 and so.bad_state == 0
    # uncoupled paths are bit-exact; the coupled ones differ by the device's pow/exp rounding
    err = compare_states(rg, ro, case)
    print(name, "relative L2", err, flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("name", ["jets2d", "multiblock3d"])
def test_full_size_batched_equals_plain(gpu, name):
    rp = cases.Runner(gpu, FULL[name](0), "ice_gpu_")
    rb = cases.Runner(gpu, FULL[name](1), "ice_gpu_")
    sp, sb = rp.step(STEPS[name]), rb.step(STEPS[name])
    assert sp.dtglobal == sb.dtglobal
    case = FULL[name](0)
    for b in range(1, len(case.blocks) + 1):
        for p in range(1, case.cfg.ngroups + 1):
            assert np.array_equal(rp.state(b, p), rb.state(b, p)), (b, p)
