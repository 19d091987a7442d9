"""ice_gpu_residual_norm_noalloc, the residual norm that allocates nothing (its
partial sums live in a file-scope device array), against ice_gpu_residual_norm on
the same state: bit for bit (the same kernels, the totals summed in the same
order), on two calls in a row, the second on a smaller grid; and its refusal of a
grid with more block-family pairs than the array holds."""
import numpy as np
import pytest

from ice_amd import cases


def both(gpu, case, steps=2):
    r = cases.Runner(gpu, case, "ice_gpu_")
    r.step(steps)
    tot, fam = r.residual_norm()
    ng = case.cfg.ngroups
    t1, f1 = np.zeros(5), np.zeros(5 * ng)
    rc = gpu.ice_gpu_residual_norm_noalloc(r.ctx, t1, f1)
    err = r.error() if rc else ""
    r.close()
    return tot, fam.ravel(), t1, f1, rc, err


@pytest.mark.gpu
def test_noalloc_residual_norm_equals_residual_norm(gpu):
    big = cases.case_multiblock(nblocks=12, emin=6, emax=14, ngroups=3, steps=2, mode=1)
    small = cases.case_multiblock(nblocks=3, emin=5, emax=8, ngroups=1, steps=2)
    for case in (big, small):
        print("case multiblock", len(case.blocks), "blocks", case.cfg.ngroups, "families", flush=True)
        tot, fam, t1, f1, rc, err = both(gpu, case)
        assert rc == 0, err
        assert np.all(tot > 0)
        assert np.array_equal(tot, t1) and np.array_equal(fam, f1)


@pytest.mark.gpu
def test_noalloc_residual_norm_limit(gpu):
    """the array holds (196608 - 5 - 5 ngroups) // 325 = 604 pairs: 302 blocks x 2
    families are served (bit for bit), 303 x 2 are refused"""
    for nb, fits in ((302, True), (303, False)):
        case = cases.case_multiblock(nblocks=nb, emin=2, emax=2, ngroups=2, steps=1)
        print("case multiblock", nb, "blocks x 2 families", flush=True)
        tot, fam, t1, f1, rc, err = both(gpu, case, steps=1)
        if fits:
            assert rc == 0, err
            assert np.array_equal(tot, t1) and np.array_equal(fam, f1)
        else:
            assert rc != 0 and "block-family pairs" in err and "ice_gpu_residual_norm" in err
