"""residual_norm (Mod_Diagnostic.f90:22-69, residual_sq :106-127) recomputed with
numpy from the two states of a step: per (block, family), the sum over the
interior cells of the squared changes of the primitive slots 1-4 (rho, u, v, w)
and of the temperature slot nbase - 1; the total over blocks then families, and
each family's over blocks; square roots.  numpy sums in its own order, so the
bar is 1e-12 relative: a wrong slot, a missing square or root, or the wrong pair
of states moves the norms by far more.  Two families of different closures (IG
and AG) and several blocks, steady (per-cell dt) and coupled."""
import numpy as np
import pytest

from ice_amd import cases, native


def two_family_case(mode=0):
    case = cases.case_multiblock(nblocks=5, emin=5, emax=9, ngroups=2, steps=3, mode=mode)
    # family 2 becomes AG, its state extended from the IG state (tensor = P on the diagonal)
    case.fams[1] = cases.make_family(native.AG, native.HLLE, rho_mat=1000.0, cs=4186.0, lv=2.26e6)
    for B in case.blocks:
        ni, nj, nk = B.dims
        q = B.prim[1].reshape((7, ni + 4, nj + 4, nk + 4), order="F")
        a = np.full((12,) + q.shape[1:], np.nan)
        a[0:4] = q[0:4]
        a[4] = a[7] = a[9] = q[4]
        a[5] = a[6] = a[8] = 0.0 * q[4]
        a[10], a[11] = q[5], q[6]
        B.prim[1] = np.ascontiguousarray(a.ravel(order="F"))
    case.cfg.time_accurate = 0
    return case


def states(r, case):
    return [[r.state(b, p) for p in range(1, case.cfg.ngroups + 1)] for b in range(1, len(case.blocks) + 1)]


def numpy_norms(case, before, after):
    ng = case.cfg.ngroups
    part = np.zeros((len(case.blocks), ng, 5))
    for b in range(len(case.blocks)):
        for p in range(ng):
            nc = native.NCOND[case.fams[p].closure]
            d = after[b][p] - before[b][p]
            for m, slot in enumerate([0, 1, 2, 3, nc - 2]):
                part[b, p, m] = np.sum(d[slot] * d[slot])
    return np.sqrt(part.sum(axis=(0, 1))), np.sqrt(part.sum(axis=0))


def check(lib, prefix, mode=0):
    case = two_family_case(mode)
    r = cases.Runner(lib, case, prefix)
    r.step(2)
    before = states(r, case)
    r.step(1)
    after = states(r, case)
    tot, fam = r.residual_norm()
    r.close()
    nt, nf = numpy_norms(case, before, after)
    assert np.all(nt > 0) and np.all(nf > 0)
    assert np.allclose(tot, nt, rtol=1e-12, atol=0), (tot, nt)
    assert np.allclose(fam, nf, rtol=1e-12, atol=0), (fam, nf)
    # the recomputation tells the slots apart: the density slot's norm is not the temperature's
    assert not np.allclose(nt[0], nt[4], rtol=1e-3)


def test_residual_norm_oracle_equals_numpy(oracle):
    check(oracle, "orc_")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_residual_norm_device_equals_numpy(gpu, mode):
    check(gpu, "ice_gpu_", mode)
