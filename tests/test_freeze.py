"""setup_freeze (Mod_Freeze.f90:37-141): which slots a mirror-faced thin mesh
freezes, and that a frozen slot stays exactly zero while the rest evolves.

The rules are checked case by case (2-D plane, 1-D, AG's tensor slots, a k face
that does not mirror, k node planes that are not planes normal to z, the
axisymmetric wedge and its refusals), then on a run: a 2-D AG cloud between
mirror (300) k faces, with a w and P13/P23 perturbation in its initial state, keeps
them zero under the freeze on the oracle, and on the device as on the oracle."""
import numpy as np
import pytest

from ice_amd import cases, geometry, native
from ice_amd.freeze import FreezeError, apply_freeze, setup_freeze
from ice_amd.native import R_F, R_TYPE


def flat_case(closure=native.AG, kcode=300, tilt=0.0, mode=0, n=12):
    case = cases.case_box(n=n, closure=closure, flat=True, steps=4, mode=mode)
    case.reci[np.isin(case.reci[:, R_F], (5, 6)), R_TYPE] = kcode
    nodes = case.meta["nodes"].copy()
    if tilt:
        nodes[2, :, :, 1] += tilt * nodes[0, :, :, 1]
    return case, [nodes]


def records(faces_codes, n=4):
    """a boundary table with one record per (face, code)"""
    rows = np.zeros((len(faces_codes), native.REC_NI), np.int32)
    for m, (f, c) in enumerate(faces_codes):
        rows[m, R_F], rows[m, R_TYPE] = f, c
    return rows


def slab(nz=1, tilt=0.0):
    x = np.linspace(0, 1, 5)
    nodes = geometry.cartesian_nodes(x, x, np.linspace(0, 0.01, nz + 1))
    if tilt:
        nodes[2, :, :, 1] += tilt * nodes[0, :, :, 1]
    return [nodes]


def test_freeze_rules():
    ig, ag, mk = native.IG, native.AG, native.MK
    # 2-D plane mesh, every k face mirrors: w (AG also P13, P23)
    fz = setup_freeze(2, records([(1, 400), (5, 300), (6, 200)]), slab(), [mk, ig, ag])
    assert (fz.freeze_v, fz.freeze_w) == (False, True)
    assert fz.frozen == [[4], [4], [4, 7, 9]]
    assert "2-D plane mesh with mirror k faces: w (with AG also P13 and P23)" in fz.warning
    # one k face that does not mirror: nothing
    fz = setup_freeze(2, records([(5, 300), (6, 400)]), slab(), [ag])
    assert fz.frozen == [[]] and fz.warning is None
    # planar (0) k faces do not mirror
    assert setup_freeze(2, records([(5, 0), (6, 0)]), slab(), [ig]).frozen == [[]]
    # 1-D: v between mirror j faces as well
    fz = setup_freeze(1, records([(3, 301), (4, 300), (5, 200), (6, 200)]), slab(), [ig, ag])
    assert (fz.freeze_v, fz.freeze_w) == (True, True)
    assert fz.frozen == [[3, 4], [3, 4, 6, 7, 9]]
    # 1-D with mirror j faces only: v (AG P12, P23)
    fz = setup_freeze(1, records([(3, 300), (4, 300), (5, 400)]), slab(), [ag])
    assert fz.frozen == [[3, 6, 9]]
    # 3-D meshes freeze nothing
    assert setup_freeze(3, records([(5, 300), (6, 300)]), slab(), [ag]).frozen == [[]]
    # k node planes tilted beyond mesh_fit: w integrated, with the reference's warning
    fz = setup_freeze(2, records([(5, 300), (6, 300)]), slab(tilt=0.05), [ig])
    assert fz.frozen == [[]] and "w is integrated" in fz.warning
    # within mesh_fit (a node off by 1e-3 of the thickness): frozen
    assert setup_freeze(2, records([(5, 300), (6, 300)]), slab(tilt=1e-5), [ig]).frozen == [[4]]


def wedge_nodes(delthe, skew=0.0):
    x = np.linspace(1.0, 2.0, 5)
    r = np.linspace(0.5, 1.0, 5)
    nodes = np.zeros((3, 5, 5, 2))
    for s in (0, 1):
        ang = (s - 0.5) * delthe + (skew if s == 1 else 0.0)
        X, R = np.meshgrid(x, r, indexing="ij")
        nodes[0, :, :, s] = X
        nodes[1, :, :, s] = R * np.cos(ang)
        nodes[2, :, :, s] = R * np.sin(ang)
    return [nodes]


def test_freeze_wedge():
    d = np.radians(1.0)
    fz = setup_freeze(2, records([(1, 400), (5, 200), (6, 200)]), wedge_nodes(d), [native.AG], delthe=d)
    assert fz.frozen == [[4, 7, 9]] and "axisymmetric wedge: the swirl w" in fz.warning
    with pytest.raises(FreezeError, match="must carry 200"):
        setup_freeze(2, records([(5, 200), (6, 300)]), wedge_nodes(d), [native.IG], delthe=d)
    with pytest.raises(FreezeError, match="not the planes at -/\\+ delthe/2"):
        setup_freeze(2, records([(5, 200), (6, 200)]), wedge_nodes(d, skew=0.2 * d), [native.IG], delthe=d)


def perturbed(case):
    """w, P13 and P23 of the initial state set to a small smooth perturbation"""
    B = case.blocks[0]
    a = B.prim[0].reshape(-1, 12)
    g = np.arange(a.shape[0])
    a[:, 3] = 1e-3 * np.sin(0.37 * g)
    a[:, 6] = 1e-6 * np.cos(0.21 * g) * a[:, 0]
    a[:, 8] = 1e-6 * np.sin(0.13 * g) * a[:, 0]
    return case


def run_frozen(lib, prefix, mode=0, freeze=True):
    case, nodes = flat_case(mode=mode)
    perturbed(case)
    if freeze:
        fz = apply_freeze(case, nodes)
        assert fz.frozen == [[4, 7, 9]]
    r = cases.Runner(lib, case, prefix)
    r.step(case.steps)
    s = r.state(1, 1)
    r.close()
    return s


def test_frozen_slots_stay_zero(oracle):
    s = run_frozen(oracle, "orc_")
    assert np.all(s[[3, 6, 8]] == 0.0)
    assert np.all(np.isfinite(s)) and np.any(s[1] != 0.0)
    # without the freeze the perturbation is integrated
    s0 = run_frozen(oracle, "orc_", freeze=False)
    assert np.any(s0[3] != 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_frozen_slots_device(gpu, oracle, mode):
    from helpers import rel_l2
    a = run_frozen(gpu, "ice_gpu_", mode)
    b = run_frozen(oracle, "orc_", mode)
    assert np.all(a[[3, 6, 8]] == 0.0)
    assert rel_l2(a, b) <= 1e-12
