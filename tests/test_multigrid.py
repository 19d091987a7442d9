"""Grid sequencing between levels (Mod_Multigrid.f90:93-131,
Lib_Multigrid.f90:40-191): restriction (volume-weighted conservative average
over the coarse cell's fine cells) and prolongation (trilinear interpolation of
the coarse primitives, neighbours clamped to the block).

The oracle is pinned by what the two operators must do: restriction keeps a
uniform state and the mass of the block; prolongation reproduces a field linear
in the coarse cell index at every fine cell whose neighbours are not clamped
(the fine cell at coarse coordinate i +- 1/4 gets f(i +- 1/4)); restriction of
such a prolongation returns the linear field.  The GPU is compared with the
oracle bit for bit, and over a coarse-then-fine sequence of steps."""
import numpy as np
import pytest

from ice_amd import cases
from ice_amd.native import AG, IG, MK

from helpers import compare_states


def levels(closure=IG, flat=False, n=8, mode=0):
    fine = cases.case_box(n, closure, flat, mode=mode)
    return fine, cases.coarsen(fine)


def put(r, case, prim_fn):
    """interior state (nc, ni, nj, nk) from prim_fn(I, J, K) into runner r"""
    B = case.blocks[0]
    prim = cases.padded_state(B.dims, prim_fn(*np.meshgrid(*[np.arange(1, d + 1) for d in B.dims],
                                                                indexing="ij")).shape[0], prim_fn)
    r.check(r.f("put_state")(r.ctx, 1, 1, prim), "put_state")


def linear_ig(I, J, K):
    z = np.zeros(I.shape)
    rho = 1.0 + 0.1 * I + 0.2 * J + 0.05 * K
    return np.stack([rho, z, z, z, 0.5 + 0.01 * I + 0.02 * J + 0.03 * K, z + 300.0, 1e9 * rho])


@pytest.mark.parametrize("flat", [False, True])
def test_restriction_keeps_uniform_state_and_mass(oracle, flat):
    fine, coarse = levels(IG, flat)
    rf, rc = cases.Runner(oracle, fine, "orc_"), cases.Runner(oracle, coarse, "orc_")
    q = rf.state(1, 1)
    assert rc.f("restrict")(rf.ctx, rc.ctx) == 0
    Q = rc.state(1, 1)
    vf = fine.blocks[0].geo["vol"].reshape(fine.blocks[0].dims, order="F")
    vc = coarse.blocks[0].geo["vol"].reshape(coarse.blocks[0].dims, order="F")
    assert abs((Q[0] * vc).sum() - (q[0] * vf).sum()) <= 1e-13 * (q[0] * vf).sum()
    assert abs((Q[6] * vc).sum() - (q[6] * vf).sum()) <= 1e-13 * (q[6] * vf).sum()
    put(rf, fine, lambda I, J, K: np.stack([np.full(I.shape, v) for v in (1.3, 2.0, -1.0, 0.5, 0.7, 310.0, 1e9)]))
    rc.f("restrict")(rf.ctx, rc.ctx)
    Q = rc.state(1, 1)
    for v, want in enumerate((1.3, 2.0, -1.0, 0.5, 0.7, 310.0, 1e9)):
        assert np.allclose(Q[v], want, rtol=1e-14, atol=0.0), v


@pytest.mark.parametrize("flat", [False, True])
def test_prolongation_is_trilinear(oracle, flat):
    fine, coarse = levels(IG, flat)
    rf, rc = cases.Runner(oracle, fine, "orc_"), cases.Runner(oracle, coarse, "orc_")
    put(rc, coarse, linear_ig)
    assert rc.f("prolong")(rc.ctx, rf.ctx) == 0
    q = rf.state(1, 1)
    ni, nj, nk = fine.blocks[0].dims
    I, J, K = np.meshgrid(np.arange(1, ni + 1), np.arange(1, nj + 1), np.arange(1, nk + 1), indexing="ij")
    want = linear_ig(I / 2.0 + 0.25, J / 2.0 + 0.25, (np.ones(K.shape) if flat else K / 2.0 + 0.25))
    inner = (slice(2, ni - 2), slice(2, nj - 2), slice(None) if flat else slice(2, nk - 2))
    for v in (0, 4, 6):
        assert np.allclose(q[v][inner], want[v][inner], rtol=1e-13, atol=0.0), v
    # restriction of the prolongation returns the linear field inside
    assert rc.f("restrict")(rf.ctx, rc.ctx) == 0
    Q = rc.state(1, 1)
    ci, cj, ck = coarse.blocks[0].dims
    I, J, K = np.meshgrid(np.arange(1, ci + 1), np.arange(1, cj + 1), np.arange(1, ck + 1), indexing="ij")
    lin = linear_ig(I, J, K)
    inner = (slice(1, ci - 1), slice(1, cj - 1), slice(None) if flat else slice(1, ck - 1))
    for v in (0, 4, 6):
        assert np.allclose(Q[v][inner], lin[v][inner], rtol=1e-13, atol=0.0), v


def test_levels_must_halve(oracle):
    fine = cases.case_box(8, IG)
    other = cases.case_box(6, IG)
    rf, ro = cases.Runner(oracle, fine, "orc_"), cases.Runner(oracle, other, "orc_")
    assert ro.f("restrict")(rf.ctx, ro.ctx) != 0


@pytest.mark.gpu
@pytest.mark.parametrize("closure,flat", [(MK, True), (IG, False), (AG, False), (AG, True)])
def test_gpu_operators_equal_oracle(gpu, oracle, closure, flat):
    fine, coarse = levels(closure, flat)
    out = []
    for lib, pre in ((gpu, "ice_gpu_"), (oracle, "orc_")):
        rf, rc = cases.Runner(lib, fine, pre), cases.Runner(lib, coarse, pre)
        assert rc.f("restrict")(rf.ctx, rc.ctx) == 0
        Q = rc.state(1, 1)
        rc.step(3)
        assert rc.f("prolong")(rc.ctx, rf.ctx) == 0
        out.append((Q, rf.state(1, 1)))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_sequence_coarse_then_fine(gpu, oracle, mode):
    """10 steady steps on the coarse level from the restricted start, prolongation,
    5 on the fine level"""
    fine, coarse = levels(IG, False, n=16, mode=mode)
    runs = []
    for lib, pre in ((gpu, "ice_gpu_"), (oracle, "orc_")):
        rf, rc = cases.Runner(lib, fine, pre), cases.Runner(lib, coarse, pre)
        rc.f("restrict")(rf.ctx, rc.ctx)
        rc.step(10)
        rc.f("prolong")(rc.ctx, rf.ctx)
        rf.step(5)
        runs.append(rf)
    assert compare_states(runs[0], runs[1], fine) <= 1e-12


def write_box_inputs(d, fine, levels, iters):
    """A case_box level set in ICE's own input files: the finest level's IC and
    boundary table, part-bc<m>.txt for the coarser ones, [ICE-Multigrid]."""
    import os

    from ice_amd import formats
    os.makedirs(os.path.join(d, "INPUT"), exist_ok=True)
    ini = {"ICE-Parameters": {"iter-threshold": str(iters[0])},
           "ICE-Numerics": {"time-scheme": "RK2", "cfl": "0.5", "time-accurate": ".false.",
                            "space-reconstruction": "MUSCL", "flux-limiter": "vanleer"},
           "ICE-Family1": {"closure": "IG"},
           "ICE-IO": {"shell-diter": "1000"},
           "ICE-Multigrid": {"levels": str(levels), **{f"level{m + 1}-iter": str(n) for m, n in enumerate(iters)}}}
    formats.write_ini(os.path.join(d, "input.ini"), ini)
    nodes = fine.meta["nodes"]
    ni, nj, nk = fine.blocks[0].dims
    prim = fine.blocks[0].prim[0].reshape((7, ni + 4, nj + 4, nk + 4), order="F")[:, 2:ni + 2, 2:nj + 2, 2:nk + 2]
    formats.write_tecplot(os.path.join(d, "INPUT", "part-ic.tec"), ["x", "y", "z"] + [f"v{m}" for m in range(7)],
                          [formats.Zone("B1", nodes.shape[1:], nodes, prim, {})])
    case = fine
    for m in range(1, levels + 1):
        name = "part-bc.txt" if m == 1 else f"part-bc{m}.txt"
        formats.write_bc(os.path.join(d, "INPUT", name), case.reci, case.recd, case.donor, case.weight)
        case = cases.coarsen(case)
    with open(os.path.join(d, "INPUT", "part-properties.dat"), "w") as fh:
        fh.write('TITLE = "Mass Thermodynamic Properties"\nVARIABLES = "Temperature", "Cp", "Density", "Enthalpy"\n')
        fh.write('ZONE T="A"\nI=3, F=POINT\n')
        for T in (1.0, 2.0, 3.0):
            fh.write(f"{T} 900.0 1000.0 {900.0 * T}\n")


def test_fortran_host_runs_the_levels(tmp_path, oracle):
    """[ICE-Multigrid] levels = 2 through the Fortran host over the oracle's C-ABI:
    the same field as the Python sequence (restrict, 6 coarse steps, prolong, 4 fine)"""
    import os

    from ice_amd import formats
    from test_fortran_host import ABI, run_driver
    fine = cases.case_box(8, IG, True)
    d = str(tmp_path / "mg")
    write_box_inputs(d, fine, 2, [4, 6])
    p = run_driver(d, {"LD_LIBRARY_PATH": ABI, "OMP_NUM_THREADS": "2"})
    assert p.returncode == 0, p.stdout[-800:] + p.stderr[-800:]
    _, zones = formats.read_tecplot(os.path.join(d, "OUTPUT", "part-field.tec"))
    rf, rc = cases.Runner(oracle, fine, "orc_"), cases.Runner(oracle, cases.coarsen(fine), "orc_")
    rc.f("restrict")(rf.ctx, rc.ctx)
    rc.step(6)
    rc.f("prolong")(rc.ctx, rf.ctx)
    rf.step(4)
    want = rf.state(1, 1)
    got = zones[0].cells.reshape(want.shape)
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-30)) <= 1e-12
