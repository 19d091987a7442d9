"""GPU path against the CPU oracle on the same seeded inputs.

Without transcendental functions in the loop (no gas coupling) the device path
computes the oracle's operations in the oracle's order, so states are compared
bit for bit.  With the drag law the device's pow/exp/tanh (ROCm ocml) and the
host's (glibc) may differ in the last bit; those runs are held to 1e-12
relative L2 per variable after N steps."""
import numpy as np
import pytest

from ice_amd import cases

from helpers import compare_states, run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("limiter", list(range(11)))
def test_riemann_1d_bit_exact(gpu, oracle, limiter):
    case = cases.case_riemann_1d(n=400, steps=40, limiter=limiter, drag=False)
    print(f"case riemann1d n=400 limiter={limiter} uncoupled", flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 40)
    assert sg.iter == so.iter == 40
    assert sg.dtglobal == so.dtglobal and sg.time == so.time
    a, b = rg.state(1, 1), ro.state(1, 1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("limiter", [1, 3])
def test_riemann_1d_drag(gpu, oracle, limiter):
    case = cases.case_riemann_1d(n=2000, steps=100, limiter=limiter)
    print(f"case riemann1d n=2000 limiter={limiter} drag", flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 100)
    assert sg.dtglobal == so.dtglobal
    err = compare_states(rg, ro, case)
    print("rel L2", err)
    assert err <= 1e-12


def test_crossing_jets_small(gpu, oracle):
    case = cases.case_crossing_jets(n=64, steps=30)
    print("case crossing_jets n=64", flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 30)
    err = compare_states(rg, ro, case)
    print("rel L2", err)
    assert err <= 1e-12


def test_tiled_equals_untiled(gpu):
    """The LDS-tiled stage kernel and the one-thread-per-cell kernel give the same bits."""
    import ctypes as C
    for case in (cases.case_riemann_1d(n=300, steps=10, limiter=3), cases.case_crossing_jets(n=48, steps=10)):
        print("case", case.name, flush=True)
        ra = cases.Runner(gpu, case, "ice_gpu_")
        rb = cases.Runner(gpu, case, "ice_gpu_")
        assert gpu.ice_gpu_set_tiles(rb.ctx, 0, 0, 0) == 0
        ra.step(10)
        rb.step(10)
        assert np.array_equal(ra.state(1, 1), rb.state(1, 1))


def test_multiblock_maps_and_state(gpu, oracle):
    """Connectivity maps bit for bit, then the state of every block and family."""
    case = cases.case_multiblock(nblocks=10, emin=6, emax=14, steps=20)
    print("case multiblock 10 blocks", flush=True)
    rg = cases.Runner(gpu, case, "ice_gpu_")
    ro = cases.Runner(oracle, case, "orc_")
    for p in range(1, case.cfg.ngroups + 1):
        for a, b in zip(rg.bc_maps(p), ro.bc_maps(p)):
            assert np.array_equal(a, b)
    rg.step(20)
    ro.step(20)
    err = compare_states(rg, ro, case)
    print("rel L2", err)
    assert err <= 1e-12


def test_multiblock_uncoupled_bit_exact(gpu, oracle):
    case = cases.case_multiblock(nblocks=10, emin=6, emax=14, steps=20)
    case.cfg.coupled = 0
    rg, ro, sg, so = run_both(gpu, oracle, case, 20)
    for b in range(1, len(case.blocks) + 1):
        for p in range(1, case.cfg.ngroups + 1):
            assert np.array_equal(rg.state(b, p), ro.state(b, p)), (b, p)


@pytest.mark.parametrize("which", ["multiblock", "jets"])
def test_batched_graph_equals_plain(gpu, which):
    """The batched path (persistent kernels over per-XCD tile queues, the step
    replayed from a hipGraph) gives the plain path's bits, replay after replay,
    with host reads of the state between replays."""
    mk = (lambda mode: cases.case_multiblock(nblocks=12, emin=6, emax=16, steps=6, mode=mode)) if which == "multiblock" \
        else (lambda mode: cases.case_crossing_jets(n=80, steps=6, mode=mode))
    plain, batched = mk(0), mk(1)
    print("case", which, "batched vs plain", flush=True)
    rp = cases.Runner(gpu, plain, "ice_gpu_")
    rb = cases.Runner(gpu, batched, "ice_gpu_")
    for step in range(6):
        sp = rp.step(1)
        sb = rb.step(1)
        assert sp.dtglobal == sb.dtglobal and sp.iter == sb.iter
        for b in range(1, len(plain.blocks) + 1):
            for p in range(1, plain.cfg.ngroups + 1):
                assert np.array_equal(rp.state(b, p), rb.state(b, p)), (step, b, p)


@pytest.mark.parametrize("mode", [0, 1])
def test_residual_norm(gpu, oracle, mode):
    """residual_norm of the last step against the oracle's (fixed-order tree sums
    on the device, sequential sums on the host: 1e-12 relative)."""
    case = cases.case_multiblock(nblocks=10, emin=6, emax=14, steps=5, mode=mode)
    case.cfg.coupled = 0  # the states are then bit-exact and only the summation order differs
    print("case multiblock residual_norm mode", mode, flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 5)
    tg, fg = rg.residual_norm()
    to, fo = ro.residual_norm()
    assert np.all(to > 0)
    assert np.allclose(tg, to, rtol=1e-12, atol=0)
    assert np.allclose(fg, fo, rtol=1e-12, atol=0)


@pytest.mark.parametrize("evap,intf,blow", [(1, 0, 0), (2, 1, 1), (3, 1, 1), (4, 1, 0), (5, 1, 0), (2, 0, 1)])
def test_evaporation_case(gpu, oracle, evap, intf, blow):
    """The evaporation-heavy case (MK, per-cell far-field vapour, drag, heat,
    radiation, evaporation with the Langmuir-Knudsen interface and blowing) at
    20^3 cells, 8 steps: 1e-12 relative L2 per variable."""
    case = cases.case_evaporation(n=20, steps=8, evap=evap, intf=intf, blow=blow)
    print("case evaporation 20^3", evap, intf, blow, flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 8)
    assert sg.n_floored == so.n_floored and sg.bad_state == so.bad_state == 0
    err = compare_states(rg, ro, case)
    print("rel L2", err)
    assert err <= 1e-12


@pytest.mark.parametrize("kw", [dict(), dict(closure=2, ngroups=2), dict(closure=2, ngroups=2, irs=True, evap=2),
                                dict(irs=True)], ids=["ig", "ag2", "ag2-irs-evap", "ig-irs"])
def test_chimera_pair(gpu, oracle, kw):
    """Two overset blocks with chimera ghosts (two donors per layer), optionally
    the combined physics of the steady multi-family case: AG, two families, IRS,
    evaporation: 1e-12 relative L2 per variable after 10 steps."""
    case = cases.case_chimera(steps=10, **kw)
    print("case chimera", kw, flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 10)
    assert sg.bad_state == so.bad_state == 0
    err = compare_states(rg, ro, case)
    print("rel L2", err)
    assert err <= 1e-12


@pytest.mark.parametrize("mode, dispatch", [(0, 0), (1, 0), (1, 1)])
def test_multiblock_irs_steady(gpu, oracle, mode, dispatch):
    """Steady run (per-cell dt) with implicit residual smoothing over 10 blocks,
    uncoupled, plain and batched (both dispatches): bit for bit."""
    case = cases.case_multiblock(nblocks=10, emin=6, emax=14, steps=10, mode=mode)
    case.cfg.coupled = 0
    case.cfg.irs = 1
    case.cfg.time_accurate = 0
    print("case multiblock irs steady mode", mode, "dispatch", dispatch, flush=True)
    rg = cases.Runner(gpu, case, "ice_gpu_")
    assert gpu.ice_gpu_set_dispatch(rg.ctx, dispatch) == 0
    ro = cases.Runner(oracle, case, "orc_")
    rg.step(10)
    ro.step(10)
    for b in range(1, len(case.blocks) + 1):
        for p in range(1, case.cfg.ngroups + 1):
            assert np.array_equal(rg.state(b, p), ro.state(b, p)), (b, p)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("ndir", [1, 2, 3])
def test_irs_each_dimension(gpu, oracle, ndir, mode):
    """Implicit residual smoothing over one, two and three resolved directions
    (the smoothing along i on its own, folded into the march along j, the last
    march doing the update), plain and batched: bit for bit against the oracle."""
    if ndir == 1:
        case = cases.case_riemann_1d(n=300, steps=6, drag=False, mode=mode)
    else:
        case = cases.case_box(n=12, flat=ndir == 2, steps=6, mode=mode)
    case.cfg.irs = 1
    case.cfg.time_accurate = 0  # the smoothing serves steady runs (Assign_Setup)
    print("case irs ndir", ndir, "mode", mode, flush=True)
    rg = cases.Runner(gpu, case, "ice_gpu_")
    ro = cases.Runner(oracle, case, "orc_")
    sg = rg.step(6)
    so = ro.step(6)
    assert sg.bad_state == so.bad_state == 0
    for p in range(1, case.cfg.ngroups + 1):
        assert np.array_equal(rg.state(1, p), ro.state(1, p)), p


@pytest.mark.parametrize("mode", [0, 1])
def test_autotune_keeps_state_and_bits(gpu, mode):
    """The autotuner's trial steps leave the run where it was, and the tile it
    picks gives the same bits as the default one."""
    import ctypes as C
    case = cases.case_multiblock(nblocks=6, emin=8, emax=20, steps=5, mode=mode)
    print("case multiblock 6 blocks autotune mode", mode, flush=True)
    ra = cases.Runner(gpu, case, "ice_gpu_")
    rb = cases.Runner(gpu, case, "ice_gpu_")
    ra.step(1)
    rb.step(1)
    best = np.zeros(4, np.int32)
    ms = np.zeros(24)
    assert gpu.ice_gpu_autotune(ra.ctx, 2, best, ms) == 0, ra.error()
    print("tile", best.tolist(), "ms/step", ms.tolist())
    assert best[0] > 0 and np.isfinite(ms).any()
    sa = ra.step(4)
    sb = rb.step(4)
    assert sa.iter == sb.iter == 5 and sa.time == sb.time
    for b in range(1, len(case.blocks) + 1):
        for p in range(1, case.cfg.ngroups + 1):
            assert np.array_equal(ra.state(b, p), rb.state(b, p))


def test_item_dispatch_equals_queues(gpu):
    """The batched stage as one workgroup per item, in either item order, gives the
    queues' bits."""
    case = cases.case_multiblock(nblocks=6, emin=8, emax=20, steps=4, mode=1)
    ra = cases.Runner(gpu, case, "ice_gpu_")
    ra.step(4)
    for dispatch in (1, 2):
        print("case multiblock 6 blocks dispatch", dispatch, flush=True)
        rb = cases.Runner(gpu, case, "ice_gpu_")
        assert gpu.ice_gpu_set_dispatch(rb.ctx, dispatch) == 0
        rb.step(4)
        for b in range(1, len(case.blocks) + 1):
            for p in range(1, case.cfg.ngroups + 1):
                assert np.array_equal(ra.state(b, p), rb.state(b, p))
        rb.close()
    assert gpu.ice_gpu_set_dispatch(ra.ctx, 3) != 0


def test_combined_batched_irs_evap_chimera(gpu, oracle):
    """The combined configuration's ingredients in one batched run: a chimera
    pair, AG with two families, evaporation with the Langmuir-Knudsen interface
    and Stefan blowing, steady with implicit residual smoothing."""
    case = cases.case_chimera(closure=2, ngroups=2, irs=True, evap=2, steps=6, mode=1)
    print("case chimera ag2 irs evap batched", flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 6)
    err = compare_states(rg, ro, case)
    print("rel L2", err)
    assert err <= 1e-12


@pytest.mark.parametrize("mode", [0, 1])
def test_combined_small(gpu, oracle, mode):
    """The combined configuration at test size: 6 tree blocks and 2 chimera pairs,
    AG x 2 families, full physics with evaporation, steady with IRS."""
    case = cases.case_combined(ntree=6, npairs=2, emin=6, emax=12, pair_n=8, steps=3, mode=mode)
    print("case combined small mode", mode, flush=True)
    rg, ro, sg, so = run_both(gpu, oracle, case, 3)
    assert sg.bad_state == so.bad_state == 0
    err = compare_states(rg, ro, case)
    print("rel L2", err)
    assert err <= 1e-12
