"""Coupler births and imports (Mod_Exchange.f90:212-324, add_births in
Lib_Residual.f90:164-210) and the restart counters (Wrap_Setup.f90:158-214).

  imports  a constant import rate on a cloud at rest, uncoupled, moves its
           conservative state by rate x dt over one step whatever the RK order
           (the residual is the same at every stage, and the SSP weights sum to
           it)
  births   packets added at once: every conservative slot grows by its packet,
           two packets into one cell in their order
  restart  2N steps equal N steps, then a new context from the state, tau and the
           counters, then N steps, bit for bit, inside the CFL ramp (without the
           counters the ramp restarts and the states differ)
The device does each as the oracle does (-m gpu)."""
import numpy as np
import pytest

from ice_amd import cases, coupler, geometry, native

CS, RHO_MAT = 900.0, 1000.0


def cons_mk(q):
    """the MK conservative state of primitive cells q (6, ...) at constant cs"""
    rho, u = q[0], q[1:4]
    return np.stack([rho, rho * u[0], rho * u[1], rho * u[2], rho * CS * q[4] + 0.5 * rho * np.sum(u * u, axis=0),
                     q[5]])


def rest_case(closure=native.MK, nrk=3, n=4, mode=0, uniform=False):
    dims = (n, n, n)
    geo = geometry.uniform_metrics(dims, 1e-3)
    nc = native.NCOND[closure]

    def fill(I, J, K):
        z = np.zeros(I.shape)
        rho = 1e-2 * (1.0 + (0.0 if uniform else 0.1) * np.sin(I + 2.0 * J + 3.0 * K))
        cols = [rho, z, z, z]
        if closure == native.IG:
            cols += [rho * 0.5]
        elif closure == native.AG:
            cols += [rho * 0.5, z, z, rho * 0.5, z, rho * 0.5]
        return np.stack(cols + [z + 300.0, rho / 4e-12])

    prim = cases.padded_state(dims, nc, fill)
    tb = cases.TableBuilder(1)
    tb.add_block(1, dims, lambda f, i, j, k: np.full(len(f), 400, np.int32))
    reci, recd, donor, weight = tb.arrays()
    cfg = cases.make_config(ngroups=1, nrk=nrk, time_accurate=True, ndir=3, limiter=3, dt_max=1e-4, mode=mode)
    fam = cases.make_family(closure, rho_mat=RHO_MAT, cs=CS)
    return cases.Case("rest", cfg, [fam], [cases.Block(dims, geo, [prim])], reci, recd, donor, weight, 1)


def imports_of(case, varying=True):
    n = case.blocks[0].dims
    nc = native.NCOND[case.fams[0].closure]
    cells = int(np.prod(n))
    pk = coupler.birth_packet(2e-3, [1.0, -2.0, 0.5], 310.0, 20e-6, RHO_MAT, CS)
    r = coupler.birth_rate(case.fams[0].closure, pk)
    imp = np.tile(r, cells)  # (nc, ni, nj, nk) Fortran: the slot fastest
    scale = 1.0 + (0.01 if varying else 0.0) * np.arange(cells)
    return (imp.reshape(cells, nc) * scale[:, None]).ravel(), r, scale


def run_imports(lib, prefix, nrk, mode=0):
    """One step with imports: with Euler on a varying cloud and varying rates (the
    fluxes of the cloud at rest are zero), with RK2/RK3 on a uniform cloud and
    uniform rates (the cloud set moving stays uniform, so its fluxes cancel)."""
    case = rest_case(nrk=nrk, mode=mode, uniform=nrk > 1)
    imp, r, scale = imports_of(case, varying=nrk == 1)
    run = cases.Runner(lib, case, prefix)
    before = run.state(1, 1)
    assert getattr(lib, prefix + "put_imports")(run.ctx, 1, 1, imp.ctypes.data) == 0
    run.step(1)
    after = run.state(1, 1)
    run.close()
    return before, after, r, scale


@pytest.mark.parametrize("nrk", [1, 2, 3])
def test_imports_move_the_state_by_rate_dt(oracle, nrk):
    before, after, r, scale = run_imports(oracle, "orc_", nrk)
    dt = 1e-4   # at rest the step is dt-max
    c0, c1 = cons_mk(before), cons_mk(after)
    n = before.shape[1:]
    want = c0 + (r[:, None] * scale[None, :]).reshape((6,) + n, order="F") * dt
    assert np.allclose(c1, want, rtol=1e-13, atol=0)


def test_absorption_packets():
    """absorption_rate: no number; IG's m |u|^2 and AG's m u u"""
    pk = coupler.absorption_packet(3e-3, [1.0, 2.0, -1.0], 300.0, CS)
    for cl in (native.MK, native.IG, native.AG):
        r = coupler.absorption_rate(cl, pk)
        nc = native.NCOND[cl]
        assert r[nc - 1] == 0.0 and np.array_equal(r[:4], pk[:4]) and r[nc - 2] == pk[4]
    assert np.isclose(coupler.absorption_rate(native.IG, pk)[4], 3e-3 * 6.0)
    ag = coupler.absorption_rate(native.AG, pk)
    assert np.allclose(ag[4:10], 3e-3 * np.array([1.0, 2.0, -1.0, 4.0, -2.0, 1.0]))
    with pytest.raises(ValueError):
        coupler.absorption_rate(native.MK, coupler.absorption_packet(0.0, [0, 0, 0], 300.0, CS))


def run_births(lib, prefix, mode=0):
    case = rest_case(mode=mode)
    run = cases.Runner(lib, case, prefix)
    before = run.state(1, 1)
    cells = np.array([[1, 1, 1, 1], [1, 2, 3, 4], [1, 1, 1, 1], [1, 4, 4, 4]], np.int32)
    pks = np.stack([coupler.birth_packet(m, [0.5, 0.0, -0.5], 320.0, 15e-6, RHO_MAT, CS)
                    for m in (1e-3, 2e-3, 3e-3, 4e-3)])
    assert getattr(lib, prefix + "add_births")(run.ctx, 1, 4, cells.ravel(), pks.ravel()) == 0
    after = run.state(1, 1)
    run.close()
    return before, after, cells, pks


def test_births_add_their_packets(oracle):
    before, after, cells, pks = run_births(oracle, "orc_")
    c0, c1 = cons_mk(before), cons_mk(after)
    add = np.zeros_like(c0)
    for (b, i, j, k), pk in zip(cells, pks):
        add[:, i - 1, j - 1, k - 1] += pk
    assert np.allclose(c1, c0 + add, rtol=1e-13, atol=1e-300)
    untouched = np.all(add == 0.0, axis=0)
    assert np.array_equal(after[:, untouched], before[:, untouched])


def restart_case(mode=0):
    case = cases.case_multiblock(nblocks=4, emin=5, emax=8, ngroups=2, steps=6, mode=mode)
    case.cfg.cfl_ramp_iter = 10
    return case


def restarted(lib, prefix, n, counters=True, mode=0):
    """(2n steps in one go, n steps + a restart of n) states of every block and family"""
    case = restart_case(mode)
    a = cases.Runner(lib, case, prefix)
    a.step(2 * n)
    whole = [[a.state(b, p) for p in (1, 2)] for b in range(1, 5)]
    a.close()
    r = cases.Runner(lib, case, prefix)
    r.step(n)
    nb, ng = len(case.blocks), 2
    saved, taus = {}, {}
    for b in range(1, nb + 1):
        ni, nj, nk = case.blocks[b - 1].dims
        for p in range(1, ng + 1):
            out = np.empty(7 * (ni + 4) * (nj + 4) * (nk + 4))
            r.check(getattr(lib, prefix + "get_state")(r.ctx, b, p, out), "get_state")
            saved[(b, p)] = out
            t = np.empty(ni * nj * nk)
            r.check(getattr(lib, prefix + "get_tau")(r.ctx, b, p, t), "get_tau")
            taus[(b, p)] = t
    import ctypes as C
    it, tm, ref = C.c_int32(), C.c_double(), C.c_double()
    r.check(getattr(lib, prefix + "get_counters")(r.ctx, C.byref(it), C.byref(tm), C.byref(ref)), "get_counters")
    assert it.value == n
    r.close()
    # the new run: the saved state in place of the initial condition
    for b in range(1, nb + 1):
        for p in range(1, ng + 1):
            case.blocks[b - 1].prim[p - 1] = saved[(b, p)]
    s = cases.Runner(lib, case, prefix)
    for b in range(1, nb + 1):
        for p in range(1, ng + 1):
            s.check(getattr(lib, prefix + "put_tau")(s.ctx, b, p, taus[(b, p)]), "put_tau")
    if counters:
        s.check(getattr(lib, prefix + "set_counters")(s.ctx, it.value, tm.value, ref.value), "set_counters")
    s.step(n)
    again = [[s.state(b, p) for p in (1, 2)] for b in range(1, 5)]
    s.close()
    return whole, again


def test_restart_is_exact(oracle):
    whole, again = restarted(oracle, "orc_", 3)
    for wb, ab in zip(whole, again):
        for w, a in zip(wb, ab):
            assert np.array_equal(w, a)
    whole, again = restarted(oracle, "orc_", 3, counters=False)
    assert any(not np.array_equal(w, a) for wb, ab in zip(whole, again) for w, a in zip(wb, ab))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_imports_births_restart_device(gpu, oracle, mode):
    for nrk in (1, 3):
        g = run_imports(gpu, "ice_gpu_", nrk, mode)
        o = run_imports(oracle, "orc_", nrk, mode)
        assert np.array_equal(g[1], o[1])
    g = run_births(gpu, "ice_gpu_", mode)
    o = run_births(oracle, "orc_", mode)
    assert np.array_equal(g[1], o[1])
    whole, again = restarted(gpu, "ice_gpu_", 3, mode=mode)
    for wb, ab in zip(whole, again):
        for w, a in zip(wb, ab):
            assert np.array_equal(w, a)
