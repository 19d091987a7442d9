"""Oracle set-up and step composition, pinned by invariants and closed forms
(the block-level Fortran modules need ORION and cannot be built here)."""
import ctypes as C

import numpy as np
import pytest

from ice_amd import cases, geometry, native


def test_pairing_invariants(oracle):
    """Pair_Connections (IO_BC.f90:360-510): partner(partner(n)) = n, exactly one owner
    per face, odd/even ownership, faces numbered by owners in table order; boundary
    cells in k, j, i order with their records by face (collect_boundary_cells)."""
    case = cases.case_multiblock(nblocks=8, emin=5, emax=9, steps=1, ngroups=2)
    r = cases.Runner(oracle, case, "orc_")
    rec = case.reci
    for p in (1, 2):
        partner, owner, cf, first, brec = r.bc_maps(p)
        conn = np.nonzero(np.isin(rec[:, native.R_TYPE], [101, 201]))[0]
        assert len(conn) > 0
        for n0 in conn:
            q = partner[n0] - 1
            assert q >= 0 and partner[q] == n0 + 1
            assert owner[n0] + owner[q] == 1
            assert cf[n0] == cf[q] > 0
            fn, fq = rec[n0, native.R_F], rec[q, native.R_F]
            if fn % 2 != fq % 2:
                assert owner[n0] == (fn % 2 == 0)
        fam = conn[rec[conn, native.R_P] == p]
        owners = [n0 for n0 in fam if owner[n0]]
        assert [cf[n0] for n0 in owners] == list(range(1, len(owners) + 1))
        # boundary cells: every non-zero record of the family once, cells in (b, k, j, i) order, faces sorted
        keys = []
        for c in range(len(first) - 1):
            rr = brec[first[c] - 1:first[c + 1] - 1] - 1
            assert len(set(tuple(rec[x, [native.R_B, native.R_I, native.R_J, native.R_K]]) for x in rr)) == 1
            faces = rec[rr, native.R_F]
            assert np.all(np.diff(faces) >= 0)
            x = rr[0]
            keys.append((rec[x, native.R_B], rec[x, native.R_K], rec[x, native.R_J], rec[x, native.R_I]))
        assert keys == sorted(keys)
        nz = np.nonzero((rec[:, native.R_P] == p) & (rec[:, native.R_TYPE] != 0))[0]
        assert sorted(brec - 1) == sorted(nz)


def test_cartesian_metrics():
    """setup_metrics on a stretched Cartesian block: volumes, face areas and normals,
    metric rows grad(xi) = 1/dx and lengths dl = dx, against their closed forms."""
    rng = np.random.default_rng(3)
    x = np.cumsum(np.r_[0, rng.uniform(0.5, 2.0, 5)])
    y = np.cumsum(np.r_[0, rng.uniform(0.5, 2.0, 4)])
    z = np.cumsum(np.r_[0, rng.uniform(0.5, 2.0, 3)])
    g = geometry.metrics(geometry.cartesian_nodes(x, y, z))
    ni, nj, nk = 5, 4, 3
    dx, dy, dz = np.diff(x), np.diff(y), np.diff(z)
    vol = g["vol"].reshape((ni, nj, nk), order="F")
    assert np.allclose(vol, dx[:, None, None] * dy[None, :, None] * dz[None, None, :], rtol=1e-14)
    fi = g["face_i"].reshape((4, ni + 1, nj, nk), order="F")
    assert np.allclose(fi[0], 1.0) and np.allclose(fi[1:3], 0.0)
    assert np.allclose(fi[3], (dy[:, None] * dz[None, :])[None], rtol=1e-14)
    M = g["M"].reshape((3, 3, ni, nj, nk), order="F")
    assert np.allclose(M[0, 0], 1.0 / dx[:, None, None], rtol=1e-14)
    assert np.allclose(M[1, 1], 1.0 / dy[None, :, None], rtol=1e-14)
    assert np.allclose(M[0, 1], 0.0) and np.allclose(M[2, 0], 0.0)
    dl = g["dl"].reshape((3, ni + 2, nj + 2, nk + 2), order="F")
    assert np.allclose(dl[0, 1:-1, 1, 1], dx, rtol=1e-14)
    # ghost lengths across a face are those of the first cell (extrapolated nodes)
    assert np.isclose(dl[0, 0, 1, 1], dx[0]) and np.isclose(dl[0, -1, 1, 1], dx[-1])


def test_rotated_metrics():
    """a rotated, sheared block: the metric tensor is the inverse transpose of the mean
    edge matrix (M A^T = I) and the volumes sum to the parallelepiped's volume"""
    rng = np.random.default_rng(4)
    A = np.eye(3) + 0.3 * rng.normal(size=(3, 3))
    n = geometry.cartesian_nodes(np.arange(4.0), np.arange(3.0), np.arange(3.0))
    nodes = np.einsum("ij,jabc->iabc", A, n)
    g = geometry.metrics(nodes)
    M = g["M"].reshape((3, 3, 3, 2, 2), order="F")
    assert np.allclose(M[:, :, 1, 1, 1] @ A, np.eye(3), atol=1e-13)
    assert np.isclose(g["vol"].sum(), abs(np.linalg.det(A)) * 3 * 2 * 2, rtol=1e-13)


def periodic(case):
    """every face of every block connected (201) to the opposite face of the same block"""
    rec = case.reci
    dims = case.blocks[0].dims
    rec[:, native.R_TYPE] = 201
    f = rec[:, native.R_F]
    d = (f - 1) // 2
    rec[:, native.R_BS] = rec[:, native.R_B]
    rec[:, native.R_FS] = np.where(f % 2 == 1, f + 1, f - 1)
    src = rec[:, [native.R_I, native.R_J, native.R_K]].copy()
    for dd in range(3):
        m = d == dd
        src[m, dd] = np.where(f[m] % 2 == 1, dims[dd], 1)
    rec[:, native.R_IS], rec[:, native.R_JS], rec[:, native.R_KS] = src[:, 0], src[:, 1], src[:, 2]
    return case


def test_mass_conservation_periodic(oracle):
    """A periodic block (every face 201 to its opposite), no coupling: each connection
    face flux enters one cell and leaves its partner, so the total of every
    conserved quantity is kept to round-off over many RK3 steps."""
    case = periodic(cases.case_multiblock(nblocks=1, emin=10, emax=12, steps=1, ngroups=1))
    case.cfg.coupled = 0
    r = cases.Runner(oracle, case, "orc_")
    vol = case.blocks[0].geo["vol"].reshape(case.blocks[0].dims, order="F")
    q0 = r.state(1, 1)
    m0, px0 = (q0[0] * vol).sum(), (q0[0] * q0[1] * vol).sum()
    r.step(20)
    q1 = r.state(1, 1)
    m1, px1 = (q1[0] * vol).sum(), (q1[0] * q1[1] * vol).sum()
    assert abs(m1 - m0) <= 1e-13 * m0
    assert abs(px1 - px0) <= 1e-12 * (q0[0] * np.abs(q0[1]) * vol).sum()
    assert not np.allclose(q0, q1)


def test_rk3_drag_relaxation(oracle):
    """Uniform cloud in a uniform gas, drag only (Stokes, Cd = 24/Re): no gradients, so
    each step is SSP-RK3 on du/dt = (ug - u)/tau: one step multiplies the slip by
    1 - h + h^2/2 - h^3/6, h = dt/tau (the RK3 stability polynomial)."""
    case = cases.case_riemann_1d(n=16, steps=1, limiter=1)
    B = case.blocks[0]
    prim = B.prim[0].reshape((7, 20, 5, 5), order="F")
    prim[0, 2:18] = 1e-3
    prim[1, 2:18] = 5.0
    prim[4, 2:18] = 1e-3 * 1.0 / 1.5
    prim[6, 2:18] = 1e9
    B.prim[0] = np.ascontiguousarray(prim.ravel(order="F"))
    case.cfg.drag_select = 2
    case.reci[:, native.R_TYPE] = np.where(case.reci[:, native.R_F] <= 2, 300, 0)
    r = cases.Runner(oracle, case, "orc_")
    st = r.step(1)
    q = r.state(1, 1)
    rho_mat, mu = 2700.0, 1.8e-5
    Rp = (0.75 * 1e-3 / (1e9 * np.pi * rho_mat)) ** (1 / 3)
    tau = 8 * rho_mat * Rp / (3 * 1.2 * 24 / (2 * 1.2 * Rp / mu))  # Cd Re = 24: tau = (2 Rp)^2 rho_mat / (18 mu)
    assert np.isclose(tau, (2 * Rp) ** 2 * rho_mat / (18 * mu), rtol=1e-12)
    h = st.dtglobal / tau
    assert h > 0.05
    expect = 5.0 * (1 - h + h * h / 2 - h ** 3 / 6)
    # interior cells away from the mirror ends
    assert np.allclose(q[1, 4:12, 0, 0], expect, rtol=1e-9)


def test_lazy_blocks_upload_as_eager(oracle):
    """A case whose blocks are made as the Runner uploads them (LazyBlock, for the
    full combined configuration) runs as the same case built at once."""
    from ice_amd import cases
    kw = dict(ntree=5, npairs=1, emin=5, emax=9, pair_n=6, steps=2)
    eager = cases.case_combined(**kw)
    lazy = cases.case_combined(lazy=True, **kw)
    assert any(isinstance(B, cases.LazyBlock) for B in lazy.blocks)
    ra = cases.Runner(oracle, eager, "orc_")
    rb = cases.Runner(oracle, lazy, "orc_", workers=3)
    ra.step(2)
    rb.step(2)
    for b in range(1, len(eager.blocks) + 1):
        for p in (1, 2):
            assert np.array_equal(ra.state(b, p), rb.state(b, p))
