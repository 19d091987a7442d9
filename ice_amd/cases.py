"""Synthetic cases: blocks, boundary tables, states and gas fields in the
layouts of include/ice_gpu.h, built the way ICE's set-up builds them.

The boundary table follows Setup_BC (src/lib/io/IO_BC.f90:65-344): block by
block, every face cell of the block (2(nj nk + ni nk + ni nj) of them), faces 1..6,
one copy for all families, record idx = (n-1)*ngroups + p.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional

import numpy as np

from . import geometry
from .native import (AG, GC, HLLE, IG, MK, NCOND, RD_ALPHA, RD_BETA, RD_MASSFLUX, RD_RADIUS, RD_TEMPERATURE,
                     RD_VELOCITY, REC_ND, REC_NI, RUSANOV, SAUREL, Config, Family, R_B, R_BS, R_DONOR0, R_F, R_FS,
                     R_I, R_IS, R_J, R_JS, R_K, R_KS, R_NCHI1, R_NCHI2, R_NORMAL_DIR, R_P, R_TYPE)

PI = 3.14159265358979323846


@dataclasses.dataclass
class Block:
    dims: tuple
    geo: dict
    prim: List[np.ndarray]            # per family, flat Fortran (nc, -1:ni+2, -1:nj+2, -1:nk+2)
    gas: Optional[dict] = None        # prim (5,ni,nj,nk), R, gam, k, mu flat Fortran
    yinf: Optional[np.ndarray] = None


class LazyBlock:
    """A block whose arrays are made only when a Runner uploads it (make() returns
    the Block), so that a large case never holds every block's arrays at once."""

    def __init__(self, dims, make):
        self.dims = tuple(dims)
        self.make = make


def block_data(B) -> Block:
    """The arrays of a block, made now if it is a LazyBlock."""
    return B.make() if isinstance(B, LazyBlock) else B


@dataclasses.dataclass
class Case:
    name: str
    cfg: Config
    fams: List[Family]
    blocks: List[Block]
    reci: np.ndarray                  # (nrec, REC_NI) int32
    recd: np.ndarray                  # (nrec, REC_ND) float64
    donor: np.ndarray                 # (nd, 4) int32
    weight: np.ndarray                # (nd,)
    steps: int = 1
    meta: Dict = dataclasses.field(default_factory=dict)

    @property
    def cells(self) -> int:
        return sum(b.dims[0] * b.dims[1] * b.dims[2] for b in self.blocks)

    def cell_stages(self, steps: int) -> int:
        """interior cells x families x RK stages x steps (the Mcell-stage count)"""
        return self.cells * self.cfg.ngroups * self.cfg.nrk * steps


def make_config(**kw) -> Config:
    c = Config()
    c.ngroups = kw.get("ngroups", 1)
    c.nrk = kw.get("nrk", 3)
    c.time_accurate = int(kw.get("time_accurate", True))
    c.cfl_ramp_iter = kw.get("cfl_ramp_iter", 0)
    c.ndir = kw.get("ndir", 3)
    c.limiter = kw.get("limiter", 3)
    c.shock_detector = int(kw.get("shock_detector", False))
    c.irs = int(kw.get("irs", False))
    c.coupled = int(kw.get("coupled", False))
    c.drag_select = kw.get("drag_select", 13)
    c.heat_select = kw.get("heat_select", 8)
    c.blow_select = kw.get("blow_select", 0)
    c.mode = kw.get("mode", 0)
    c.var_yinf = int(kw.get("var_yinf", False))
    c.cfl = kw.get("cfl", 0.5)
    c.dt_max = kw.get("dt_max", 1e-4)
    c.tau_factor = kw.get("tau_factor", 1.0)
    c.irs_beta = kw.get("irs_beta", 0.5)
    return c


def make_family(closure: int, riemann: Optional[int] = None, rho_mat=2700.0, cs=1598.0, lv=10.8e6, emiss=0.0,
                frozen=(), evap=0, intf=0, ep=None) -> Family:
    f = Family()
    f.closure = closure
    f.riemann = (SAUREL if closure == MK else RUSANOV) if riemann is None else riemann
    f.nfrozen = len(frozen)
    for m, s in enumerate(frozen):
        f.frozen[m] = s
    f.evap_select, f.intf_select, f.surrogate = evap, intf, 0
    f.rho_mat, f.cs, f.lv, f.emiss = rho_mat, cs, lv, emiss
    for m, v in enumerate(ep if ep is not None else [0.0] * 8):
        f.ep[m] = v
    return f


def face_cells(dims):
    """(f, i, j, k) of every face cell of a block, faces 1..6, first tangential index fastest."""
    return [tuple(int(x) for x in row) for row in face_cell_array(dims)]


def face_cell_array(dims) -> np.ndarray:
    """(n, 4) int array of (f, i, j, k) of every face cell, in face_cells order."""
    ni, nj, nk = dims
    out = []
    for f in range(1, 7):
        if f <= 2:
            J, K = np.meshgrid(np.arange(1, nj + 1), np.arange(1, nk + 1), indexing="xy")
            I = np.full_like(J, 1 if f == 1 else ni)
        elif f <= 4:
            I, K = np.meshgrid(np.arange(1, ni + 1), np.arange(1, nk + 1), indexing="xy")
            J = np.full_like(I, 1 if f == 3 else nj)
        else:
            I, J = np.meshgrid(np.arange(1, ni + 1), np.arange(1, nj + 1), indexing="xy")
            K = np.full_like(I, 1 if f == 5 else nk)
        out.append(np.stack([np.full(I.size, f), I.ravel(), J.ravel(), K.ravel()], axis=1))
    return np.concatenate(out).astype(np.int32)


class TableBuilder:
    """Accumulates the boundary table of a case, block by block: one row per face
    cell, fanned out over the families (one copy for all families)."""

    def __init__(self, ngroups: int):
        self.ng = ngroups
        self.ints = []      # (n, REC_NI) arrays, family column filled later
        self.reals = []
        self.donor = []
        self.weight = []

    def add_block(self, b: int, dims, code_fn, ints_fn=None, reals_fn=None):
        """code_fn(f, i, j, k arrays) -> codes; ints_fn/reals_fn(rows, fijk) fill payload columns in place."""
        fijk = face_cell_array(dims)
        n = fijk.shape[0]
        ri = np.zeros((n, REC_NI), np.int32)
        rd = np.zeros((n, REC_ND), np.float64)
        ri[:, R_B] = b
        ri[:, R_F], ri[:, R_I], ri[:, R_J], ri[:, R_K] = fijk[:, 0], fijk[:, 1], fijk[:, 2], fijk[:, 3]
        ri[:, R_TYPE] = code_fn(fijk[:, 0], fijk[:, 1], fijk[:, 2], fijk[:, 3])
        if ints_fn is not None:
            ints_fn(ri, fijk)
        if reals_fn is not None:
            reals_fn(rd, fijk)
        self.ints.append(ri)
        self.reals.append(rd)

    def add(self, b, f, i, j, k, code, ints=None, reals=None):
        ri = np.zeros((1, REC_NI), np.int32)
        rd = np.zeros((1, REC_ND), np.float64)
        ri[0, R_B], ri[0, R_I], ri[0, R_J], ri[0, R_K], ri[0, R_F], ri[0, R_TYPE] = b, i, j, k, f, code
        for key, v in (ints or {}).items():
            ri[0, key] = v
        for key, v in (reals or {}).items():
            rd[0, key] = v
        self.ints.append(ri)
        self.reals.append(rd)

    def arrays(self):
        ri = np.concatenate(self.ints)
        rd = np.concatenate(self.reals)
        reci = np.repeat(ri, self.ng, axis=0)
        recd = np.repeat(rd, self.ng, axis=0)
        reci[:, R_P] = np.tile(np.arange(1, self.ng + 1, dtype=np.int32), ri.shape[0])
        donor = np.asarray(self.donor, np.int32).reshape(-1, 4) if self.donor else np.zeros((0, 4), np.int32)
        weight = np.asarray(self.weight, np.float64)
        return np.ascontiguousarray(reci), np.ascontiguousarray(recd), donor, weight


def padded_state(dims, nc, fill_fn) -> np.ndarray:
    """Flat Fortran (nc, -1:ni+2, -1:nj+2, -1:nk+2) array, NaN ghosts, interior from fill_fn(i,j,k arrays) -> (nc, ...)"""
    ni, nj, nk = dims
    a = np.full((nc, ni + 4, nj + 4, nk + 4), np.nan)
    I, J, K = np.meshgrid(np.arange(1, ni + 1), np.arange(1, nj + 1), np.arange(1, nk + 1), indexing="ij")
    a[:, 2:ni + 2, 2:nj + 2, 2:nk + 2] = fill_fn(I, J, K)
    return np.ascontiguousarray(a.ravel(order="F"))


def padded_fields(dims, fields) -> np.ndarray:
    """padded_state from the nc interior fields (each (ni, nj, nk), best laid out
    in Fortran order, or scalars), written straight into the Fortran layout."""
    ni, nj, nk = dims
    a = np.full((nk + 4, nj + 4, ni + 4, len(fields)), np.nan)   # the Fortran layout, read backwards
    for v, f in enumerate(fields):
        a[2:nk + 2, 2:nj + 2, 2:ni + 2, v] = np.asarray(f).T
    return a.ravel()


def gas_arrays(prim5: np.ndarray, R, gam, k, mu) -> dict:
    """prim5: (5, ni, nj, nk) rho u v w T; scalars or (ni,nj,nk) arrays for the rest."""
    shp = prim5.shape[1:]
    full = lambda x: np.broadcast_to(np.asarray(x, float), shp)  # noqa: E731
    f = lambda a: np.ascontiguousarray(np.asarray(a, float).ravel(order="F"))  # noqa: E731
    return {"prim": f(prim5), "R": f(full(R)), "gam": f(full(gam)), "k": f(full(k)), "mu": f(full(mu))}


# ---------------------------------------------------------------- case 1
def case_riemann_1d(n: int = 2000, steps: int = 1000, limiter: int = 1, drag: bool = True, mode: int = 0) -> Case:
    """Plain method, CPU-runnable: 1-D block of n cells on [0,1], IG, Riemann problem
    split at x = 0.5, frozen uniform gas at rest, MUSCL minmod, Rusanov, SSP-RK3,
    CFL 0.5, time-accurate, transmissive (400) ends, type-0 side faces.

    Granular energy e maps to IG's pseudo-pressure as P = rho e / 1.5
    (e = 3 theta / 2 with P = rho theta, cons(5) = rho |u|^2 + 3 P)."""
    dims = (n, 1, 1)
    dx = 1.0 / n
    nodes = geometry.cartesian_nodes(np.linspace(0.0, 1.0, n + 1), [0.0, dx], [0.0, dx])
    geo = geometry.metrics(nodes)
    T = 300.0

    def fill(I, J, K):
        x = (I - 0.5) * dx
        left = x < 0.5
        rho = np.where(left, 1e-3, 5e-4)
        num = np.where(left, 1e9, 5e8)
        u = np.where(left, 10.0, -10.0)
        e = np.where(left, 1.0, 0.5)
        z = np.zeros_like(x)
        return np.stack([rho, u, z, z, rho * e / 1.5, z + T, num])

    prim = padded_state(dims, 7, fill)
    g5 = np.zeros((5,) + dims)
    g5[0] = 1.2
    g5[4] = 300.0
    gas = gas_arrays(g5, 287.05, 1.4, 0.026, 1.8e-5) if drag else None
    tb = TableBuilder(1)
    tb.add_block(1, dims, lambda f, i, j, k: np.where(f <= 2, 400, 0))
    reci, recd, donor, weight = tb.arrays()
    cfg = make_config(ngroups=1, nrk=3, time_accurate=True, ndir=1, limiter=limiter, coupled=drag,
                      drag_select=4, heat_select=8, cfl=0.5, mode=mode)
    fam = make_family(IG, RUSANOV)
    return Case("riemann1d", cfg, [fam], [Block(dims, geo, [prim], gas)], reci, recd, donor, weight, steps,
                meta={"dx": dx})


def case_duct400(closure: int = MK, u: float = -10.0, cells=None, codes=None, inlet=None, tilt: float = 0.0,
                  gas_u: Optional[float] = None, rho0: Optional[float] = None, steps: int = 20, mode: int = 0) -> Case:
    """A duct of 16 cells on [0, 0.16] (cross-section 0.01 x 0.01) with extrapolation
    faces (400) on faces 1-4 and type 0 on 5-6, a uniform cloud moving along x at u
    in `cells` (1-based, every cell by default; the others empty), 20 time-accurate
    RK2 steps at CFL 0.5, MUSCL van Leer: the setting of ICE's inflow-400 test
    (test/fast/inflow-400/run.py:57-140).  MK at 1e-3 kg/m3, IG and AG at 1 kg/m3
    (IG P = 25; AG P11 = 100, P22 = P33 = 1), T = 300 K, drops of 10 um radius of a
    1000 kg/m3, 900 J/(kg K) material.  `codes` replaces the face codes, with the
    402 payload `inlet` (mass flux, speed, temperature, radius, along the normal);
    `tilt` moves the nodes of i = 0 by -tilt*y along x; `gas_u` adds a gas moving
    along x that drags the cloud (Stokes, no heat, dt-max 1e-6)."""
    nx, lx, rad, rho_al, cs = 16, 0.16, 1e-5, 1000.0, 900.0
    mass = rho_al * 4.0 / 3.0 * PI * rad ** 3
    dims = (nx, 1, 1)
    nodes = geometry.cartesian_nodes(np.linspace(0.0, lx, nx + 1), [0.0, lx / nx], [0.0, lx / nx])
    if tilt:
        nodes[0, 0] = nodes[0, 0] - tilt * nodes[1, 0]
    geo = geometry.metrics(nodes)
    if rho0 is None:
        rho0 = 1e-3 if closure == MK else 1.0
    held = np.array([1.0 if cells is None or i in cells else 0.0 for i in range(1, nx + 1)])

    def fill(I, J, K):
        a = held[I - 1]
        rho, z = rho0 * a, np.zeros(I.shape)
        cols = [rho, u * a, z, z]
        if closure == IG:
            cols += [25.0 * a]
        elif closure == AG:
            cols += [100.0 * a, z, z, a, z, a]
        return np.stack(cols + [z + 300.0, rho / mass])

    prim = padded_state(dims, NCOND[closure], fill)
    codes = codes or {1: 400, 2: 400, 3: 400, 4: 400, 5: 0, 6: 0}
    tb = TableBuilder(1)

    def payload(rd, fijk):
        if inlet is not None:
            on = np.isin(fijk[:, 0], [f for f, c in codes.items() if c == 402])
            rd[on, RD_MASSFLUX], rd[on, RD_VELOCITY], rd[on, RD_TEMPERATURE], rd[on, RD_RADIUS] = inlet

    def normal_dir(ri, fijk):
        ri[:, R_NORMAL_DIR] = 1

    tb.add_block(1, dims, lambda f, i, j, k: np.vectorize(codes.get)(f), normal_dir, payload)
    reci, recd, donor, weight = tb.arrays()
    gas = None
    if gas_u is not None:
        g5 = np.zeros((5,) + dims)
        g5[0], g5[1], g5[4] = 1.2, gas_u, 300.0
        gas = gas_arrays(g5, 287.05, 1.4, 0.026, 1.8e-5)
    cfg = make_config(ngroups=1, nrk=2, time_accurate=True, ndir=1, limiter=3, coupled=gas is not None,
                      drag_select=2, heat_select=8, cfl=0.5, dt_max=1e-6 if gas is not None else 1e-4, mode=mode)
    fam = make_family(closure, rho_mat=rho_al, cs=cs)
    return Case("duct400", cfg, [fam], [Block(dims, geo, [prim], gas)], reci, recd, donor, weight, steps)


def case_box(n: int = 16, closure: int = IG, flat: bool = False, seed: int = 20261209, steps: int = 10,
             mode: int = 0, nodes=None, steady: bool = True) -> Case:
    """One block of n x n x n cells (flat: n x n x 1) on the unit box, a smooth random
    cloud (density, velocity, pressure or tensor, temperature), extrapolation (400)
    faces, steady RK2, MUSCL van Leer: a level of ICE's grid sequencing.  `nodes`
    gives the block's nodes instead (a coarser level: coarsen())."""
    rng = np.random.default_rng(seed)
    if nodes is None:
        x = np.linspace(0.0, 1.0, n + 1)
        nodes = geometry.cartesian_nodes(x, x, [0.0, 1.0 / n] if flat else x)
    geo = geometry.metrics(nodes)
    dims = geo["dims"]
    nc = NCOND[closure]

    def fill(I, J, K):
        X, Y, Z = (I - 0.5) / dims[0], (J - 0.5) / dims[1], (K - 0.5) / dims[2]
        ph = rng.uniform(0, 2 * PI, 8)
        s = lambda m: np.sin(2 * PI * X + ph[m]) * np.cos(2 * PI * Y + ph[m + 1])  # noqa: E731
        rho = 1.0 + 0.3 * s(0)
        cols = [rho, 2.0 * s(1), 2.0 * s(2), (0.0 if flat else 1.0) * s(3) + 0.0 * Z]
        if closure == IG:
            cols += [rho * (1.0 + 0.2 * s(4))]
        elif closure == AG:
            z = np.zeros(I.shape)
            cols += [rho * (1.0 + 0.2 * s(4)), 0.05 * rho * s(5), z, rho * (1.0 + 0.1 * s(6)), z, rho + z]
        return np.stack(cols + [300.0 + 10.0 * s(2), rho / 1e-12])

    prim = padded_state(dims, nc, fill)
    tb = TableBuilder(1)
    tb.add_block(1, dims, lambda f, i, j, k: np.where(f <= (4 if flat else 6), 400, 0))
    reci, recd, donor, weight = tb.arrays()
    cfg = make_config(ngroups=1, nrk=2, time_accurate=not steady, ndir=2 if flat else 3, limiter=3, mode=mode)
    fam = make_family(closure, rho_mat=1000.0, cs=900.0)
    return Case("box", cfg, [fam], [Block(dims, geo, [prim])], reci, recd, donor, weight, steps,
                meta={"nodes": nodes, "closure": closure, "flat": flat})


def coarsen(case: Case) -> Case:
    """The next coarser level of a case_box (Setup_Multigrid, Mod_Multigrid.f90:37-70):
    every other node (a flat block keeps its two k nodes, Coarse_Grid), the same faces."""
    nodes = case.meta["nodes"]
    flat = case.meta["flat"]
    coarse = nodes[:, ::2, ::2, :] if flat else nodes[:, ::2, ::2, ::2]
    c = case_box(closure=case.meta["closure"], flat=flat, nodes=np.ascontiguousarray(coarse), mode=case.cfg.mode,
                 steady=not case.cfg.time_accurate)
    return c


# ---------------------------------------------------------------- case 2
def fourier_velocity(X, Y, seed: int, modes: int = 16, umax: float = 5.0):
    """Divergence-free-looking 2-D field: sum of `modes` sine modes with wave vectors
    2 pi (kx, ky), kx, ky in 1..8, random amplitudes and phases, scaled to max |u| = umax."""
    rng = np.random.default_rng(seed)
    u = np.zeros_like(X)
    v = np.zeros_like(X)
    for _ in range(modes):
        kx, ky = rng.integers(1, 9, size=2)
        amp = rng.normal()
        ph = rng.uniform(0, 2 * np.pi)
        th = rng.uniform(0, 2 * np.pi)
        s = np.sin(2 * np.pi * (kx * X + ky * Y) + ph)
        u += amp * np.cos(th) * s
        v += amp * np.sin(th) * s
    mag = np.sqrt(u * u + v * v).max()
    return u * umax / mag, v * umax / mag


def case_crossing_jets(n: int = 1024, steps: int = 500, seed: int = 20261204, mode: int = 0,
                       time_accurate: bool = True) -> Case:
    """2-D crossing jets: n x n cells on the unit square, AG, HLLE, MUSCL van Leer
    with the Jameson sensor, SSP-RK3, FP64.  Two 402 inlet jets on face 1 at +-30
    degrees (y in [0.25,0.35] and [0.65,0.75]), rho_p = 1e-2, n = 1e10, |u| = 20;
    the rest of face 1 symmetry (300), faces 2-4 extrapolation (400), faces 5/6
    planar (0).  Frozen gas: 16 random Fourier modes (max |u| = 5 m/s), 300 K.
    Drag Schiller-Naumann, heat Ranz-Marshall.  The domain starts empty."""
    dims = (n, n, 1)
    d = 1.0 / n
    nodes = geometry.cartesian_nodes(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1), [0.0, d])
    geo = geometry.metrics(nodes)
    rho_mat = 1000.0
    rho_j, n_j, spd = 1e-2, 1e10, 20.0
    radius = (0.75 * rho_j / (n_j * PI * rho_mat)) ** (1.0 / 3.0)
    xc = (np.arange(1, n + 1) - 0.5) * d
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    ug, vg = fourier_velocity(X, Y, seed)
    g5 = np.zeros((5,) + dims)
    g5[0] = 1.2
    g5[1, :, :, 0] = ug
    g5[2, :, :, 0] = vg
    g5[4] = 300.0
    gas = gas_arrays(g5, 287.05, 1.4, 0.026, 1.8e-5)
    rho0 = 1e-8
    m_drop = rho_mat * 4.0 / 3.0 * PI * radius ** 3

    def fill(I, J, K):
        z = np.zeros(I.shape)
        uu = ug[I - 1, J - 1]
        vv = vg[I - 1, J - 1]
        th = rho0 * 1e-2
        return np.stack([z + rho0, uu, vv, z, z + th, z, z, z + th, z, z + th, z + 300.0, z + rho0 / m_drop])

    prim = padded_state(dims, 12, fill)
    tb = TableBuilder(1)
    g = rho_j * spd * math.cos(math.radians(30.0))

    def jet(f, j):
        y = (j - 0.5) * d
        return (f == 1) & (((y >= 0.25) & (y <= 0.35)) | ((y >= 0.65) & (y <= 0.75)))

    def codes(f, i, j, k):
        return np.where(jet(f, j), 402, np.where(f == 1, 300, np.where(f <= 4, 400, 0)))

    def reals(rd, fijk):
        m = jet(fijk[:, 0], fijk[:, 2])
        y = (fijk[:, 2] - 0.5) * d
        rd[m, RD_MASSFLUX] = g
        rd[m, RD_VELOCITY] = spd
        rd[m, RD_ALPHA] = np.where(y[m] < 0.5, math.radians(30.0), -math.radians(30.0))
        rd[m, RD_BETA] = 0.0
        rd[m, RD_TEMPERATURE] = 300.0
        rd[m, RD_RADIUS] = radius

    tb.add_block(1, dims, codes, reals_fn=reals)
    reci, recd, donor, weight = tb.arrays()
    cfg = make_config(ngroups=1, nrk=3, time_accurate=time_accurate, ndir=2, limiter=3, shock_detector=True,
                      coupled=True, drag_select=4, heat_select=6, cfl=0.5, mode=mode)
    fam = make_family(AG, HLLE, rho_mat=rho_mat, cs=4186.0, lv=2.26e6)
    return Case("crossing_jets", cfg, [fam], [Block(dims, geo, [prim], gas)], reci, recd, donor, weight, steps,
                meta={"radius": radius})


# ---------------------------------------------------------------- driving
def _made_ahead(blocks, lazy, workers: int):
    """The LazyBlocks `lazy` (1-based) made in order, up to `workers` of them
    being made at a time on threads (numpy releases the interpreter lock in its
    array loops), so the host holds only the blocks in flight."""
    import concurrent.futures as cf
    if not lazy:
        return iter(())

    def gen():
        with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
            pending = []
            nxt = 0
            while nxt < len(lazy) and len(pending) < max(1, workers):
                pending.append(ex.submit(blocks[lazy[nxt] - 1].make))
                nxt += 1
            while pending:
                blk = pending.pop(0).result()
                if nxt < len(lazy):
                    pending.append(ex.submit(blocks[lazy[nxt] - 1].make))
                    nxt += 1
                yield blk
                del blk
    return gen()


class Runner:
    """Feeds a Case to a library with the ice_gpu_ C-ABI (prefix 'ice_gpu_') or its
    CPU restatement (prefix 'orc_'), and reads states back."""

    def __init__(self, lib, case: Case, prefix: str = "ice_gpu_", device: int = 0, workers: int = 4):
        import ctypes as C
        self.lib, self.case, self.pre = lib, case, prefix
        self.C = C
        fams = (Family * len(case.fams))(*case.fams)
        dims = np.asarray([d for b in case.blocks for d in b.dims], np.int32)
        ctx = C.c_void_p()
        rc = self.f("create")(device, C.byref(case.cfg), fams, len(case.blocks), dims, C.byref(ctx))
        self.ctx = ctx
        self.check(rc, "create")
        for b, B in enumerate(case.blocks, start=1):
            if not isinstance(B, LazyBlock):
                self.put_geometry(b, B)
        self.check(self.f("set_boundaries")(ctx, case.reci.shape[0], np.ascontiguousarray(case.reci),
                                            np.ascontiguousarray(case.recd), case.donor.shape[0],
                                            np.ascontiguousarray(case.donor.reshape(-1)),
                                            np.ascontiguousarray(case.weight)), "set_boundaries")
        for p, tab in case.meta.get("tables", {}).items():
            self.check(self.f("set_properties")(ctx, p, tab.t_min, tab.n, tab.cols, tab.flags), "set_properties")
        lazy = [b for b, B in enumerate(case.blocks, start=1) if isinstance(B, LazyBlock)]
        made = _made_ahead(case.blocks, lazy, workers)
        for b, B in enumerate(case.blocks, start=1):
            if isinstance(B, LazyBlock):
                B = next(made)
                self.put_geometry(b, B)
            for p in range(1, case.cfg.ngroups + 1):
                self.check(self.f("put_state")(ctx, b, p, B.prim[p - 1]), "put_state")
            if B.gas is not None and case.cfg.coupled:
                gg = B.gas
                self.check(self.f("put_gas")(ctx, b, gg["prim"], gg["R"], gg["gam"], gg["k"], gg["mu"]), "put_gas")
            if B.yinf is not None and case.cfg.var_yinf:
                self.check(self.f("put_gas_yinf")(ctx, b, B.yinf), "put_gas_yinf")

    def put_geometry(self, b: int, B: Block):
        g = B.geo
        self.check(self.f("set_geometry")(self.ctx, b, g["vol"], g["dl"], g["M"], g["face_i"], g["face_j"],
                                          g["face_k"]), "set_geometry")

    def f(self, name):
        return getattr(self.lib, self.pre + name)

    def error(self) -> str:
        buf = self.C.create_string_buffer(512)
        self.f("last_error")(self.ctx, buf, 512)
        return buf.value.decode()

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{self.pre}{what} failed ({rc}): {self.error()}")

    def step(self, n: int):
        from .native import Status
        st = Status()
        rc = self.f("step")(self.ctx, n, self.C.byref(st))
        if rc != 0:
            raise RuntimeError(f"{self.pre}step failed ({rc}): {self.error()}")
        return st

    def state(self, b: int, p: int) -> np.ndarray:
        """interior state (nc, ni, nj, nk) of family p in block b"""
        B = self.case.blocks[b - 1]
        nc = NCOND[self.case.fams[p - 1].closure]
        ni, nj, nk = B.dims
        out = np.empty(nc * (ni + 4) * (nj + 4) * (nk + 4))
        self.check(self.f("get_state")(self.ctx, b, p, out), "get_state")
        a = out.reshape((nc, ni + 4, nj + 4, nk + 4), order="F")
        return a[:, 2:ni + 2, 2:nj + 2, 2:nk + 2].copy()

    def residual_norm(self):
        """(total[5], fam[ngroups, 5]) norms of rho, u, v, w, T changes over the last step"""
        tot = np.zeros(5)
        fam = np.zeros(5 * self.case.cfg.ngroups)
        self.check(self.f("residual_norm")(self.ctx, tot, fam), "residual_norm")
        return tot, fam.reshape(self.case.cfg.ngroups, 5)

    def bc_maps(self, p: int):
        nrec = self.case.reci.shape[0]
        partner = np.zeros(nrec, np.int32)
        owner = np.zeros(nrec, np.int32)
        cf = np.zeros(nrec, np.int32)
        cap = nrec + 2
        first = np.zeros(cap, np.int32)
        rec = np.zeros(cap, np.int32)
        ncell = self.C.c_int32()
        self.check(self.f("get_bc_maps")(self.ctx, partner, owner, cf, p, self.C.byref(ncell), first, rec, cap),
                   "get_bc_maps")
        n = ncell.value
        return partner, owner, cf, first[:n + 1].copy(), rec[:first[n] - 1].copy()

    def set_res_threshold(self, threshold: float):
        self.check(self.f("set_res_threshold")(self.ctx, threshold), "set_res_threshold")

    def res_stop(self):
        """(stopped, reference) of the last step call"""
        s, r = self.C.c_int32(), self.C.c_double()
        self.check(self.f("res_stop")(self.ctx, self.C.byref(s), self.C.byref(r)), "res_stop")
        return bool(s.value), r.value

    def source(self, b: int, p: int):
        """(source (nc, ni, nj, nk), tau (ni, nj, nk)) kept by the last step"""
        B = self.case.blocks[b - 1]
        nc = NCOND[self.case.fams[p - 1].closure]
        ni, nj, nk = B.dims
        src, tau = np.zeros(nc * ni * nj * nk), np.zeros(ni * nj * nk)
        self.check(self.f("get_source")(self.ctx, b, p, src, tau), "get_source")
        return src.reshape((nc, ni, nj, nk), order="F"), tau.reshape((ni, nj, nk), order="F")

    def inflow400_lines(self):
        """(lines, dropped): the 400-inflow lines logged since the last call"""
        from .native import INFLOW400_LOG, Inflow400Line
        out = (Inflow400Line * INFLOW400_LOG)()
        n, dropped = self.C.c_int32(), self.C.c_int64()
        self.check(self.f("inflow400_lines")(self.ctx, out, INFLOW400_LOG, self.C.byref(n), self.C.byref(dropped)),
                   "inflow400_lines")
        return [out[k] for k in range(n.value)], dropped.value

    def inflow400_totals(self):
        from .native import Inflow400Total
        cap = 6 * len(self.case.blocks) * self.case.cfg.ngroups
        out = (Inflow400Total * cap)()
        n = self.C.c_int32()
        self.check(self.f("inflow400_totals")(self.ctx, out, cap, self.C.byref(n)), "inflow400_totals")
        return [out[k] for k in range(n.value)]

    def close(self):
        if self.ctx:
            self.f("destroy")(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- case 3
def smooth_field(rng, X, Y, Z, modes: int = 16, kmax: int = 4):
    """sum of `modes` random sine modes over the box, wave numbers 1..kmax per axis"""
    f = np.zeros_like(X)
    for _ in range(modes):
        k = rng.integers(1, kmax + 1, size=3) * rng.choice([-1, 1], size=3)
        f += rng.normal() * np.sin(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) + rng.uniform(0, 2 * np.pi))
    return f


def block_tree(nblocks: int, emin: int, emax: int, rng):
    """A seeded random tree of face-matching blocks: each new block takes a free face
    of an existing block, copies its two edges, draws its third (log-uniform), and
    sits against that face with the same index directions.  Returns dims, origins
    (in cells) and the joins (a, fa, b, fb)."""
    logu = lambda: int(round(np.exp(rng.uniform(np.log(emin), np.log(emax)))))  # noqa: E731
    dims = [(logu(), logu(), logu())]
    orig = [(0, 0, 0)]
    free = [(0, f) for f in range(1, 7)]
    joins = []
    while len(dims) < nblocks:
        a, fa = free.pop(int(rng.integers(len(free))))
        d = (fa - 1) // 2
        fb = fa + 1 if fa % 2 == 1 else fa - 1
        nd = list(dims[a])
        nd[d] = logu()
        o = list(orig[a])
        o[d] = orig[a][d] + dims[a][d] if fa % 2 == 0 else orig[a][d] - nd[d]
        b = len(dims)
        dims.append(tuple(nd))
        orig.append(tuple(o))
        joins.append((a, fa, b, fb))
        free += [(b, f) for f in range(1, 7) if f != fb]
    return dims, orig, joins


def case_multiblock(nblocks: int = 64, emin: int = 16, emax: int = 128, seed: int = 20261205, steps: int = 100,
                    ngroups: int = 3, mode: int = 0, limiter: int = 3, dx: float = 1e-3) -> Case:
    """3-D multi-block batching: `nblocks` blocks with edges log-uniform on [emin, emax],
    joined by face-matching connections (101) into a seeded random tree, extrapolation
    (400) elsewhere.  IG closure, `ngroups` families with lognormal radii (median
    5/20/60 um, sigma 0.4) giving n = rho / (rho_mat 4/3 pi r^3); rho_p ~ 1e-3 (1 +
    0.3 noise), P = rho a^2 / 3 with a = 10 m/s.  Frozen gas: smooth random velocity
    field (|u| <= 30 m/s) at 300-600 K.  Rusanov, MUSCL, SSP-RK3, drag, FP64."""
    rng = np.random.default_rng(seed)
    dims, orig, joins = block_tree(nblocks, emin, emax, rng)
    rho_mat = 1000.0
    medians = [5e-6, 20e-6, 60e-6][:ngroups] + [30e-6] * max(0, ngroups - 3)
    lo = np.min(np.asarray(orig), axis=0)
    hi = np.max(np.asarray(orig) + np.asarray(dims), axis=0)
    span = (hi - lo).astype(float)
    gseed = rng.integers(1 << 30)
    blocks = []
    for b, (dd, o) in enumerate(zip(dims, orig)):
        ni, nj, nk = dd
        xs = [(o[m] + np.arange(dd[m] + 1)) * dx for m in range(3)]
        geo = geometry.metrics(geometry.cartesian_nodes(*xs))
        c = [((o[m] - lo[m]) + np.arange(dd[m]) + 0.5) / span[m] for m in range(3)]
        X, Y, Z = np.meshgrid(*c, indexing="ij")
        g_rng = np.random.default_rng(gseed)  # same field for every block, sampled at the block's cells
        u = smooth_field(g_rng, X, Y, Z)
        v = smooth_field(g_rng, X, Y, Z)
        w = smooth_field(g_rng, X, Y, Z)
        t = smooth_field(g_rng, X, Y, Z)
        blocks.append([dd, geo, (X, Y, Z), (u, v, w, t)])
    umax = max(np.sqrt(bb[3][0] ** 2 + bb[3][1] ** 2 + bb[3][2] ** 2).max() for bb in blocks)
    tmin = min(bb[3][3].min() for bb in blocks)
    tmax = max(bb[3][3].max() for bb in blocks)
    out = []
    for b, (dd, geo, (X, Y, Z), (u, v, w, t)) in enumerate(blocks):
        g5 = np.zeros((5,) + dd)
        g5[0] = 1.2
        g5[1], g5[2], g5[3] = 30.0 * u / umax, 30.0 * v / umax, 30.0 * w / umax
        g5[4] = 300.0 + 300.0 * (t - tmin) / (tmax - tmin)
        gas = gas_arrays(g5, 287.05, 1.4, 0.026, 1.8e-5)
        prims = []
        for p in range(ngroups):
            prng = np.random.default_rng([seed, b, p])

            def fill(I, J, K, prng=prng, p=p):
                shp = I.shape
                rho = 1e-3 * (1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) * np.sin(2 * np.pi * Z))
                r = medians[p] * np.exp(0.4 * prng.standard_normal(shp))
                num = rho / (rho_mat * 4.0 / 3.0 * PI * r ** 3)
                z = np.zeros(shp)
                return np.stack([rho, g5[1] * 0.9, g5[2] * 0.9, g5[3] * 0.9, rho * 100.0 / 3.0, z + 300.0, num])

            prims.append(padded_state(dd, 7, fill))
        out.append(Block(dd, geo, prims, gas))
    # boundary table: 101 across the joins, 400 elsewhere
    partner_face = {}
    for a, fa, b, fb in joins:
        partner_face[(a, fa)] = (b, fb)
        partner_face[(b, fb)] = (a, fa)
    tb = TableBuilder(ngroups)
    for b, dd in enumerate(dims):
        tb.add_block(b + 1, dd, tree_codes(partner_face, b), ints_fn=tree_ints(partner_face, dims, b))
    reci, recd, donor, weight = tb.arrays()
    cfg = make_config(ngroups=ngroups, nrk=3, time_accurate=True, ndir=3, limiter=limiter, coupled=True,
                      drag_select=4, heat_select=8, cfl=0.5, mode=mode)
    fams = [make_family(IG, RUSANOV, rho_mat=rho_mat, cs=4186.0, lv=2.26e6) for _ in range(ngroups)]
    return Case("multiblock", cfg, fams, out, reci, recd, donor, weight, steps,
                meta={"joins": joins, "origins": orig})


# ---------------------------------------------------------------- case 4
def water_ep(yinf: float = 0.0, alphae: float = 0.04, cpv: float = 0.0):
    """ep(1:8) of water vapour as Setup_Materials packs it (Mv, Lv, cpv, Le, Yinf,
    Lv Mv / Ru, 1/Tboil, alphaE), values of test/verification/F-evaporation/run.py"""
    Mv, Lv, Tb = 18.015, 2.26e6, 373.15
    return [Mv, Lv, cpv, 1.0, yinf, Lv * Mv / 8314.46, 1.0 / Tb, alphae]


def water_table(t_min: int = 250, t_max: int = 450):
    """A property table of liquid water on t_min..t_max K for ice_gpu_set_properties,
    every column varying: cp = 4180 + 0.8 (T - 300) with its exact enthalpy, density
    1000 - 0.25 (T - 280), Clausius-Clapeyron Psat at the nodes, Lv = 2.5e6 - 2400
    (T - 273.15), Cpv = 1850 + 0.3 (T - 300), emissivity 0.85 + 5e-4 (T - 250)."""
    from . import properties
    T = np.arange(t_min, t_max + 1, dtype=float)
    cp = 4180.0 + 0.8 * (T - 300.0)
    h = 4180.0 * T + 0.4 * (T - 300.0) ** 2 - 0.4 * 300.0 ** 2
    ep = water_ep()
    psat = 101325.0 * np.exp(-ep[5] * (1.0 / T - ep[6]))
    return properties.material_table(T, cp, 1000.0 - 0.25 * (T - 280.0), h, psat=psat,
                                      lv=2.5e6 - 2400.0 * (T - 273.15), cpv=1850.0 + 0.3 * (T - 300.0),
                                      emiss=0.85 + 5e-4 * (T - 250.0), evaporates=True)


def case_evaporation(n: int = 256, steps: int = 50, seed: int = 20261206, mode: int = 0, evap: int = 2,
                     intf: int = 1, blow: int = 1, box: float = 25.6e-3, limiter: int = 3,
                     table: bool = False) -> Case:
    """Evaporation-heavy: one n^3 block of side `box` (25.6 mm at n = 256), MK closure
    (Saurel flux), water drops in a frozen gas with seeded smooth random fields
    (16 modes each): T 300-2000 K, p 1-5 bar (rho_g = p / (R T)), far-field vapour
    mass fraction 0-0.3 per cell, velocity |u| <= 10 m/s.  Drops: radius lognormal
    (median 20 um, sigma 0.6, smooth), clipped to 1-100 um; T_d 280-350 K; rho_p =
    1e-2 (1 + 0.3 noise); at rest.  Drag Schiller-Naumann, Ranz-Marshall heat,
    grey-body radiation (emissivity 0.9), evaporation `evap` (2 CEM, 4 ASM) with
    the Langmuir-Knudsen interface (intf 1) and Stefan blowing (blow 1).  SSP-RK3,
    time-accurate, transmissive (400) faces."""
    rng = np.random.default_rng(seed)
    dims = (n, n, n)
    dx = box / n
    geo = geometry.uniform_metrics(dims, dx)
    c = (np.arange(n) + 0.5) / n

    def unit(f):
        return (f - f.min()) / (f.max() - f.min())

    def field():
        return geometry.smooth_field_axes(rng, c, c, c)

    u, v, w = field(), field(), field()
    umax = np.sqrt(u * u + v * v + w * w).max()
    Tg = 300.0 + 1700.0 * unit(field())
    p = 1e5 + 4e5 * unit(field())
    yinf = 0.3 * unit(field())
    zr = field()
    r = np.clip(20e-6 * np.exp(0.6 * zr / zr.std()), 1e-6, 100e-6)
    Td = 280.0 + 70.0 * unit(field())
    noise = field()
    rho = 1e-2 * (1.0 + 0.3 * noise / np.abs(noise).max())
    R, gam = 287.05, 1.4
    g5 = np.stack([p / (R * Tg), 10.0 * u / umax, 10.0 * v / umax, 10.0 * w / umax, Tg])
    mu = 1.8e-5 * (Tg / 300.0) ** 0.7
    kg = 0.026 * (Tg / 300.0) ** 0.7
    gas = gas_arrays(g5, R, gam, kg, mu)
    rho_mat = 1000.0
    num = rho / (rho_mat * 4.0 / 3.0 * PI * r ** 3)

    def fill(I, J, K):
        z = np.zeros(I.shape)
        return np.stack([rho, z, z, z, Td, num])

    prim = padded_state(dims, 6, fill)
    blk = Block(dims, geo, [prim], gas, yinf=np.ascontiguousarray(yinf.ravel(order="F")))
    tb = TableBuilder(1)
    tb.add_block(1, dims, lambda f, i, j, k: np.full(len(f), 400, np.int32))
    reci, recd, donor, weight = tb.arrays()
    cfg = make_config(ngroups=1, nrk=3, time_accurate=True, ndir=3, limiter=limiter, coupled=True,
                      drag_select=4, heat_select=6, blow_select=blow, var_yinf=True, cfl=0.5, dt_max=2e-6,
                      tau_factor=0.5, mode=mode)
    fam = make_family(MK, SAUREL, rho_mat=rho_mat, cs=4180.0, lv=2.26e6, emiss=0.9, evap=evap, intf=intf,
                      ep=water_ep())
    meta = {"tables": {1: water_table()}} if table else {}
    return Case("evaporation", cfg, [fam], [blk], reci, recd, donor, weight, steps, meta=meta)


# ---------------------------------------------------------------- chimera (overset) pair
def case_chimera(n: int = 12, na: int = 16, shift: float = 3.7, closure: int = IG, ngroups: int = 1,
                 irs: bool = False, evap: int = 0, steps: int = 10, seed: int = 20261207, mode: int = 0,
                 limiter: int = 3, connect: bool = False, closed_form: bool = False) -> Case:
    """Two overlapping blocks joined by chimera (102) records: A is na x n x n cells,
    B (na x n x n) starts `shift` cells before A's upper i-end, so that the two
    ghost layers of B's face 1 lie inside A and those of A's face 2 inside B.  Each
    ghost takes the conservative state of the two A (or B) cells bracketing its
    centre along i, with linear weights summing to 1 per layer (the donor lists of
    IO_BC.f90:209-235).  Other faces extrapolate (400).  A smooth particle field
    moves with +x velocity through the overlap; the gas is a smooth random field.
    `closure` IG or AG, `ngroups` families, IRS (steady) and evaporation on demand.
    With shift = 0 the blocks abut; `connect` then joins them by face-matching
    connections (101) instead, the same interface by another route.  closed_form:
    the blocks' metrics in closed form (geometry.uniform_metrics), for large pairs."""
    rng = np.random.default_rng(seed)
    dx = 1e-3
    dims = (na, n, n)
    x0 = [0.0, (na - shift) * dx]
    gseed = int(rng.integers(1 << 30))
    blocks = []
    rho_mat = 1000.0
    nc = NCOND[closure]
    ext = (x0[1] + na * dx)
    for b in range(2):
        if closed_form:
            geo = geometry.uniform_metrics(dims, dx)
        else:
            xs = [x0[b] + np.arange(na + 1) * dx, np.arange(n + 1) * dx, np.arange(n + 1) * dx]
            geo = geometry.metrics(geometry.cartesian_nodes(*xs))
        xc = [x0[b] + (np.arange(na) + 0.5) * dx, (np.arange(n) + 0.5) * dx, (np.arange(n) + 0.5) * dx]
        X, Y, Z = np.meshgrid(xc[0] / ext, xc[1] / (n * dx), xc[2] / (n * dx), indexing="ij")
        g_rng = np.random.default_rng(gseed)
        u, v, w, t = (smooth_field(g_rng, X, Y, Z, modes=6, kmax=2) for _ in range(4))
        g5 = np.stack([np.full(X.shape, 1.2), 8.0 + u, v, w, 900.0 + 100.0 * t])
        gas = gas_arrays(g5, 287.05, 1.4, 0.026, 1.8e-5)
        prims = []
        for p in range(ngroups):
            r0 = [20e-6, 50e-6][p % 2]

            def fill(I, J, K, X=X, Y=Y, Z=Z, r0=r0):
                rho = 1e-3 * (1.0 + 0.3 * np.sin(2 * PI * X) * np.cos(2 * PI * Y) * np.cos(2 * PI * Z))
                num = rho / (rho_mat * 4.0 / 3.0 * PI * r0 ** 3)
                z = np.zeros(I.shape)
                uu, vv = 5.0 + 0.5 * np.sin(2 * PI * Y), 0.5 * np.cos(2 * PI * X)
                if closure == AG:
                    th = rho * 1.0
                    return np.stack([rho, uu, vv, z, th, z, z, th, z, th, z + 330.0, num])
                return np.stack([rho, uu, vv, z, rho * 1.0, z + 330.0, num])

            prims.append(padded_state(dims, nc, fill))
        blocks.append(Block(dims, geo, prims, gas,
                            yinf=np.zeros(dims[0] * dims[1] * dims[2]) if evap else None))
    tb = TableBuilder(ngroups)

    def donors(other, xg):
        """(cells, weights) of block `other` bracketing centre x = xg along i"""
        s = (xg - x0[other]) / dx + 0.5  # fractional 1-based index
        i0 = int(np.floor(s + 1e-9))
        f = s - i0
        if abs(f) < 1e-9:
            return [(i0, 1.0)]
        return [(i0, 1.0 - f), (i0 + 1, f)]

    for b in range(2):
        other = 1 - b
        face = 1 if b == 1 else 2
        for (f, i, j, k) in face_cell_array(dims):
            if f != face:
                tb.add(b + 1, int(f), int(i), int(j), int(k), 400)
                continue
            if connect:
                # the same interface as a face-matching connection (shift 0)
                src = 1 if face == 2 else na
                tb.add(b + 1, int(f), int(i), int(j), int(k), 101,
                       ints={R_BS: other + 1, R_IS: src, R_JS: int(j), R_KS: int(k), R_FS: 3 - face})
                continue
            d0 = len(tb.weight)
            nchi = []
            for g in (1, 2):
                ig = i - g if face == 1 else i + g
                xg = x0[b] + (ig - 0.5) * dx
                dd = donors(other, xg)
                nchi.append(len(dd))
                for (io, wgt) in dd:
                    assert 1 <= io <= na
                    tb.donor.extend([other + 1, io, int(j), int(k)])
                    tb.weight.append(wgt)
            tb.add(b + 1, int(f), int(i), int(j), int(k), 102,
                   ints={R_NCHI1: nchi[0], R_NCHI2: nchi[1], R_DONOR0: d0})
    reci, recd, donor, weight = tb.arrays()
    cfg = make_config(ngroups=ngroups, nrk=3, time_accurate=not irs, ndir=3, limiter=limiter, coupled=True,
                      drag_select=4, heat_select=6, blow_select=1 if evap else 0, var_yinf=bool(evap), irs=irs,
                      cfl=0.5, dt_max=1e-4, tau_factor=0.5, mode=mode)
    fams = [make_family(closure, HLLE if closure == AG else RUSANOV, rho_mat=rho_mat, cs=4180.0, lv=2.26e6,
                        emiss=0.9 if evap else 0.0, evap=evap, intf=1 if evap else 0, ep=water_ep(alphae=0.04))
            for _ in range(ngroups)]
    return Case("chimera", cfg, fams, blocks, reci, recd, donor, weight, steps, meta={"x0": x0})


# ---------------------------------------------------------------- cases from ICE's input files
_LIMITER_WORDS = {"none": 0, "minmod": 1, "vanalbada": 2, "vanleer": 3, "ospre": 4, "umist": 5, "osher": 6,
                  "sweby": 7, "mc": 8, "koren": 9, "superbee": 10}
_CLOSURE_WORDS = {"MK": MK, "IG": IG, "AG": AG}
_RIEMANN_WORDS = {"saurel": SAUREL, "rusanov": RUSANOV, "hlle": HLLE}


def _ini_bool(v: str) -> bool:
    return v.strip().lower() in (".true.", "true", "t", "1", "yes")


def config_from_ini(ini: dict, ndir: int, mode: int = 0):
    """ice_config_t and the families of an uncoupled run from the INI sections
    [ICE-Numerics] and [ICE-FamilyN] (defaults of Register_Numerics.f90:28-80)."""
    num = ini.get("ICE-Numerics", {})
    phys = ini.get("ICE-Physics", {})
    if phys.get("drag", "NoDrag") != "NoDrag" or phys.get("heat-transfer", "NoHeat") != "NoHeat":
        raise ValueError("a coupled run needs its gas fields: build it with the case builders")
    scheme = num.get("time-scheme", "RK2")
    nrk = {"euler": 1, "RK2": 2, "RK3": 3}[scheme]
    recon = num.get("space-reconstruction", "first-order")
    lim = _LIMITER_WORDS[num.get("flux-limiter", "none").lower()] if recon == "MUSCL" else 0
    fams = []
    p = 1
    while f"ICE-Family{p}" in ini:
        fams.append(_CLOSURE_WORDS[ini[f"ICE-Family{p}"].get("closure", "IG")])
        p += 1
    cfg = make_config(ngroups=len(fams), nrk=nrk, time_accurate=_ini_bool(num.get("time-accurate", ".true.")),
                      cfl_ramp_iter=int(num.get("cfl-rise-threshold", "0")), ndir=ndir, limiter=lim,
                      shock_detector=num.get("shock-detector", "none").lower() == "jameson",
                      irs=_ini_bool(num.get("irs", ".false.")), coupled=False, mode=mode,
                      cfl=float(num.get("cfl", "0.5")), dt_max=float(num.get("dt-max", "1e-4")),
                      tau_factor=float(num.get("tau-factor", "1.0")), irs_beta=float(num.get("irs-beta", "0.5")))
    rs = num.get("riemann-solver", "").strip().lower()
    return cfg, [(c, _RIEMANN_WORDS[rs] if rs else None) for c in fams]


def case_from_inputs(name: str, ini: dict, nodes: List[np.ndarray], ic: List[np.ndarray], reci, recd, donor, weight,
                     rho_mat: float, cs: float, steps: int = 1, mode: int = 0) -> Case:
    """A Case from what ICE's set-up reads: the INI sections, each block's nodes
    (3, ni+1, nj+1, nk+1), its initial primitive cells (nc, ni, nj, nk) of the one
    family, the boundary table, and the material's density and heat capacity."""
    geos = [geometry.metrics(n) for n in nodes]
    dims = [g["dims"] for g in geos]
    ndir = sum(1 for d in range(3) if max(dm[d] for dm in dims) > 1)
    cfg, fspec = config_from_ini(ini, ndir, mode)
    if len(fspec) != 1:
        raise ValueError("the initial condition holds one family")
    closure, riemann = fspec[0]
    fam = make_family(closure, riemann, rho_mat=rho_mat, cs=cs)
    nc = NCOND[closure]
    blocks = []
    for g, c in zip(geos, ic):
        if c.shape[0] != nc:
            raise ValueError(f"initial condition with {c.shape[0]} variables for a closure of {nc}")
        prim = padded_state(g["dims"], nc, lambda I, J, K, c=c: c[:, I - 1, J - 1, K - 1])
        blocks.append(Block(g["dims"], g, [prim]))
    return Case(name, cfg, [fam], blocks, reci, recd, donor, weight, steps)


# ---------------------------------------------------------------- combined configuration
# The seed of the full combined configuration's tree: the first seed from 20261208
# up whose 88-block tree (edges log-uniform on [32, 192]) holds 140-160 M cells.
# Such trees are rare: the median tree of this generator holds 64 M cells, and 1
# seed in 400 reaches 140 M; this one holds 142.6 M.
COMBINED_FULL_SEED = 20261309


def tree_codes(partner_face, b):
    """code_fn of block b of a connection tree: 101 on joined faces, 400 elsewhere."""
    joined = np.array([(b, f) in partner_face for f in range(0, 7)])

    def codes(f, i, j, k):
        return np.where(joined[np.asarray(f)], 101, 400).astype(np.int32)
    return codes


def tree_ints(partner_face, dims, b):
    """ints_fn of block b of a connection tree: the source cell across each joined face."""
    def ints(ri, fijk):
        for f in range(1, 7):
            if (b, f) not in partner_face:
                continue
            ob, of = partner_face[(b, f)]
            m = fijk[:, 0] == f
            d = (f - 1) // 2
            src = fijk[m, 1:4].copy()
            # the cell across the face: same tangential indices, first interior cell of the other block
            src[:, d] = 1 if of % 2 == 1 else dims[ob][d]
            ri[m, R_BS] = ob + 1
            ri[m, R_IS], ri[m, R_JS], ri[m, R_KS] = src[:, 0], src[:, 1], src[:, 2]
            ri[m, R_FS] = of
    return ints


def case_combined(ntree: int = 88, npairs: int = 4, emin: int = 32, emax: int = 192, pair_n: int = 24,
                  seed: int = 20261208, steps: int = 10, mode: int = 0, dx: float = 1e-4, lazy: bool = False,
                  pair_closed_form: bool = False) -> Case:
    """Combined: a seeded random tree of `ntree` face-matching blocks (edges
    log-uniform on [emin, emax], connections 101) plus `npairs` chimera overset
    block pairs (102 donor lists with linear weights, as case_chimera, blocks of
    2 pair_n x pair_n x pair_n cells), AG closure with 2 families (HLLE, MUSCL van
    Leer), full interphase physics: drag Schiller-Naumann, Ranz-Marshall heat,
    grey-body radiation, CEM evaporation with the Langmuir-Knudsen interface and
    Stefan blowing.  Gas fields as the evaporation case: seeded smooth random T
    300-2000 K, p 1-5 bar, far-field vapour fraction 0-0.3, |u| <= 10 m/s.  Steady
    SSP-RK3 with implicit residual smoothing (beta 0.5).  Uniform Cartesian cells
    of side dx (closed-form metrics).  case_combined_full() is the full
    configuration; smaller arguments give the same physics at test sizes.
    lazy: the tree's blocks are LazyBlocks, made one at a time as a Runner uploads
    them (the full configuration's arrays would hold some 70 GB of host memory).
    pair_closed_form: the chimera pairs' metrics in closed form (large pairs)."""
    rng = np.random.default_rng(seed)
    dims, orig, joins = block_tree(ntree, emin, emax, rng)
    rho_mat, ng = 1000.0, 2
    lo = np.min(np.asarray(orig), axis=0)
    hi = np.max(np.asarray(orig) + np.asarray(dims), axis=0)
    span = (hi - lo).astype(float)
    fseed = int(rng.integers(1 << 30))
    R, gam = 287.05, 1.4
    medians = [20e-6, 60e-6]

    def make_block(dd, o):
        geo = geometry.uniform_metrics(dd, dx)
        c = [((o[m] - lo[m]) + np.arange(dd[m]) + 0.5) / span[m] for m in range(3)]
        frng = np.random.default_rng(fseed)   # the same fields everywhere, sampled at the block's cells

        def field():
            f = geometry.smooth_field_product(frng, c[0], c[1], c[2], modes=8, kmax=3)
            return 0.5 + 0.5 * np.tanh(f / 2.0)   # in (0, 1)

        u, v, w = (2.0 * field() - 1.0 for _ in range(3))
        Tg = 300.0 + 1700.0 * field()
        p = 1e5 + 4e5 * field()
        yinf = 0.3 * field()
        Td = 280.0 + 70.0 * field()
        rr = field()
        g5 = np.stack([p / (R * Tg), 10.0 * u / np.sqrt(3.0), 10.0 * v / np.sqrt(3.0), 10.0 * w / np.sqrt(3.0), Tg])
        del u, v, w, p
        gas = gas_arrays(g5, R, gam, 0.026 * (Tg / 300.0) ** 0.7, 1.8e-5 * (Tg / 300.0) ** 0.7)
        prims = []
        rho = 1e-2 * (0.7 + 0.6 * rr)
        for q in range(ng):
            r = np.clip(medians[q] * np.exp(0.6 * (2.0 * rr - 1.0) * 2.0), 1e-6, 100e-6)
            num = rho / (rho_mat * 4.0 / 3.0 * PI * r ** 3)
            prims.append(padded_fields(dd, [rho, 0.9 * g5[1], 0.9 * g5[2], 0.9 * g5[3], rho, 0.0, 0.0, rho, 0.0, rho,
                                            Td, num]))
        return Block(dd, geo, prims, gas, yinf=np.ascontiguousarray(yinf.ravel(order="F")))

    blocks = []
    for dd, o in zip(dims, orig):
        blocks.append(LazyBlock(dd, lambda dd=dd, o=o: make_block(dd, o)) if lazy else make_block(dd, o))
    partner_face = {}
    for a, fa, b, fb in joins:
        partner_face[(a, fa)] = (b, fb)
        partner_face[(b, fb)] = (a, fa)
    tb = TableBuilder(ng)
    for b, dd in enumerate(dims):
        tb.add_block(b + 1, dd, tree_codes(partner_face, b), ints_fn=tree_ints(partner_face, dims, b))
    reci, recd, donor, weight = tb.arrays()
    # the chimera pairs, appended block by block after the tree
    ri_all, rd_all, dn_all, w_all = [reci], [recd], [donor], [weight]
    nb = len(blocks)
    nd = donor.shape[0]
    for q in range(npairs):
        pc = case_chimera(n=pair_n, na=2 * pair_n, shift=3.7, closure=AG, ngroups=ng, irs=True, evap=2,
                          seed=seed + 1 + q, mode=mode, closed_form=pair_closed_form)
        ri = pc.reci.copy()
        ri[:, R_B] += nb
        has = ri[:, R_BS] > 0
        ri[has, R_BS] += nb
        chim = ri[:, R_TYPE] == 102
        ri[chim, R_DONOR0] += nd
        dn = pc.donor.copy()
        dn[:, 0] += nb
        ri_all.append(ri)
        rd_all.append(pc.recd)
        dn_all.append(dn)
        w_all.append(pc.weight)
        blocks.extend(pc.blocks)
        nb += len(pc.blocks)
        nd += dn.shape[0]
    reci = np.ascontiguousarray(np.concatenate(ri_all))
    recd = np.ascontiguousarray(np.concatenate(rd_all))
    donor = np.ascontiguousarray(np.concatenate(dn_all).astype(np.int32))
    weight = np.ascontiguousarray(np.concatenate(w_all))
    cfg = make_config(ngroups=ng, nrk=3, time_accurate=False, ndir=3, limiter=3, coupled=True, drag_select=4,
                      heat_select=6, blow_select=1, var_yinf=True, irs=True, irs_beta=0.5, cfl=0.5, dt_max=2e-6,
                      tau_factor=0.5, mode=mode)
    fams = [make_family(AG, HLLE, rho_mat=rho_mat, cs=4180.0, lv=2.26e6, emiss=0.9, evap=2, intf=1,
                        ep=water_ep(alphae=0.04)) for _ in range(ng)]
    return Case("combined", cfg, fams, blocks, reci, recd, donor, weight, steps,
                meta={"joins": joins, "tree_blocks": ntree, "pairs": npairs})


def case_combined_full(mode: int = 0, steps: int = 10, lazy: bool = True) -> Case:
    """The full combined configuration: 96 blocks, 88 in the connection tree (edges
    log-uniform on [32, 192]) and 4 chimera pairs of 128 x 64 x 64 blocks, 146.8 M
    cells in all, AG x 2 families with the full interphase physics, steady SSP-RK3
    with IRS."""
    return case_combined(ntree=88, npairs=4, emin=32, emax=192, pair_n=64, seed=COMBINED_FULL_SEED, steps=steps,
                         mode=mode, lazy=lazy, pair_closed_form=True)
