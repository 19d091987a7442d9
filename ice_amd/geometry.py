"""Block geometry from its nodes: ICE's setup_metrics (src/lib/numerics/space/
Mod_Metrics.f90:19-59) restated with numpy, operation for operation, so that a
mesh built here gives the kernels the same metrics the reference computes.

Nodes are given over (0:ni, 0:nj, 0:nk); the ghost layer -1 / n+1 is
extrapolated linearly (extrapolate_nodes, :460-495).  Outputs are flat arrays
in the Fortran layouts of include/ice_gpu.h (component index fastest).
"""
from __future__ import annotations

import numpy as np


def _f(a: np.ndarray) -> np.ndarray:
    """Flatten in Fortran (column-major) order."""
    return np.ascontiguousarray(a.ravel(order="F"))


def extrapolate(nodes: np.ndarray) -> np.ndarray:
    """(3, ni+1, nj+1, nk+1) nodes -> (3, ni+3, nj+3, nk+3), index -1 at 0 (Mod_Metrics.f90:460-495)."""
    _, a, b, c = nodes.shape
    im, jm, km = a - 1, b - 1, c - 1
    n = np.full((3, im + 3, jm + 3, km + 3), np.nan)
    n[:, 1:im + 2, 1:jm + 2, 1:km + 2] = nodes
    o = 1  # offset of index 0
    # i-faces, k = 0..km, j = 0..jm
    J, K = slice(o, jm + o + 1), slice(o, km + o + 1)
    n[:, 0, J, K] = 2.0 * n[:, o, J, K] - n[:, o + 1, J, K]
    n[:, im + o + 1, J, K] = 2.0 * n[:, im + o, J, K] - n[:, im + o - 1, J, K]
    # j-faces, k = 0..km, i = -1..im+1
    n[:, :, 0, o:km + o + 1] = 2.0 * n[:, :, o, o:km + o + 1] - n[:, :, o + 1, o:km + o + 1]
    n[:, :, jm + o + 1, o:km + o + 1] = 2.0 * n[:, :, jm + o, o:km + o + 1] - n[:, :, jm + o - 1, o:km + o + 1]
    # k-faces, all i, j
    n[:, :, :, 0] = 2.0 * n[:, :, :, o] - n[:, :, :, o + 1]
    n[:, :, :, km + o + 1] = 2.0 * n[:, :, :, km + o] - n[:, :, :, km + o - 1]
    return n


def _cross(u, v):
    return np.stack([u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]])


def _dot(u, v):
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]


def normal_signs(n: np.ndarray):
    """normal_signs (Mod_Metrics.f90:500-545) from the first cell; n has index -1 at 0."""
    N = lambda i, j, k: n[:, i + 1, j + 1, k + 1]  # noqa: E731
    out = []
    # i direction
    d1 = N(1, 1, 1) - N(1, 0, 0)
    d2 = N(1, 0, 1) - N(1, 1, 0)
    d3 = _cross(d1, d2)
    c1 = 0.25 * (N(1, 1, 1) + N(1, 0, 1) + N(1, 0, 0) + N(1, 1, 0))
    c2 = 0.25 * (N(0, 1, 1) + N(0, 0, 1) + N(0, 0, 0) + N(0, 1, 0))
    out.append(np.copysign(1.0, (c1[0] - c2[0]) * d3[0] + (c1[1] - c2[1]) * d3[1] + (c1[2] - c2[2]) * d3[2]))
    # j direction
    d1 = N(0, 1, 1) - N(1, 1, 0)
    d2 = N(1, 1, 1) - N(0, 1, 0)
    d3 = _cross(d1, d2)
    c1 = 0.25 * (N(1, 1, 1) + N(0, 1, 1) + N(0, 1, 0) + N(1, 1, 0))
    c2 = 0.25 * (N(1, 0, 1) + N(0, 0, 1) + N(0, 0, 0) + N(1, 0, 0))
    out.append(np.copysign(1.0, (c1[0] - c2[0]) * d3[0] + (c1[1] - c2[1]) * d3[1] + (c1[2] - c2[2]) * d3[2]))
    # k direction
    d1 = N(0, 1, 1) - N(1, 0, 1)
    d2 = N(0, 0, 1) - N(1, 1, 1)
    d3 = _cross(d1, d2)
    c1 = 0.25 * (N(1, 1, 1) + N(0, 1, 1) + N(0, 0, 1) + N(1, 0, 1))
    c2 = 0.25 * (N(1, 1, 0) + N(0, 1, 0) + N(0, 0, 0) + N(1, 0, 0))
    out.append(np.copysign(1.0, (c1[0] - c2[0]) * d3[0] + (c1[1] - c2[1]) * d3[1] + (c1[2] - c2[2]) * d3[2]))
    return out


def metrics(nodes: np.ndarray) -> dict:
    """vol, dl, M, face_i, face_j, face_k of one block (compute_metrics :548-690,
    compute_metric_tensor :700-779), as flat Fortran-layout float64 arrays."""
    n = extrapolate(np.asarray(nodes, dtype=np.float64))
    im, jm, km = n.shape[1] - 3, n.shape[2] - 3, n.shape[3] - 3
    o = 1
    signs = normal_signs(n)

    def nd(i0, i1, j0, j1, k0, k1):
        """node slab over Fortran index ranges (inclusive)"""
        return n[:, i0 + o:i1 + o + 1, j0 + o:j1 + o + 1, k0 + o:k1 + o + 1]

    def face(d1, d2, sign):
        d3 = _cross(d1, d2) * 0.5
        A = np.sqrt(d3[0] * d3[0] + d3[1] * d3[1] + d3[2] * d3[2])
        with np.errstate(invalid="ignore", divide="ignore"):
            N = np.where(A == 0.0, 0.0, d3 / A * sign)
        return np.concatenate([N, A[None]], axis=0)

    # i faces: k = 1..km, j = 1..jm, i = 0..im
    fi = face(nd(0, im, 1, jm, 1, km) - nd(0, im, 0, jm - 1, 0, km - 1),
              nd(0, im, 0, jm - 1, 1, km) - nd(0, im, 1, jm, 0, km - 1), signs[0])
    # j faces: k = 1..km, j = 0..jm, i = 1..im
    fj = face(nd(0, im - 1, 0, jm, 1, km) - nd(1, im, 0, jm, 0, km - 1),
              nd(1, im, 0, jm, 1, km) - nd(0, im - 1, 0, jm, 0, km - 1), signs[1])
    # k faces: k = 0..km, j = 1..jm, i = 1..im
    fk = face(nd(0, im - 1, 1, jm, 0, km) - nd(1, im, 0, jm - 1, 0, km),
              nd(0, im - 1, 0, jm - 1, 0, km) - nd(1, im, 1, jm, 0, km), signs[2])

    # volumes: corners relative to the first one, 1 + a + 2b + 4c = node (i-1+a, j-1+b, k-1+c)
    c0 = nd(0, im - 1, 0, jm - 1, 0, km - 1)
    r = [None, np.zeros_like(c0)]
    for m in range(1, 8):
        a, b, c = m & 1, (m >> 1) & 1, (m >> 2) & 1
        r.append(nd(a, im - 1 + a, b, jm - 1 + b, c, km - 1 + c) - c0)
    ea1 = 0.25 * (r[2] + (r[6] - r[5]) + (r[4] - r[3]) + (r[8] - r[7]))
    ea2 = 0.25 * (r[3] + (r[7] - r[5]) + (r[4] - r[2]) + (r[8] - r[6]))
    ea3 = 0.25 * (r[5] + (r[7] - r[3]) + (r[6] - r[2]) + (r[8] - r[4]))
    vol = _dot(ea1, _cross(r[7], r[5] - r[3]) + _cross(r[8] - r[2], r[6] - r[4]))
    vol = vol + _dot(ea2, _cross(r[6], r[2] - r[5]) + _cross(r[8] - r[3], r[4] - r[7]))
    vol = vol + _dot(ea3, _cross(r[4], r[3] - r[2]) + _cross(r[8] - r[5], r[7] - r[6]))
    vol = vol / 12.0
    det_e = _dot(ea1, _cross(ea2, ea3))
    if not np.all(vol * det_e > 0.0):
        raise ValueError("[ICE::compute_metrics] a cell whose volume is zero or has not the sign of det at its centre")
    vol = np.abs(vol)

    # metric tensor and lengths over 0..n+1
    P = lambda a, b, c: nd(a - 1, im + a, b - 1, jm + b, c - 1, km + c)  # noqa: E731  node (i-1+a, ...)
    A1 = P(1, 1, 1) - P(0, 1, 1) + P(1, 0, 1) - P(0, 0, 1) + P(1, 1, 0) - P(0, 1, 0) + P(1, 0, 0) - P(0, 0, 0)
    A2 = P(1, 1, 1) - P(1, 0, 1) + P(0, 1, 1) - P(0, 0, 1) + P(1, 1, 0) - P(1, 0, 0) + P(0, 1, 0) - P(0, 0, 0)
    A3 = P(1, 1, 1) - P(1, 1, 0) + P(0, 1, 1) - P(0, 1, 0) + P(1, 0, 1) - P(1, 0, 0) + P(0, 0, 1) - P(0, 0, 0)
    A = np.stack([0.25 * A1, 0.25 * A2, 0.25 * A3])  # A[row, comp, ...]
    det = (A[0, 0] * A[1, 1] * A[2, 2] - A[0, 0] * A[1, 2] * A[2, 1]
           - A[0, 1] * A[1, 0] * A[2, 2] + A[0, 1] * A[1, 2] * A[2, 0]
           + A[0, 2] * A[1, 0] * A[2, 1] - A[0, 2] * A[1, 1] * A[2, 0])
    cof = np.empty_like(A)
    cof[0, 0] = (A[1, 1] * A[2, 2] - A[1, 2] * A[2, 1])
    cof[0, 1] = -(A[1, 0] * A[2, 2] - A[1, 2] * A[2, 0])
    cof[0, 2] = (A[1, 0] * A[2, 1] - A[1, 1] * A[2, 0])
    cof[1, 0] = -(A[0, 1] * A[2, 2] - A[0, 2] * A[2, 1])
    cof[1, 1] = (A[0, 0] * A[2, 2] - A[0, 2] * A[2, 0])
    cof[1, 2] = -(A[0, 0] * A[2, 1] - A[0, 1] * A[2, 0])
    cof[2, 0] = (A[0, 1] * A[1, 2] - A[0, 2] * A[1, 1])
    cof[2, 1] = -(A[0, 0] * A[1, 2] - A[0, 2] * A[1, 0])
    cof[2, 2] = (A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0])
    with np.errstate(invalid="ignore", divide="ignore"):
        M = np.where(np.abs(det) == 0.0, np.eye(3)[:, :, None, None, None], cof / det)
    dl = np.sqrt(A[:, 0] * A[:, 0] + A[:, 1] * A[:, 1] + A[:, 2] * A[:, 2])  # (3, ni+2, nj+2, nk+2)
    Mi = M[:, :, 1:im + 1, 1:jm + 1, 1:km + 1]  # interior (3, 3, ni, nj, nk)
    return {
        "dims": (im, jm, km),
        "vol": _f(vol),
        "dl": _f(dl),
        "M": _f(Mi),
        "face_i": _f(fi),
        "face_j": _f(fj),
        "face_k": _f(fk),
    }


def cartesian_nodes(x: np.ndarray, y: np.ndarray, z: np.ndarray) -> np.ndarray:
    """(3, nx, ny, nz) node coordinates of a Cartesian block from its node lines."""
    X, Y, Z = np.meshgrid(np.asarray(x, float), np.asarray(y, float), np.asarray(z, float), indexing="ij")
    return np.stack([X, Y, Z])


def uniform_metrics(dims, dx: float) -> dict:
    """The metrics() of a uniform Cartesian block of cubes of side dx, in closed form
    (equal to metrics() on such a grid to rounding), for blocks too large for the
    node-based construction: vol = dx^3, lengths dx, M = I/dx, unit normals along
    the axes with area dx^2.  Each array is flat in its Fortran layout, built by
    repeating the one cell's (or face's) entries."""
    ni, nj, nk = dims
    vol = np.full(ni * nj * nk, dx * dx * dx)
    dl = np.full(3 * (ni + 2) * (nj + 2) * (nk + 2), dx)
    M = np.tile(np.eye(3).ravel(order="F") / dx, ni * nj * nk)

    def face(d, n):
        e = np.zeros(4)
        e[d] = 1.0
        e[3] = dx * dx
        return np.tile(e, n)

    return {"dims": (ni, nj, nk), "vol": vol, "dl": dl, "M": M,
            "face_i": face(0, (ni + 1) * nj * nk), "face_j": face(1, ni * (nj + 1) * nk),
            "face_k": face(2, ni * nj * (nk + 1))}


def smooth_field_axes(rng, x, y, z, modes: int = 16, kmax: int = 4):
    """A sum of `modes` random sine modes (wave numbers 1..kmax per axis) on the
    tensor grid of the axes x, y, z, by the angle-addition form of the sine: the
    same draws as a pointwise evaluation, one complex product per cell and mode."""
    f = np.zeros((len(x), len(y), len(z)))
    for _ in range(modes):
        k = rng.integers(1, kmax + 1, size=3) * rng.choice([-1, 1], size=3)
        amp, ph = rng.normal(), rng.uniform(0, 2 * np.pi)
        ex = np.exp(1j * (2 * np.pi * k[0] * np.asarray(x) + ph))
        ey = np.exp(1j * 2 * np.pi * k[1] * np.asarray(y))
        ez = np.exp(1j * 2 * np.pi * k[2] * np.asarray(z))
        f += amp * (ex[:, None, None] * (ey[:, None] * ez[None, :])[None]).imag
    return f


def smooth_field_product(rng, x, y, z, modes: int = 16, kmax: int = 4):
    """smooth_field_axes with the same draws, summed over the modes as one matrix
    product: sin(a + b) = sin a cos b + cos a sin b splits each mode into an
    x-factor times a (y, z)-factor, so the field is a (ny nz, 2 modes) by
    (2 modes, nx) product.  The values equal smooth_field_axes' to rounding; the
    (nx, ny, nz) result is laid out in Fortran order (x fastest)."""
    x, y, z = (np.asarray(t, float) for t in (x, y, z))
    A = np.empty((2 * modes, len(x)))
    B = np.empty((len(z) * len(y), 2 * modes))
    for m in range(modes):
        k = rng.integers(1, kmax + 1, size=3) * rng.choice([-1, 1], size=3)
        amp, ph = rng.normal(), rng.uniform(0, 2 * np.pi)
        ax = 2 * np.pi * k[0] * x + ph
        ezy = np.exp(1j * 2 * np.pi * k[2] * z)[:, None] * np.exp(1j * 2 * np.pi * k[1] * y)[None, :]
        A[2 * m], A[2 * m + 1] = amp * np.sin(ax), amp * np.cos(ax)
        B[:, 2 * m], B[:, 2 * m + 1] = ezy.real.ravel(), ezy.imag.ravel()
    return (B @ A).reshape(len(z), len(y), len(x)).T
