"""Implicit residual smoothing of the oracle (Lib_IRS.f90:15-106) pinned by the
closed form of its two Jacobi sweeps per direction (no GPU).

With R* fixed to zero-gradient copies in the ghosts, one sweep is
Rnew = (R + b (R*[m-1] + R*[m+1])) / (1 + 2b), and R on the right-hand side is
always the pre-pass residual.  For a unit residual at cell m of a line, far
from its ends, two sweeps give 1/(1+2b) + 2b^2/(1+2b)^2 at m, b/(1+2b)^2 next to it and
b^2/(1+2b)^2 two cells away; at a line's end the ghost keeps the pre-pass
value through both sweeps; a constant field is kept; the operator is linear
and, per direction, acts along that direction only."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


@pytest.fixture(scope="module")
def irs():
    lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    f = lib.orc_fn_irs
    f.argtypes = [C.c_int] * 5 + [C.c_double, P]
    f.restype = None

    def run(R, ndir, beta):
        nk, nj, ni, nc = R.shape
        out = np.ascontiguousarray(R.copy())
        f(ni, nj, nk, nc, ndir, beta, out)
        return out
    return run


@pytest.mark.parametrize("beta", [0.5, 0.3])
def test_delta_response_1d(irs, beta):
    n, m = 15, 7
    R = np.zeros((1, 1, n, 1))
    R[0, 0, m, 0] = 1.0
    out = irs(R, 1, beta)[0, 0, :, 0]
    s = 1.0 + 2.0 * beta
    # sweep 1 from R* = R: a1[m] = 1/s, a1[m+-1] = b/s; sweep 2 from R* = a1
    want = np.zeros(n)
    want[m] = (1.0 + beta * (beta / s + beta / s)) / s
    want[m - 1] = want[m + 1] = beta * (1.0 / s) / s
    want[m - 2] = want[m + 2] = beta * (beta / s) / s
    assert np.allclose(out, want, rtol=1e-15, atol=1e-300)
    # the textbook form of the same numbers
    assert np.isclose(out[m], 1 / s + 2 * beta ** 2 / s ** 2, rtol=1e-15)
    assert np.isclose(out[m + 1], beta / s ** 2, rtol=1e-15)
    assert np.isclose(out[m + 2], beta ** 2 / s ** 2, rtol=1e-15)


def test_constant_kept_and_ends_neumann(irs):
    R = np.full((3, 4, 5, 2), 2.5)
    assert np.allclose(irs(R, 3, 0.5), 2.5, rtol=2e-16, atol=0)
    # a unit residual in the end cell of a line: the ghost copies it (zero gradient)
    n, b = 6, 0.5
    R = np.zeros((1, 1, n, 1))
    R[0, 0, 0, 0] = 1.0
    out = irs(R, 1, b)[0, 0, :, 0]
    s = 1 + 2 * b
    a1 = np.array([(1 + b * (1 + 0)) / s, b / s] + [0.0] * (n - 2))
    want0 = (1 + b * (1.0 + a1[1])) / s  # the ghost keeps the pre-pass value through both sweeps
    assert np.isclose(out[0], want0, rtol=1e-15)


def test_linear_and_separable(irs):
    rng = np.random.default_rng(5)
    A, B = rng.normal(size=(2, 4, 5, 6, 3))
    lhs = irs(np.ascontiguousarray(2.0 * A - 3.0 * B), 3, 0.5)
    rhs = 2.0 * irs(A, 3, 0.5) - 3.0 * irs(B, 3, 0.5)
    assert np.allclose(lhs, rhs, rtol=1e-13, atol=1e-13)
    # with ndir = 1 each i-line is smoothed on its own: lines do not mix
    one = irs(A, 1, 0.5)
    line = irs(np.ascontiguousarray(A[1:2, 2:3]), 1, 0.5)
    assert np.array_equal(one[1:2, 2:3], line)
