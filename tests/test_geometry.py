"""Host-side geometry helpers (no GPU)."""
import numpy as np

from ice_amd import cases, geometry


def test_uniform_metrics_equal_node_metrics():
    dx = 1e-4
    a = geometry.metrics(geometry.cartesian_nodes(*[np.arange(n + 1) * dx for n in (5, 4, 3)]))
    b = geometry.uniform_metrics((5, 4, 3), dx)
    for k in ("vol", "dl", "M", "face_i", "face_j", "face_k"):
        assert a[k].shape == b[k].shape
        assert np.allclose(a[k], b[k], rtol=1e-14, atol=0), k


def test_smooth_field_axes_is_the_pointwise_field():
    c = (np.arange(8) + 0.5) / 8
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    f1 = cases.smooth_field(np.random.default_rng(3), X, Y, Z)
    f2 = geometry.smooth_field_axes(np.random.default_rng(3), c, c, c)
    assert np.allclose(f1, f2, rtol=0, atol=1e-12)
