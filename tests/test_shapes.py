"""Host-side Shapes.orient (js/shapes.js:212-247) against the reference's
oriented mesh, and the rotation it applies for other axis choices."""
import numpy as np

from _util import load
from windtunnel_amd import synth
from windtunnel_amd.shapes import orient


def test_orient_reproduces_reference_bytes():
    fx = load("lbm_c2")["mesh"]
    P = synth.ellipsoid_box()
    assert synth.sha256(P) == fx["sha256"]
    k = orient(P, fx["addModel"]["forward"], fx["addModel"]["up"], fx["addModel"]["flip"], fx["addModel"]["ground"])
    assert k == fx["k"]
    assert synth.sha256(P) == fx["oriented_sha256"]


def _expected(V, forward, up, flip, ground):
    """Rows of the rotation: nose (-f), up (u), and f x u, spelled out per case."""
    a, b, c = V[:, 0].astype(np.float64), V[:, 1].astype(np.float64), V[:, 2].astype(np.float64)
    cols = {0: a, 1: b, 2: c}
    sgn = -1.0 if flip else 1.0
    x = -sgn * cols[forward]
    y = cols[up]
    third = 3 - forward - up
    # f x u with f = sgn e_forward, u = e_up is sgn * (e_forward x e_up) = sgn * eps * e_third;
    # the third row is its negative applied to the vertex
    eps = 1.0 if (forward, up) in ((0, 1), (1, 2), (2, 0)) else -1.0
    z = -sgn * eps * cols[third]
    x0, x1, y0, y1, z0, z1 = x.min(), x.max(), y.min(), y.max(), z.min(), z.max()
    k = 1 / max(1e-9, x1 - x0)
    cy = y0 if ground else (y0 + y1) / 2
    return np.stack([(x - (x0 + x1) / 2) * k, (y - cy) * k, (z - (z0 + z1) / 2) * k], 1), k


def test_orient_axes_and_normalisation():
    rng = np.random.default_rng(3)
    base = (rng.normal(size=(300, 9)) * [3, 1, 2, 3, 1, 2, 3, 1, 2]).astype(np.float32).reshape(-1)
    for forward in range(3):
        for up in range(3):
            if forward == up:
                continue
            for flip in (False, True):
                for ground in (False, True):
                    P = base.copy()
                    k = orient(P, forward, up, flip, ground)
                    want, k2 = _expected(base.reshape(-1, 3), forward, up, flip, ground)
                    assert abs(k - k2) <= 1e-12 * k2
                    np.testing.assert_allclose(P.reshape(-1, 3), want, rtol=0, atol=2e-7)
                    V = P.reshape(-1, 3).astype(np.float64)
                    assert abs((V[:, 0].max() - V[:, 0].min()) - 1) < 1e-6
                    if ground:
                        assert abs(V[:, 1].min()) < 1e-7


def test_orient_same_axes_falls_back():
    """forward == up picks the next axis after up (js/shapes.js:215)."""
    rng = np.random.default_rng(4)
    base = rng.normal(size=90).astype(np.float32)
    P, Q = base.copy(), base.copy()
    assert orient(P, 1, 1, False, False) == orient(Q, 2, 1, False, False)
    assert np.array_equal(P, Q)
