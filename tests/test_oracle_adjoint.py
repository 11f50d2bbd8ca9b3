"""The oracle's gray bounce-back and FP64 discrete adjoint.

Pinned by mathematics, not by the reference (which has neither): weights of
exactly 0/1 must reproduce the plain solver bit for bit, and dJ/dw from the
adjoint must match central finite differences of the forward solver.
"""
import numpy as np
import pytest

from oracle.oracle import OracleLBM, lib

DIMS = (20, 12, 12)
C = np.array([7.0, 6.0, 6.0])


def gray_field(dims=DIMS, r_in=2.0, r_out=3.6):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt((x + 0.5 - C[0]) ** 2 + (y + 0.5 - C[1]) ** 2 + (z + 0.5 - C[2]) ** 2)
    w = np.clip((r_out - r) / (r_out - r_in), 0, 1)
    return w.reshape(-1).astype(np.float32)


def lattice(w, les, trt, prec=8):
    L = OracleLBM(*DIMS, prec=prec)
    L.set_params(tau=0.6, les=les, trt=trt, csm=0.14, u0=0.05, **{"lambda": 3 / 16})
    L.reset()
    solid = (w >= 1).reshape(DIMS[::-1])
    idx = np.argwhere(solid)
    lo = idx.min(0)
    hi = idx.max(0)
    box = [int(lo[2]), int(lo[1]), int(lo[0]), int(hi[2]), int(hi[1]), int(hi[0])]
    mask = solid[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1].astype(np.uint8).reshape(-1)
    L.voxelize_mask(box, mask)
    gw = np.where(w >= 1, 0, w).astype(np.float32)
    L.set_gray(gw)
    return L


def objective(forces, Fbar):
    return float(np.sum(forces * Fbar))


def test_zero_one_weights_reduce_to_reference():
    """A gray field holding only 0 keeps the plain solver's trajectory bit for bit."""
    w = (gray_field() >= 1).astype(np.float32)
    a = lattice(w, True, True, prec=4)
    b = OracleLBM(*DIMS)
    b.set_params(tau=0.6, les=True, trt=True, csm=0.14, u0=0.05, **{"lambda": 3 / 16})
    b.reset()
    solid = (w >= 1).reshape(DIMS[::-1])
    idx = np.argwhere(solid)
    lo, hi = idx.min(0), idx.max(0)
    b.voxelize_mask([lo[2], lo[1], lo[0], hi[2], hi[1], hi[0]],
                    solid[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1].astype(np.uint8).reshape(-1))
    fa, fb = a.step(20), b.step(20)
    assert np.array_equal(fa, fb)
    assert np.array_equal(a.array("f"), b.array("f"))


@pytest.mark.parametrize("les,trt", [(False, False), (False, True), (True, True)])
def test_adjoint_matches_central_differences(les, trt):
    N = 40
    w = gray_field()
    Fbar = np.zeros((N, 3))
    Fbar[-10:, 0] = 0.1
    Fbar[-10:, 1] = 0.03
    L = lattice(w, les, trt)
    forces, g = L.adjoint(N, Fbar)
    gray = np.flatnonzero((w > 0) & (w < 1))
    assert gray.size > 50
    rng = np.random.default_rng(4)
    cells = rng.choice(gray, 6, replace=False)
    eps = 1e-5
    for c in cells:
        J, wv = [], []
        for s in (1, -1):
            wp = w.copy()
            wp[c] = np.float32(w[c] + s * eps)      # weights are float32: use the stored value
            wv.append(float(wp[c]))
            Lp = lattice(wp, les, trt)
            J.append(objective(Lp.step(N), Fbar))
        fd = (J[0] - J[1]) / (wv[0] - wv[1])
        assert abs(g[c] - fd) <= 1e-6 * max(1e-12, np.abs(g[gray]).max()) + 1e-4 * abs(fd), (c, g[c], fd)
    # the adjoint run's own forward forces equal a plain forward run
    assert np.array_equal(forces, lattice(w, les, trt).step(N))
    assert np.all(g[(w <= 0) | (w >= 1)] == 0)


