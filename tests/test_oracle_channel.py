"""Channel boundaries of the oracle (configuration variant with free-slip
side walls and an equilibrium outflow; no reference function).

Pinned by what the mathematics fixes:
* uniform flow u0 with no body stays uniform (the reflection of the
  equilibrium at u = (u0, 0, 0) is itself, and the outlet equilibrium takes
  the same velocity);
* a free-slip wall is a mirror: a body next to the wall gives the same flow
  as the body and its mirror image in a lattice twice as wide, whose
  mirror plane carries no wall at all. A wrong reflection (an e_y left
  unflipped, a ghost one cell off) breaks the symmetry at the first step.
"""
import numpy as np
import pytest

from oracle.oracle import OracleLBM


def lattice(dims, mask, prec=8):
    L = OracleLBM(*dims, prec=prec)
    L.set_params(tau=0.6, les=False, trt=True, csm=0.14, u0=0.05, movingWalls=False, **{"lambda": 3 / 16})
    L.set_boundary_mode("channel")
    L.reset()
    if mask is not None:
        nx, ny, nz = dims
        L.voxelize_mask([1, 1, 1, nx - 2, ny - 2, nz - 2], mask[1:-1, 1:-1, 1:-1].astype(np.uint8).reshape(-1))
    return L


def fields(L, dims):
    nx, ny, nz = dims
    return {k: L.array(k).reshape(nz, ny, nx).astype(np.float64) for k in ("rho", "ux", "uy", "uz")}


def test_uniform_flow_stays_uniform():
    dims = (24, 10, 9)
    L = lattice(dims, None)
    L.step(60)
    F = fields(L, dims)
    assert np.abs(F["ux"] - 0.05).max() < 1e-9
    assert np.abs(F["uy"]).max() < 1e-9 and np.abs(F["uz"]).max() < 1e-9
    assert np.abs(F["rho"] - 1).max() < 1e-9


def sphere(dims, c, r):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    return (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r


@pytest.mark.parametrize("axis", ["y", "z"])
def test_free_slip_wall_is_a_mirror(axis):
    nx, m = 40, 12                          # m interior rows across the mirrored axis
    other = 11                              # the other transverse extent (cells incl. ghosts)
    half = (nx, m + 2, other) if axis == "y" else (nx, other, m + 2)
    full = (nx, 2 * m + 2, other) if axis == "y" else (nx, other, 2 * m + 2)
    ax = 1 if axis == "y" else 0            # axis index in [z, y, x] arrays
    c = [14.0, 5.0, 5.5] if axis == "y" else [14.0, 5.5, 5.0]
    mh = sphere(half, c, 2.2)
    mf = np.zeros(full[::-1], dtype=bool)
    for r in range(1, m + 1):               # half row r -> full rows m + r and m + 1 - r
        src = np.take(mh, r, axis=ax)
        idx = [slice(None)] * 3
        idx[ax] = m + r
        mf[tuple(idx)] = src
        idx[ax] = m + 1 - r
        mf[tuple(idx)] = src
    H, Fl = lattice(half, mh), lattice(full, mf)
    H.step(250)
    Fl.step(250)
    A, B = fields(H, half), fields(Fl, full)
    sgn = {"uy": -1.0 if axis == "y" else 1.0, "uz": -1.0 if axis == "z" else 1.0, "ux": 1.0, "rho": 1.0}
    rows = np.arange(1, m + 1)
    for k in A:
        a = np.take(A[k], rows, axis=ax)
        b_up = np.take(B[k], m + rows, axis=ax)
        b_dn = np.take(B[k], m + 1 - rows, axis=ax)
        scale = max(1e-3, np.abs(a).max())
        assert np.abs(a - b_up).max() <= 1e-10 * scale, (k, np.abs(a - b_up).max())
        assert np.abs(a - sgn[k] * b_dn).max() <= 1e-10 * scale, (k, "mirror")
    assert np.abs(A["uy" if axis == "y" else "uz"]).max() > 1e-4  # the body deflects flow across the axis
