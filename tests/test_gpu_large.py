"""A lattice beyond 2^31 / 19 cells (the old 32-bit population index range):
1280x320x320 = 131 M cells in float32 storage with two buffers (25 GB).

Property: information travels one cell per step, so after T steps the cells
farther than T from the outlet see nothing of the lattice's length. A sphere
placed the same way in a 1280-long and a 256-long lattice of the same
cross-section gives, after T = 4 steps, the same speed and pressure slices
bit for bit over x < 240, and the same force to the order of its sum (the
near-wall cells are summed in blocks of the row width)."""
import numpy as np
import pytest

from windtunnel_amd.lbm import LBM

pytestmark = pytest.mark.gpu

NY = NZ = 320
T = 4


def run(nx):
    L = LBM(nx, NY, NZ)
    L.set_params(tau=0.6, les=True, csm=0.14, trt=True, u0=0.05, **{"lambda": 3 / 16})
    L.reset()
    c = np.array([64.0, NY / 2, NZ / 2])
    lo, hi = (c - 12).astype(int), (c + 12).astype(int)
    z, y, x = np.meshgrid(*(np.arange(lo[k], hi[k] + 1) + 0.5 for k in (2, 1, 0)), indexing="ij")
    mask = (((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) <= 100).astype(np.uint8).reshape(-1)
    L.voxelize_mask([lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]], mask)
    F = L.step(T)
    sl = {f: L.slice("xy", NZ // 2, f)[0][:, :240].copy() for f in ("speed", "pressure")}
    L.close()
    return F, sl


def test_lattice_beyond_the_old_index_range():
    print("[large] 1280x320x320 float32", flush=True)
    Fb, sb = run(1280)
    print("[large] 256x320x320 float32", flush=True)
    Fs, ss = run(256)
    assert np.isfinite(Fb).all()
    np.testing.assert_allclose(Fb, Fs, rtol=1e-12, atol=1e-15)
    for f in sb:
        assert np.array_equal(sb[f], ss[f]), f
