"""Parity at the throughput workload's own size: the 512x256x256 TRT+LES
lattice (tau 0.505, u0 0.08, Cs 0.1) around the ~20k-triangle mesh body
scaled 2x, voxelized on the GPU. In binary64 arithmetic (float storage, the
production 64-cell blocks, two buffers and in place) two steps must leave
the populations, fields and solid/near maps equal to the oracle's (the C
restatement of js/lbm.js) bit for bit, and the forces within 1e-11 (the
device sums per-workgroup partials, the reference cell by cell). The
float32 and binary16 paths must stay within 1e-5 and 2e-3 of the binary64
forces over the same steps."""
import numpy as np
import pytest

from oracle.oracle import OracleLBM
from test_gpu_lbm import stored_mask
from windtunnel_amd.lbm import LBM, Voxel
from windtunnel_amd.synth import ellipsoid_box

pytestmark = pytest.mark.gpu

DIMS = (512, 256, 256)
PARAMS = dict(tau=0.505, les=True, trt=True, csm=0.1, u0=0.08, **{"lambda": 3 / 16})


def lattice(vox, **kw):
    L = LBM(*DIMS, **kw)
    L.set_params(**PARAMS)
    L.reset()
    b = np.array(vox["bbox"])
    pos = np.array([0.3 * DIMS[0], DIMS[1] / 2, DIMS[2] / 2])
    lo, hi = pos + 2 * b[:3] - 2, pos + 2 * b[3:] + 2
    L.voxelize_grid(vox, np.eye(3).reshape(-1), pos, 2.0, dict(x0=lo[0], y0=lo[1], z0=lo[2], x1=hi[0], y1=hi[1],
                                                                  z1=hi[2]))
    return L


def test_full_size_two_steps():
    vox = Voxel.build(ellipsoid_box(), 200)
    print("[full] binary64 arithmetic on the GPU", flush=True)
    L = lattice(vox, arith="f64")
    assert L.block_cells == 64
    s, nr = L.solid_near()
    sg = s.reshape(DIMS[::-1])
    idx = np.argwhere(sg == 1)
    lo, hi = idx.min(0), idx.max(0)
    box = [int(lo[2]), int(lo[1]), int(lo[0]), int(hi[2]), int(hi[1]), int(hi[0])]
    mask = (sg[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] == 1).astype(np.uint8).reshape(-1)
    print("[full] oracle", flush=True)
    O = OracleLBM(*DIMS)
    O.set_params(**PARAMS)
    O.reset()
    st = O.voxelize_mask(box, mask)
    assert (L.frontalArea, L.planformArea, L.solidCells) == (st["frontalArea"], st["planformArea"], st["solidCells"])
    assert np.array_equal(s, O.array("solid")) and np.array_equal(nr, O.array("near"))
    fg = L.step(2)
    fo = O.step(2)
    np.testing.assert_allclose(fg, fo, rtol=1e-11, atol=1e-14)
    print("[full] populations and fields", flush=True)
    assert np.array_equal(L.f, O.array("f"))
    assert np.array_equal(L.g, O.array("g"))
    F = L.fields()
    for k in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(F[k], O.array(k)), k
    L.close()
    print("[full] in place (esoteric pull), binary64 arithmetic", flush=True)
    E = lattice(vox, arith="f64", layout="esoteric")
    fe = E.step(2)
    np.testing.assert_allclose(fe, fo, rtol=1e-11, atol=1e-14)
    FE = E.fields()
    for k in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(FE[k], O.array(k)), k
    m = stored_mask(DIMS, O.array("solid"))
    assert np.array_equal(E.f[m], O.array("f")[m])
    E.close()
    del O, m
    for storage, tol in (("f32", 1e-5), ("f16", 2e-3)):
        print(f"[full] {storage} storage, float arithmetic", flush=True)
        G = lattice(vox, storage=storage)
        f = G.step(2)
        G.close()
        err = float(np.abs(f - fg).max() / np.abs(fg).max())
        print(f"[full] {storage}: max force deviation {err:.2e}", flush=True)
        assert err < tol
