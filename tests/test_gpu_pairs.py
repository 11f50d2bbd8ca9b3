"""The step with two cells per thread (binary16 and float32 storage, two
buffers, rows of 128-cell multiples) against the one-cell step: the same populations and fields bit for
bit and the same forces to the order of their sum, with a body (solid and near-wall cells, a solid cell in the
column next to the outlet), a rolling-road floor, moving walls, the slab
form and the steps that write the fields. WT_PAIRS (read when a lattice is
created) selects the one-cell step (0) or two cells for float32 too (2; by
default only binary16 takes two)."""
import numpy as np
import pytest

from windtunnel_amd.lbm import LBM

pytestmark = pytest.mark.gpu

DIMS = (256, 24, 20)


def run(monkeypatch, storage, pairs, les, trt, ground, dims=DIMS):
    monkeypatch.setenv("WT_PAIRS", str(pairs))
    L = LBM(*dims, storage=storage)
    assert L.block_cells == (128 if storage == "f16" else 64)
    L.set_params(tau=0.56, les=les, trt=trt, csm=0.1, u0=0.05, **{"lambda": 3 / 16})
    L.reset()
    nx, ny, nz = dims
    # box x 60..nx-2, y 5..16, z 4..15: a spinning sphere (wall velocity:
    # moving-wall links) and a solid column next to the outlet (x = nx - 2)
    z, y, x = np.meshgrid(np.arange(4, 16), np.arange(5, 17), np.arange(60, nx - 1), indexing="ij")
    sphere = (x + 0.5 - 66) ** 2 + (y + 0.5 - 11) ** 2 + (z + 0.5 - 10) ** 2 <= 20
    column = (x == nx - 2) & (y >= 9) & (y <= 11) & (z >= 9) & (z <= 11)
    mask = np.where(sphere, 2, np.where(column, 1, 0)).astype(np.uint8).reshape(-1)
    wv = np.tile(np.array([0.0, 0.01, 0.0], np.float32), int(np.count_nonzero(mask == 2)))
    L.voxelize_mask([60, 5, 4, nx - 2, 16, 15], mask, wv, ground, 0.05 if ground else 0.0)
    out = {"f1": L.step(5)}
    out["f2"] = L.step(4)
    out["f3"] = L.step_slabs(3, 2)
    out["pops"] = L.pops(0)
    out.update({k: v.copy() for k, v in L.fields().items()})
    L.close()
    return out


@pytest.mark.parametrize("storage", ["f16", "f32"])
@pytest.mark.parametrize("les,trt,ground", [(True, True, False), (False, False, True), (True, False, True)])
def test_pairs_equal_one_cell_step(monkeypatch, storage, les, trt, ground):
    print(f"[pairs] {storage} les={les} trt={trt} ground={ground}", flush=True)
    one = run(monkeypatch, storage, 0, les, trt, ground)
    two = run(monkeypatch, storage, 2, les, trt, ground)
    for k in one:
        if k.startswith("f"):
            # the force partials of near-wall cells are summed per block: a
            # different order with two cells per thread
            np.testing.assert_allclose(two[k], one[k], rtol=1e-12, atol=1e-18, err_msg=k)
        else:
            assert np.array_equal(one[k], two[k], equal_nan=True), k
    assert np.isfinite(one["pops"]).all()


def test_pairs_five_wave_rows(monkeypatch):
    """Rows of 384 cells: three-wave workgroups, which take the pair step's
    five-waves-per-SIMD variant; the same populations and fields as the
    one-cell step bit for bit."""
    dims = (384, 24, 20)
    print("[pairs] f16 384-cell rows", flush=True)
    one = run(monkeypatch, "f16", 0, True, True, True, dims)
    two = run(monkeypatch, "f16", 1, True, True, True, dims)
    for k in one:
        if k.startswith("f"):
            np.testing.assert_allclose(two[k], one[k], rtol=1e-12, atol=1e-18, err_msg=k)
        else:
            assert np.array_equal(one[k], two[k], equal_nan=True), k
