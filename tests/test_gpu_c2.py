"""The 256x128x128 TRT (Lambda = 3/16) + LES (Cs = 0.1) wake of the seeded
ellipsoid+box mesh (~20k triangles), through the app's own mesh pipeline:
Shapes.orient to unit length, Voxel.build at res 200, the grid-backed
Body.makeTest at scale 1/k, then 5000 steps (tau = 0.505, u0 = 0.08).

Against the reference run under Node (lbm_c2): the voxel grid and the
placed body bit for bit, the whole state (populations and fields) bit for
bit after 1, 100, 1000 and 5000 steps in binary64 arithmetic, every step's
force, and the float32-arithmetic force series over the steps before the
wake's chaos separates runs of different rounding.

tau = 0.505 is this flow's stability edge in the reference's own
arithmetic: the same run at tau = 0.503 or 0.504 ends in NaN near step 500.
The float32 fast path follows that trajectory and crosses the edge at
tau = 0.505 (NaN near step 520), so it is compared over the first steps only.
"""
import numpy as np
import pytest

from _util import arr, load, set_params, sha
from windtunnel_amd import synth
from windtunnel_amd.lbm import LBM, Voxel
from windtunnel_amd.shapes import orient

pytestmark = pytest.mark.gpu


def placed(fx, **kw):
    """Lattice with the fixture's parameters and the mesh body placed on the GPU."""
    P = synth.ellipsoid_box()
    assert synth.sha256(P) == fx["mesh"]["sha256"]
    k = orient(P, 0, 1, True, False)
    assert k == fx["mesh"]["k"] and synth.sha256(P) == fx["mesh"]["oriented_sha256"]
    vox = Voxel.build(P, fx["mesh"]["addModel"]["res"])
    L = LBM(*fx["dims"], **kw)
    set_params(L, fx["params"])
    L.reset()
    b = fx["voxelize"]["body"]
    bi = fx["voxelize"]["box_in"]
    L.voxelize_grid(vox, b["M"], b["pos"], b["scale"], dict(x0=bi[0], y0=bi[1], z0=bi[2], x1=bi[3], y1=bi[4], z1=bi[5]))
    return L, vox


def test_c2_voxel_grid_and_placement():
    fx = load("lbm_c2")
    print("[c2] mesh pipeline", flush=True)
    L, vox = placed(fx)
    want = fx["vox"]
    for key in ("nx", "ny", "nz", "cell", "min", "inv", "ext"):
        assert vox[key] == want[key], key
    assert sha(vox["grid"]) == want["grid_sha256"]
    assert vox["stats"] == want["stats"]
    after = fx["voxelize"]["after"]
    s, nr = L.solid_near()
    assert sha(s) == after["solid"] and sha(nr) == after["near"]
    res = fx["voxelize"]["result"]
    assert (L.frontalArea, L.planformArea, L.solidCells, L.rideHeight) == (
        res["frontalArea"], res["planformArea"], res["solidCells"], res["rideHeight"])


def test_c2_exact_mode_5000_steps():
    fx = load("lbm_c2")
    L, _ = placed(fx, arith="f64")
    want = arr(fx["forces"])
    done = 0
    for stop in sorted(int(k) for k in fx["snaps"]):
        print(f"[c2] exact mode to step {stop}", flush=True)
        f = L.step(stop - done)
        np.testing.assert_allclose(f, want[done:stop], rtol=1e-9, atol=1e-12)
        done = stop
        snap = fx["snaps"][str(stop)]
        got = {"f": sha(L.f), "g": sha(L.g)}
        F = L.fields()
        for k in ("rho", "ux", "uy", "uz"):
            got[k] = sha(F[k])
        bad = [k for k in got if got[k] != snap[k]]
        assert not bad, f"step {stop}: arrays differ from the reference: {bad}"
        assert L.parity == snap["parity"] and L.steps == snap["steps"]




@pytest.mark.parametrize("storage,steps,tol", [("f32", 300, 1e-5), ("f16", 100, 2e-3)])
def test_c2_fast_paths_track_reference(storage, steps, tol):
    """Float32 arithmetic (float32 or binary16 storage) follows the reference's
    force series until the wake's chaos separates the runs: max |dF| over the
    window relative to max |F| of the reference."""
    fx = load("lbm_c2")
    L, _ = placed(fx, storage=storage)
    want = arr(fx["forces"])[:steps]
    got = L.step(steps)
    err = float(np.abs(got - want).max() / np.abs(want).max())
    print(f"[c2] {storage}: first {steps} steps, max force deviation {err:.2e}", flush=True)
    assert err < tol
