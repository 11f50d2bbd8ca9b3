"""GPU voxelizer against the reference's Voxel.build (golden grids, bit-exact).

The golden grids were written from js/voxel.js under Node; occupancy,
solid count and the openness (axis-disagreement) fraction must match
exactly, including the 512-crossing cap of a dense ray (vox_hitcap) and
leaky meshes whose axes disagree (vox_leaky*).
"""
import numpy as np
import pytest

from _util import arr, load, sha, unbits
from windtunnel_amd import synth
from windtunnel_amd.lbm import LBM, Voxel

pytestmark = pytest.mark.gpu

GEN = {
    "vox_c2body": lambda: synth.ellipsoid_box(),
    "vox_c2body_r200": lambda: synth.ellipsoid_box(),
    "vox_leaky": lambda: synth.noisy_sphere(level=4, R=20, holes=0.02, seed=5),
    "vox_leaky_l6": lambda: synth.noisy_sphere(level=6, R=56, holes=0.005, seed=5),
    "vox_leaky_l7": lambda: synth.noisy_sphere(level=7, R=56, holes=0.005, seed=5),
}


def mesh(name, fx):
    if "mesh" in fx:
        return arr(fx["mesh"]).reshape(-1)
    P = GEN[name]()
    assert synth.sha256(P) == fx["mesh_sha256"]
    return P


@pytest.mark.parametrize("name", ["vox_box", "vox_sphere", "vox_c2body", "vox_c2body_r200", "vox_leaky",
                                  "vox_leaky_l6", "vox_leaky_l7", "vox_hitcap"])
def test_build_bit_exact(name):
    fx = load(name)
    P = mesh(name, fx)
    print(f"[voxel] {name} ntri={P.size // 9} res={fx['res']}", flush=True)
    v = Voxel.build(P, fx["res"])
    for k in ("nx", "ny", "nz", "cell", "inv", "min", "ext", "centre", "bbox"):
        assert v[k] == fx[k], k
    assert v["stats"]["solid"] == fx["stats"]["solid"]
    assert v["stats"]["openness"] == fx["stats"]["openness"]
    assert sha(v["grid"]) == fx["grid_sha256"]
    assert Voxel.bounds(P) == fx["bounds"]


def test_errors_match_reference_messages():
    fx = load("vox_errors")
    cases = {"empty": np.zeros(0, np.float32), "short": np.zeros(8, np.float32),
             "degenerate": np.full(9, 1.5, np.float32),
             "too_large": np.array([0, 0, 0, 1, 0, 0, 0, 1, 1], np.float32)}
    for key, P in cases.items():
        with pytest.raises(ValueError) as e:
            Voxel.build(P, 400 if key == "too_large" else 96)
        assert str(e.value) == fx[key]


def test_mesh_cube_pipeline_on_gpu():
    """testmodels.html mesh-cube path: GPU voxel grid, then the GPU grid lookup
    into an 80x36x36 lattice, equal to the reference's solid map."""
    fx = load("vox_meshcube_lookup")
    vx = fx["vox"]
    v = Voxel.build(arr(vx["mesh"]).reshape(-1), 96)
    assert np.array_equal(v["grid"], unbits(vx["grid_bits"], vx["nx"] * vx["ny"] * vx["nz"]))
    rec = fx["meshcube"]
    b = rec["body"]
    L = LBM(80, 36, 36)
    L.voxelize_grid(v, b["M"], b["pos"], b["scale"], rec["box"])
    assert sha(L.solid) == rec["after"]["solid"]
    assert L.frontalArea == rec["result"]["frontalArea"]
