"""The CPU oracle against golden vectors written from the reference under Node.

Every solver fixture is replayed through the oracle: same lattice, same
parameters, the same body (replayed from the recorded inside/wallVel mask),
the same number of steps. Hashes of the population buffers and macroscopic
fields and the force series must match bit for bit.
"""
import numpy as np
import pytest

from _util import arr, fixture_names, load, replay_voxelize, set_params, sha, unbits
from oracle.oracle import OracleLBM, grid_mask, voxel_bounds, voxel_build
from windtunnel_amd import synth

HASH_KEYS = ("f", "g", "rho", "ux", "uy", "uz", "solid", "near", "uwx", "uwy", "uwz")


def check_hashes(L, want: dict, keys=HASH_KEYS):
    bad = [k for k in keys if sha(L.array(k)) != want[k]]
    st = L.state()
    assert st["parity"] == want["parity"] and st["steps"] == want["steps"]
    assert not bad, f"mismatching arrays: {bad}"


def check_voxelize(res: dict, rec: dict):
    for k, v in rec["result"].items():
        assert res[k] == v, (k, res[k], v)


def test_init_and_voxelize():
    fx = load("lbm_init")
    L = OracleLBM(*fx["dims"])
    L.set_params(u0=fx["u0"])
    L.reset()
    check_hashes(L, fx["reset"])
    for rec in fx["voxelize"]:
        check_voxelize(replay_voxelize(L, rec), rec)
        check_hashes(L, rec["after"])


SMALL = ["lbm_bgk_sphere", "lbm_trt_les", "lbm_moving", "lbm_outlet_body", "lbm_revox"]


@pytest.mark.parametrize("name", SMALL)
def test_small_trajectories(name):
    fx = load(name)
    L = OracleLBM(*fx["dims"])
    set_params(L, fx["params"])
    L.reset()
    vres = replay_voxelize(L, fx["voxelize"])
    check_voxelize(vres, fx["voxelize"])
    check_hashes(L, fx["voxelize"]["after"])
    assert vres["movingWalls"] == fx["params"]["movingWalls"]
    want_f = arr(fx["forces"])
    got = np.zeros_like(want_f)
    for s in range(fx["nsteps"]):
        got[s] = L.step(1)[0]
        snap = fx["snaps"].get(str(s + 1))
        if snap:
            check_hashes(L, snap)
    assert np.array_equal(got.view(np.uint64), want_f.view(np.uint64))
    raw = fx["snaps"].get("raw")
    if raw:
        for k, rec in raw.items():
            assert np.array_equal(L.array(k).view(np.uint32), arr(rec).reshape(-1).view(np.uint32))
    for smp in fx.get("samples", []):
        assert np.array_equal(L.sample(*smp["p"]), np.array(smp["u"]))
        assert L.is_solid(*smp["p"]) == smp["solid"]
    if "revox" in fx:
        check_voxelize(replay_voxelize(L, fx["revox"]), fx["revox"])
        check_hashes(L, fx["revox"]["after"])
        f2 = L.step(30)
        assert np.array_equal(f2, arr(fx["forces2"]))
        check_hashes(L, fx["snaps2"]["60"])


def test_pool_emulation():
    """Worker-pool path: slab split, worker-order force totals, and the
    README's claim that the fields equal the single-threaded run."""
    fx = load("lbm_pool3")
    wp = fx["worker_params"]
    for threaded in (False, True):
        L = OracleLBM(*fx["dims"])
        set_params(L, wp)
        L.reset()
        check_voxelize(replay_voxelize(L, fx["voxelize"]), fx["voxelize"])
        f = L.step_pool(fx["nsteps"], fx["nWorkers"], threaded=threaded)
        assert np.array_equal(f, arr(fx["forces"]))
        check_hashes(L, fx["pooled"])
    S = OracleLBM(*fx["dims"])
    set_params(S, wp)
    S.reset()
    replay_voxelize(S, fx["voxelize"])
    f = S.step(fx["nsteps"])
    assert np.array_equal(f, arr(fx["single_forces"]))
    check_hashes(S, fx["single"])
    # the pool's worker-major force sum differs from the cell-major one only by rounding
    np.testing.assert_allclose(arr(fx["forces"]), f, rtol=1e-9, atol=1e-12)


@pytest.mark.slow
def test_c1_sphere_2000_steps():
    """96x48x48 BGK sphere, 2000 steps: full trajectory bitwise."""
    fx = load("lbm_c1")
    L = OracleLBM(*fx["dims"])
    set_params(L, fx["params"])
    L.reset()
    check_voxelize(replay_voxelize(L, fx["voxelize"]), fx["voxelize"])
    want = arr(fx["forces"])
    done = 0
    for stop in (1, 100, 500, 2000):
        f = L.step(stop - done)
        assert np.array_equal(f, want[done:stop]), f"forces differ in steps {done}..{stop}"
        done = stop
        check_hashes(L, fx["snaps"][str(stop)])


@pytest.mark.slow
@pytest.mark.parametrize("shape", ["cube", "sphere"])
def test_testpage_bodies(shape):
    """test.html procedure: 400 warm-up + 300 averaged steps at 96x44x44."""
    name = "lbm_testpage_" + shape
    if name not in fixture_names("lbm_testpage_"):
        pytest.skip("fixture not generated")
    fx = load(name)
    L = OracleLBM(*fx["dims"])
    set_params(L, fx["params"])
    L.reset()
    v = replay_voxelize(L, fx["voxelize"])
    f = L.step(fx["warm"] + fx["avg"])
    assert np.array_equal(f, arr(fx["forces"]))
    check_hashes(L, fx["final"])
    Fx = 0.0
    Fy = 0.0
    for t in range(fx["warm"], fx["warm"] + fx["avg"]):
        Fx += f[t, 0]
        Fy += f[t, 1]
    q = 0.5 * fx["params"]["u0"] * fx["params"]["u0"] * v["frontalArea"]
    assert Fx / fx["avg"] / q == fx["cd"]


# ------------------------------------------------------------- voxelizer

def _mesh(fx, name):
    if "mesh" in fx:
        return arr(fx["mesh"]).reshape(-1)
    gens = {
        "vox_c2body": lambda: synth.ellipsoid_box(),
        "vox_c2body_r200": lambda: synth.ellipsoid_box(),
        "vox_leaky": lambda: synth.noisy_sphere(level=4, R=20, holes=0.02, seed=5),
        "vox_leaky_l6": lambda: synth.noisy_sphere(level=6, R=56, holes=0.005, seed=5),
        "vox_leaky_l7": lambda: synth.noisy_sphere(level=7, R=56, holes=0.005, seed=5),
    }
    P = gens[name]()
    assert synth.sha256(P) == fx["mesh_sha256"], "synthetic mesh bytes differ from the fixture's"
    return P


VOX = ["vox_box", "vox_sphere", "vox_c2body", "vox_c2body_r200", "vox_leaky", "vox_leaky_l6", "vox_leaky_l7", "vox_hitcap"]


@pytest.mark.parametrize("name", VOX)
def test_voxel_build(name):
    if name not in fixture_names("vox_"):
        pytest.skip("fixture not generated")
    fx = load(name)
    P = _mesh(fx, name)
    v = voxel_build(P, fx["res"])
    for k in ("nx", "ny", "nz", "cell", "inv"):
        assert v[k] == fx[k], k
    for k in ("min", "ext", "centre", "bbox"):
        assert v[k] == fx[k], k
    assert voxel_bounds(P) == fx["bounds"]
    assert v["stats"]["solid"] == fx["stats"]["solid"]
    assert v["stats"]["tris"] == fx["stats"]["tris"]
    assert v["stats"]["openness"] == fx["stats"]["openness"]
    assert sha(v["grid"]) == fx["grid_sha256"]


def test_voxel_errors():
    fx = load("vox_errors")
    cases = {"empty": np.zeros(0, np.float32), "short": np.zeros(8, np.float32),
             "degenerate": np.full(9, 1.5, np.float32),
             "too_large": np.array([0, 0, 0, 1, 0, 0, 0, 1, 1], np.float32)}
    for key, P in cases.items():
        res = 400 if key == "too_large" else 96
        with pytest.raises(ValueError) as e:
            voxel_build(P, res)
        assert str(e.value) == fx[key]


def test_grid_lookup_meshcube():
    """Grid-backed Body.makeTest over the lattice box (js/shapes.js:369-388)."""
    fx = load("vox_meshcube_lookup")
    vx = fx["vox"]
    n = vx["nx"] * vx["ny"] * vx["nz"]
    grid = unbits(vx["grid_bits"], n)
    # the mesh cube's voxel grid itself is reproduced by the oracle voxelizer
    v = voxel_build(arr(vx["mesh"]).reshape(-1), 96)
    assert np.array_equal(v["grid"], grid)
    rec = fx["meshcube"]
    b = rec["body"]
    mask = grid_mask(grid, [vx["nx"], vx["ny"], vx["nz"]], vx["min"], vx["inv"],
                     b["M"], b["pos"], b["scale"], rec["box"])
    want = arr(rec["mask"]).reshape(-1)
    assert np.array_equal(mask, (want > 0).astype(np.uint8))
    # frontal areas of mesh cube and analytic cube agree as testmodels.html requires (<6%)
    fa_m, fa_a = rec["result"]["frontalArea"], fx["cube"]["result"]["frontalArea"]
    assert abs(fa_m - fa_a) / fa_a < 0.06


@pytest.mark.slow
@pytest.mark.parametrize("k", [2, 6])
def test_testpage_case_table(k):
    """test.html's ground vehicle (sedan, on the road) and wing at 8 degrees
    (planform area): every force and the final state bit for bit."""
    fx = load("lbm_testpage_cases")
    c = fx["cases"][k]
    L = OracleLBM(*fx["dims"])
    set_params(L, fx["params"])
    L.set_params(tau=fx["tau"])
    L.reset()
    v = replay_voxelize(L, c["voxelize"])
    assert (v["planformArea"] if c["planform"] else v["frontalArea"]) == c["area"]
    f = L.step(fx["warm"] + fx["avg"])
    assert np.array_equal(f, arr(c["forces"]))
    check_hashes(L, c["final"])
