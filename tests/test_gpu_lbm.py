"""GPU parity of the fused D3Q19 step (libwindtunnel on gfx950).

* binary64 collision with float storage must reproduce the reference's own
  Float32Array contents bit for bit: population buffers, macroscopic fields,
  solid/near maps against the golden hashes written from js/lbm.js;
* the float32 fast path must stay within relative L2 1e-4 of the oracle on
  density, velocity and C_p, and of the force-coefficient series, at
  laminar Reynolds numbers;
* FP64 storage must equal the oracle's FP64-storage run bit for bit.

The bitwise tests run in both population layouts: planar planes and
blocks of 2^k cells (WT_BLOCK_SHIFT = k, any k whose block divides nx).
"""
import numpy as np
import pytest

from _util import arr, load, rel_l2, replay_voxelize, set_params, sha, unbits
from oracle.oracle import OracleLBM
from windtunnel_amd.lbm import LBM

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "ux", "uy", "uz")


# storage layouts: planar, and 8-cell blocks (every golden grid has nx % 8 == 0)
LAYOUTS = [0, 3]


@pytest.fixture
def block_shift(request, monkeypatch):
    """Select the population layout of lattices created in the test."""
    monkeypatch.setenv("WT_BLOCK_SHIFT", str(request.param))
    return request.param


def expect_layout(L, shift):
    assert L.block_cells == (1 << shift if shift else 0)


def make(fx, **kw):
    L = LBM(*fx["dims"], **kw)
    set_params(L, fx["params"])
    L.reset()
    return L


def oracle_for(fx, prec=4):
    O = OracleLBM(*fx["dims"], prec=prec)
    set_params(O, fx["params"])
    O.reset()
    replay_voxelize(O, fx["voxelize"])
    return O


def check_hashes(L, want, keys=("f", "g", "rho", "ux", "uy", "uz")):
    got = {"f": sha(L.f), "g": sha(L.g)}
    F = L.fields()
    for k in FIELDS:
        got[k] = sha(F[k])
    bad = [k for k in keys if got[k] != want[k]]
    assert not bad, f"arrays differ from the reference: {bad}"
    assert L.parity == want["parity"] and L.steps == want["steps"]


@pytest.mark.parametrize("block_shift", LAYOUTS, indirect=True)
@pytest.mark.parametrize("name", ["lbm_bgk_sphere", "lbm_trt_les", "lbm_moving", "lbm_outlet_body", "lbm_revox"])
def test_exact_mode_bitwise(name, block_shift):
    fx = load(name)
    L = make(fx, arith="f64")
    expect_layout(L, block_shift)
    print(f"[exact] {name} dims={fx['dims']} block_shift={block_shift}", flush=True)
    replay_voxelize(L, fx["voxelize"])
    s, nr = L.solid_near()
    assert sha(s) == fx["voxelize"]["after"]["solid"]
    assert sha(nr) == fx["voxelize"]["after"]["near"]
    want_f = arr(fx["forces"])
    done = 0
    got = []
    for stop in sorted(int(k) for k in fx["snaps"] if k != "raw"):
        got.append(L.step(stop - done))
        done = stop
        check_hashes(L, fx["snaps"][str(stop)])
    got = np.concatenate(got)
    # the device sums the force in a tree order; the reference sums cell by cell
    np.testing.assert_allclose(got, want_f[:done], rtol=1e-11, atol=1e-14)
    if "revox" in fx:
        replay_voxelize(L, fx["revox"])
        f2 = L.step(30)
        check_hashes(L, fx["snaps2"]["60"])
        np.testing.assert_allclose(f2, arr(fx["forces2"]), rtol=1e-11, atol=1e-14)


def test_exact_mode_init_and_voxelize():
    fx = load("lbm_init")
    L = LBM(*fx["dims"], arith="f64")
    L.u0 = fx["u0"]
    L.reset()
    check_hashes(L, fx["reset"])
    for rec in fx["voxelize"]:
        replay_voxelize(L, rec)
        check_hashes(L, rec["after"])
        assert L.frontalArea == rec["result"]["frontalArea"]
        assert L.rideHeight == rec["result"]["rideHeight"]
        assert L.movingWalls == rec["result"]["movingWalls"]


@pytest.mark.parametrize("block_shift", [0, 5], indirect=True)
def test_exact_mode_c1_2000_steps(block_shift):
    """96x48x48 BGK sphere, the CPU-runnable workload, 2000 steps bitwise."""
    fx = load("lbm_c1")
    L = make(fx, arith="f64")
    expect_layout(L, block_shift)
    replay_voxelize(L, fx["voxelize"])
    want = arr(fx["forces"])
    done = 0
    for stop in (1, 100, 500, 2000):
        f = L.step(stop - done)
        np.testing.assert_allclose(f, want[done:stop], rtol=1e-10, atol=1e-13)
        done = stop
        check_hashes(L, fx["snaps"][str(stop)])


@pytest.mark.parametrize("name", ["lbm_bgk_sphere", "lbm_trt_les", "lbm_moving"])
def test_fp32_fields_within_1e4(name):
    fx = load(name)
    L = make(fx)
    replay_voxelize(L, fx["voxelize"])
    O = oracle_for(fx)
    fg = L.step(fx["nsteps"])
    fo = O.step(fx["nsteps"])
    F = L.fields()
    for k in FIELDS:
        e = rel_l2(F[k], O.array(k))
        assert e < 1e-4, f"{k}: rel-L2 {e:.3e}"
    u0 = fx["params"]["u0"]
    cp_g = (F["rho"].astype(np.float64) - 1) / 3 / (0.5 * u0 * u0)
    cp_o = (O.array("rho").astype(np.float64) - 1) / 3 / (0.5 * u0 * u0)
    assert rel_l2(cp_g, cp_o) < 1e-4
    assert rel_l2(fg[:, 0], fo[:, 0]) < 1e-4


def test_fp32_c1_force_and_fields():
    """C1 (laminar, Re 40): 2000 fp32 steps against the oracle."""
    fx = load("lbm_c1")
    L = make(fx)
    replay_voxelize(L, fx["voxelize"])
    O = oracle_for(fx)
    fg = L.step(2000)
    fo = O.step(2000)
    F = L.fields()
    for k in FIELDS:
        e = rel_l2(F[k], O.array(k))
        assert e < 1e-4, f"{k}: rel-L2 {e:.3e}"
    A = L.frontalArea
    q = 0.5 * 0.05 ** 2 * A
    assert rel_l2(fg[:, 0] / q, fo[:, 0] / q) < 1e-4
    cd_g, cd_o = fg[-500:, 0].mean() / q, fo[-500:, 0].mean() / q
    assert abs(cd_g - cd_o) / abs(cd_o) < 1e-4


def test_uniform_flow_stays_uniform():
    """No body: uniform inflow stays uniform. In the reference's own
    arithmetic the fields are an exact fixed point while the stored
    populations drift by rounding only (below 1e-9), and the binary64 path
    reproduces that drift bit for bit."""
    L = LBM(40, 20, 18, arith="f64")
    L.set_params(tau=0.6, les=True, trt=True, u0=0.07)
    L.reset()
    f0 = L.f.copy()
    L.step(25)
    O = OracleLBM(40, 20, 18)
    O.set_params(tau=0.6, les=True, trt=True, u0=0.07)
    O.reset()
    O.step(25)
    assert np.array_equal(L.f, O.array("f"))
    assert np.max(np.abs(L.f - f0)) < 1e-9
    assert np.all(L.ux == np.float32(0.07)) and np.all(L.rho == 1)
    L32 = LBM(40, 20, 18)
    L32.set_params(tau=0.6, les=True, trt=True, u0=0.07)
    L32.reset()
    L32.step(25)
    assert np.max(np.abs(L32.ux - np.float32(0.07))) < 1e-6
    assert np.max(np.abs(L32.rho - 1)) < 1e-6


def test_fp64_storage_matches_fp64_oracle():
    fx = load("lbm_trt_les")
    L = make(fx, storage="f64", arith="f64")
    replay_voxelize(L, fx["voxelize"])
    O = oracle_for(fx, prec=8)
    L.step(50)
    O.step(50)
    assert np.array_equal(L.f, O.array("f"))
    for k in FIELDS:
        assert np.array_equal(L.fields()[k], O.array(k))


@pytest.mark.parametrize("block_shift", LAYOUTS, indirect=True)
@pytest.mark.parametrize("les", [False, True])
def test_exact_mode_tau_one_half(les, block_shift):
    """tau = 0.5 exactly: TRT's Lambda / (tau_eff - 1/2) divides by zero in
    every cell at rest, where IEEE division gives an infinity and the
    antisymmetric rate 1 / (1/2 + inf) = 0. Binary64 arithmetic must still
    equal the oracle (IEEE division in the reference's order) bit for bit,
    in the planar layout and in blocks (the interior cells' straight path)."""
    fx = load("lbm_trt_les")
    p = dict(fx["params"], tau=0.5, les=les, trt=True)
    L = LBM(*fx["dims"], arith="f64")
    set_params(L, p)
    L.reset()
    replay_voxelize(L, fx["voxelize"])
    O = OracleLBM(*fx["dims"], prec=4)
    set_params(O, p)
    O.reset()
    replay_voxelize(O, fx["voxelize"])
    L.step(20)
    O.step(20)
    want = O.array("f")
    assert np.isfinite(want).any()
    assert np.array_equal(L.f, want, equal_nan=True)
    for k in FIELDS:
        assert np.array_equal(L.fields()[k], O.array(k), equal_nan=True)


@pytest.mark.parametrize("block_shift", LAYOUTS, indirect=True)
def test_pool_slabs_exact(block_shift):
    """Worker-pool form: 3 slabs, per-slab forces, fields equal to the fixture."""
    fx = load("lbm_pool3")
    L = LBM(*fx["dims"], arith="f64")
    expect_layout(L, block_shift)
    set_params(L, fx["worker_params"])
    L.reset()
    replay_voxelize(L, fx["voxelize"])
    per = L.step_slabs(fx["nsteps"], fx["nWorkers"])
    check_hashes(L, fx["pooled"])
    tot = np.zeros((fx["nsteps"], 3))
    for w in range(fx["nWorkers"]):
        tot = tot + per[:, w, :]
    np.testing.assert_allclose(tot, arr(fx["forces"]), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("layout", ["ab", "esoteric"])
def test_worker_protocol_through_step_slab(layout):
    """The lbmWorker.js step loop (js/lbmWorker.js:59-79) driven through the
    solver-object calls: per worker stepSlab(z0,z1) then its share of
    _boundaries, parity flip after all workers; fields equal the reference's
    pooled run and the per-step force is the worker-ordered sum."""
    fx = load("lbm_pool3")
    L = LBM(*fx["dims"], arith="f64", layout=layout)
    set_params(L, fx["worker_params"])
    L.reset()
    replay_voxelize(L, fx["voxelize"])
    nz, NW = fx["dims"][2], fx["nWorkers"]
    lo, hi = 1, nz - 1
    got = np.zeros((fx["nsteps"], 3))
    for s in range(fx["nsteps"]):
        for w in range(NW):
            z0 = lo + (hi - lo) * w // NW
            z1 = lo + (hi - lo) * (w + 1) // NW
            bz0 = 0 if w == 0 else z0
            bz1 = nz if w == NW - 1 else z1
            got[s] += L.step_slab(z0, z1)
            L.boundaries(bz0, bz1, w == 0, w == NW - 1)
        L.parity = L.parity ^ 1
    for k in FIELDS:
        assert sha(L.fields()[k]) == fx["pooled"][k], k
    np.testing.assert_allclose(got, arr(fx["forces"]), rtol=1e-11, atol=1e-14)


def test_fp16_storage_tracks_fp64():
    """binary16 deviations with float32 collision against the FP64 oracle
    (a short transient: stochastic rounding's noise, about 1.4x that of
    round-to-nearest per store, sets the bound here)."""
    fx = load("lbm_trt_les")
    L = make(fx, storage="f16")
    replay_voxelize(L, fx["voxelize"])
    O = oracle_for(fx, prec=8)
    L.step(fx["nsteps"])
    O.step(fx["nsteps"])
    F = L.fields()
    u = np.sqrt(F["ux"].astype(np.float64) ** 2 + F["uy"].astype(np.float64) ** 2)
    uo = np.sqrt(O.array("ux") ** 2 + O.array("uy") ** 2)
    assert rel_l2(u, uo) < 5e-3
    assert rel_l2(F["rho"], O.array("rho")) < 1e-4


def test_fp16_steady_drag_matches_fp64():
    """A steady laminar wake over 3000 steps: binary16 storage with
    stochastic rounding keeps the time-averaged C_d and the velocity field
    at the binary64 values; round-to-nearest stalls updates below half a
    binary16 step and ends ~0.3 % off in C_d (1 % and more at lower
    viscosity)."""
    dims = (96, 40, 40)
    runs = {}
    for storage, arith in (("f64", "f64"), ("f16", "f32")):
        L = LBM(*dims, storage=storage, arith=arith)
        L.set_params(tau=0.58, les=False, trt=True, csm=0.14, u0=0.05, **{"lambda": 3 / 16})
        L.reset()
        z, y, x = np.meshgrid(np.arange(12, 29) + 0.5, np.arange(12, 29) + 0.5, np.arange(22, 39) + 0.5, indexing="ij")
        m = ((x - 30) ** 2 + (y - 20) ** 2 + (z - 20) ** 2 <= 36).astype(np.uint8).reshape(-1)
        L.voxelize_mask([22, 12, 12, 38, 28, 28], m)
        F = L.step(3000)
        runs[storage] = (np.mean(F[-1000:, 0]), L.fields()["ux"].astype(np.float64))
        L.close()
    cd_err = abs(runs["f16"][0] / runs["f64"][0] - 1)
    u_err = rel_l2(runs["f16"][1], runs["f64"][1])
    print(f"[f16 steady] C_d rel err {cd_err:.2e}, ux rel-L2 {u_err:.2e}", flush=True)
    assert cd_err < 5e-4 and u_err < 2e-3


def test_determinism_fp32():
    fx = load("lbm_trt_les")
    out = []
    for _ in range(2):
        L = make(fx)
        replay_voxelize(L, fx["voxelize"])
        f = L.step(40)
        out.append((f, L.f.copy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_grid_lookup_on_gpu():
    """Grid-backed body lookup (K5) against the reference's own mask."""
    fx = load("vox_meshcube_lookup")
    vx = fx["vox"]
    grid = unbits(vx["grid_bits"], vx["nx"] * vx["ny"] * vx["nz"])
    rec = fx["meshcube"]
    b = rec["body"]
    L = LBM(80, 36, 36)
    L.voxelize_grid(dict(grid=grid, nx=vx["nx"], ny=vx["ny"], nz=vx["nz"], min=vx["min"], inv=vx["inv"],
                         cell=vx["cell"]), b["M"], b["pos"], b["scale"], rec["box"])
    assert sha(L.solid) == rec["after"]["solid"]
    assert sha(L.near) == rec["after"]["near"]
    assert L.frontalArea == rec["result"]["frontalArea"]
    assert L.solidCells == rec["result"]["solidCells"]


# ------------------------------------------------------------ esoteric pull

def stored_mask(dims, solid):
    """Entries of f[q*n+cell] an in-place lattice stores: non-solid cells,
    and for odd q only where the neighbour along c_q is inside the lattice."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
    ok = []
    for q in range(19):
        m = solid == 0
        if q % 2 == 1:
            X, Y, Zc = x + LBM.EX[q], y + LBM.EY[q], z + LBM.EZ[q]
            m = m & (X >= 0) & (X < nx) & (Y >= 0) & (Y < ny) & (Zc >= 0) & (Zc < nz)
        ok.append(m)
    return np.concatenate(ok)


@pytest.mark.parametrize("block_shift", LAYOUTS, indirect=True)
@pytest.mark.parametrize("name", ["lbm_bgk_sphere", "lbm_trt_les", "lbm_moving", "lbm_revox"])
def test_esoteric_exact_mode(name, block_shift):
    """In-place streaming with binary64 collision: the macroscopic fields are
    the reference's bit for bit, and every stored population equals the
    oracle's f."""
    fx = load(name)
    L = make(fx, arith="f64", layout="esoteric")
    expect_layout(L, block_shift)
    print(f"[esoteric exact] {name}", flush=True)
    replay_voxelize(L, fx["voxelize"])
    O = oracle_for(fx)
    want_f = arr(fx["forces"])
    done = 0
    for stop in sorted(int(k) for k in fx["snaps"] if k != "raw"):
        f = L.step(stop - done)
        O.step(stop - done)
        np.testing.assert_allclose(f, want_f[done:stop], rtol=1e-11, atol=1e-14)
        done = stop
        F = L.fields()
        for k in FIELDS:
            assert sha(F[k]) == fx["snaps"][str(stop)][k], (stop, k)
        m = stored_mask(fx["dims"], O.array("solid"))
        assert np.array_equal(L.f[m], O.array("f")[m]), f"populations differ at step {stop}"
    if "revox" in fx:
        replay_voxelize(L, fx["revox"])
        L.step(30)
        for k in FIELDS:
            assert sha(L.fields()[k]) == fx["snaps2"]["60"][k]


def test_esoteric_fp32_equals_ab_fp32():
    fx = load("lbm_moving")
    out = {}
    for lay in ("ab", "esoteric"):
        L = make(fx, layout=lay)
        replay_voxelize(L, fx["voxelize"])
        f = L.step(fx["nsteps"])
        out[lay] = (f, {k: v.copy() for k, v in L.fields().items()})
    for k in FIELDS:
        assert np.array_equal(out["ab"][1][k], out["esoteric"][1][k]), k
    np.testing.assert_allclose(out["ab"][0], out["esoteric"][0], rtol=1e-12, atol=1e-15)


def test_esoteric_pool_slabs_and_f16():
    fx = load("lbm_pool3")
    L = LBM(*fx["dims"], arith="f64", layout="esoteric")
    set_params(L, fx["worker_params"])
    L.reset()
    replay_voxelize(L, fx["voxelize"])
    L.step_slabs(fx["nsteps"], fx["nWorkers"])
    for k in FIELDS:
        assert sha(L.fields()[k]) == fx["pooled"][k], k
    fx = load("lbm_trt_les")
    a = make(fx, storage="f16")
    b = make(fx, storage="f16", layout="esoteric")
    for L in (a, b):
        replay_voxelize(L, fx["voxelize"])
        L.step(40)
    for k in FIELDS:
        assert np.array_equal(a.fields()[k], b.fields()[k]), k


def sphere_lattice(dims, **kw):
    L = LBM(*dims, **kw)
    L.set_params(tau=0.52, les=True, csm=0.1, trt=True, u0=0.08, **{"lambda": 3 / 16})
    L.reset()
    L.voxelize_mask(*sphere_mask(dims))
    return L


@pytest.mark.parametrize("arith,storage", [("f64", "f32"), ("f64", "f64"), ("f32", "f32"), ("f32", "f16")])
def test_layouts_agree_bitwise(arith, storage, monkeypatch):
    """128x40x36 sphere at the production defaults (64-cell blocks for 4- and
    8-byte storage, 128 for binary16): planar, blocked and in-place lattices
    compute the same populations, fields and forces bit for bit, and in
    binary64 arithmetic they equal the oracle (float or double storage)."""
    dims = (128, 40, 36)
    runs = {}
    for shift in (0, 7, 6, 4):
        for layout in ("ab", "esoteric"):
            monkeypatch.setenv("WT_BLOCK_SHIFT", str(shift))
            L = sphere_lattice(dims, arith=arith, storage=storage, layout=layout)
            expect_layout(L, shift)
            print(f"[layouts] {arith}/{storage} shift={shift} {layout}", flush=True)
            f = L.step(37)
            runs[(shift, layout)] = (f, {k: v.copy() for k, v in L.fields().items()}, L.solid.copy())
            L.close()
    base = runs[(0, "ab")]
    for key, (f, F, _) in runs.items():
        for k in FIELDS:
            assert np.array_equal(F[k], base[1][k]), (key, k)
        np.testing.assert_allclose(f, base[0], rtol=1e-12, atol=1e-15, err_msg=str(key))
    if arith == "f64":
        monkeypatch.delenv("WT_BLOCK_SHIFT")
        L = sphere_lattice(dims, arith=arith, storage=storage)
        assert L.block_cells == 64
        O = OracleLBM(*dims, prec=8 if storage == "f64" else 4)
        O.set_params(tau=0.52, les=True, csm=0.1, trt=True, u0=0.08, **{"lambda": 3 / 16})
        O.reset()
        O.voxelize_mask(*sphere_mask(dims))
        assert np.array_equal(O.array("solid"), base[2])
        fo = O.step(37)
        L.step(37)
        assert np.array_equal(L.f, O.array("f"))
        for k in FIELDS:
            assert np.array_equal(base[1][k], O.array(k)), k
        np.testing.assert_allclose(base[0], fo, rtol=1e-11, atol=1e-14)


def sphere_mask(dims):
    """Off-centre sphere (radius 0.22 ny) at cell centres, over its box."""
    nx, ny, nz = dims
    c = np.array([0.3 * nx, ny / 2 - 0.3, nz / 2 + 0.2])
    R = 0.22 * ny
    lo = np.floor(c - R).astype(int) - 1
    hi = np.ceil(c + R).astype(int) + 1
    z, y, x = np.meshgrid(*(np.arange(lo[k], hi[k] + 1) for k in (2, 1, 0)), indexing="ij")
    m = ((x + 0.5 - c[0]) ** 2 + (y + 0.5 - c[1]) ** 2 + (z + 0.5 - c[2]) ** 2 <= R * R).astype(np.uint8)
    return [int(lo[0]), int(lo[1]), int(lo[2]), int(hi[0]), int(hi[1]), int(hi[2])], m.reshape(-1)


@pytest.mark.parametrize("storage,layout", [("f32", "ab"), ("f64", "ab"), ("f16", "ab"), ("f32", "esoteric")])
def test_set_pops_round_trip_and_restore(storage, layout):
    """wt_set_pops inverts wt_get_pops (binary16: rounded to nearest), and a
    state saved mid-run and restored into a fresh lattice continues bit for
    bit like the run it was taken from."""
    fx = load("lbm_bgk_sphere")
    arith = "f64" if storage == "f64" else "f32"
    A = make(fx, storage=storage, arith=arith, layout=layout)
    replay_voxelize(A, fx["voxelize"])
    A.step(40)
    saved, parity = A.f.copy(), A.parity
    B = make(fx, storage=storage, arith=arith, layout=layout)
    replay_voxelize(B, fx["voxelize"])
    if B.parity != parity:
        B.parity = parity
    B.set_pops(saved)
    assert np.array_equal(B.f, saved)
    fa, fb = A.step(25), B.step(25)
    if storage != "f16":  # binary16 rounding draws on the step count, which a fresh lattice restarts
        assert np.array_equal(fa, fb)
        assert np.array_equal(A.f, B.f)
    assert np.all(np.isfinite(fb))
    rng = np.random.default_rng(7)
    dt = np.float64 if storage == "f64" else np.float32
    r = (rng.random(19 * A.n) * 2e-3 - 1e-3).astype(dt)
    A.set_pops(r)
    got = A.f
    if layout == "esoteric":
        keep = got != 0
        r = np.where(keep, r, 0)
    if storage == "f16":
        np.testing.assert_allclose(got, r, rtol=2 ** -11, atol=2 ** -24 / 32768)
    else:
        assert np.array_equal(got, r)
