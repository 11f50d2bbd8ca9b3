"""A sequence of re-voxelizes on one lattice (body moved, ground and road
velocity toggled, wall-velocity cells added and removed, an empty box)
against the oracle, whose voxelize scans the whole lattice as
js/lbm.js:191-266 does: the library keeps only the previous body's box and
touches only the changed box (on the device, or through the host mirrors
when the ground plane is switched), and must end in the same state bit for
bit (populations, fields, solid and near maps, wall velocities, body
statistics) after every call and after steps in between."""
import numpy as np
import pytest

from _util import sha
from oracle.oracle import OracleLBM
from windtunnel_amd.lbm import LBM

pytestmark = pytest.mark.gpu

DIMS = (40, 20, 18)


def sphere_job(c, r, wheel=False):
    lo = [int(np.floor(c[k] - r)) - 1 for k in range(3)]
    hi = [int(np.ceil(c[k] + r)) + 1 for k in range(3)]
    lo = [max(1, v) for v in lo]
    hi = [min(DIMS[k] - 2, hi[k]) for k in range(3)]
    z, y, x = np.meshgrid(np.arange(lo[2], hi[2] + 1) + 0.5, np.arange(lo[1], hi[1] + 1) + 0.5,
                          np.arange(lo[0], hi[0] + 1) + 0.5, indexing="ij")
    inside = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r
    mask = inside.astype(np.uint8)
    wv = np.zeros(0, np.float32)
    if wheel:  # the lower half spins: wall velocity cells
        low = inside & (y < c[1])
        mask[low] = 2
        k = int(low.sum())
        wv = np.tile(np.array([0.02, 0.0, -0.01], np.float32), k)
    return lo + hi, mask.reshape(-1), wv


JOBS = [
    (sphere_job((12, 8, 9), 3.5), False, 0.0),
    (sphere_job((14, 8, 9), 3.5), True, 0.03),           # moved, ground with a moving road
    (sphere_job((14, 9, 9), 4.0, wheel=True), True, 0.03),
    (sphere_job((20, 10, 8), 3.0, wheel=True), False, 0.0),
    ([5, 5, 5, 4, 4, 4], False, 0.0),                     # empty box: the body leaves
    (sphere_job((25, 6, 10), 3.0), True, 0.0),
]


def state(L, oracle):
    if oracle:
        keys = ("f", "g", "rho", "ux", "uy", "uz", "solid", "near", "uwx", "uwy", "uwz")
        return {k: sha(L.array(k)) for k in keys}
    F = L.fields()
    s, nr = L.solid_near()
    out = {"f": sha(L.f), "g": sha(L.g), "solid": sha(s), "near": sha(nr)}
    out.update({k: sha(F[k]) for k in ("rho", "ux", "uy", "uz")})
    out.update(dict(zip(("uwx", "uwy", "uwz"), (sha(w) for w in L.walls()))))
    return out


def test_revoxelize_sequence_equals_oracle():
    O = OracleLBM(*DIMS)
    G = LBM(*DIMS, arith="f64")
    for L in (O, G):
        L.set_params(tau=0.6, les=True, trt=True, csm=0.14, u0=0.05, **{"lambda": 3 / 16})
        L.reset()
    for j, (job, ground, road) in enumerate(JOBS):
        box, mask, wv = job if isinstance(job, tuple) else (job, np.zeros(0, np.uint8), np.zeros(0, np.float32))
        print(f"[revox] job {j}: box {box} ground {ground} road {road}", flush=True)
        so = O.voxelize_mask(box, mask, wv, ground, road)
        G.voxelize_mask(box, mask, wv, ground, road)
        assert (G.frontalArea, G.planformArea, G.solidCells, G.rideHeight) == (
            so["frontalArea"], so["planformArea"], so["solidCells"], so["rideHeight"]), j
        assert bool(G.movingWalls) == bool(so["movingWalls"]), j
        assert state(G, False) == state(O, True), j
        fo, fg = O.step(7), G.step(7)
        np.testing.assert_allclose(fg, fo, rtol=1e-11, atol=1e-14)
        assert state(G, False) == state(O, True), j


def test_dragged_mesh_body_matches_a_fresh_voxelize():
    """The mesh body dragged through five positions on one 128x64x64 lattice
    (changed boxes uploaded, masks and the near-wall list rebuilt over the
    changed planes only) ends in the same solid and near maps, wall
    velocities, statistics and step forces as a lattice voxelized once at
    the final position."""
    from windtunnel_amd.lbm import Voxel
    from windtunnel_amd.synth import ellipsoid_box
    vox = Voxel.build(ellipsoid_box(), 200)
    b = np.array(vox["bbox"])

    def place(L, px, py):
        pos = np.array([px, py, 32.0])
        lo, hi = pos + b[:3] - 2, pos + b[3:] + 2
        L.voxelize_grid(vox, np.eye(3).reshape(-1), pos, 1.0,
                        dict(x0=lo[0], y0=lo[1], z0=lo[2], x1=hi[0], y1=hi[1], z1=hi[2]))

    dims = (128, 64, 64)
    A, B = LBM(*dims), LBM(*dims)
    for L in (A, B):
        L.set_params(tau=0.6, les=True, trt=True, csm=0.14, u0=0.05, **{"lambda": 3 / 16})
        L.reset()
    for k, (px, py) in enumerate([(40, 32), (41, 32), (43, 31), (47, 33), (46, 30)]):
        print(f"[drag] move {k} to ({px}, {py})", flush=True)
        place(A, px, py)
    place(B, 46, 30)
    assert (A.frontalArea, A.planformArea, A.solidCells, A.rideHeight) == (
        B.frontalArea, B.planformArea, B.solidCells, B.rideHeight)
    sa, na = A.solid_near()
    sb, nb = B.solid_near()
    assert sha(sa) == sha(sb) and sha(na) == sha(nb)
    for wa, wb in zip(A.walls(), B.walls()):
        assert sha(wa) == sha(wb)
    np.testing.assert_array_equal(A.step(3), B.step(3))
    fluid = np.asarray(sa).reshape(-1) == 0
    Fa, Fb = A.fields(), B.fields()
    for key in ("rho", "ux", "uy", "uz"):
        np.testing.assert_array_equal(np.asarray(Fa[key]).reshape(-1)[fluid], np.asarray(Fb[key]).reshape(-1)[fluid])
    A.close()
    B.close()


def test_device_revoxelize_equals_host_bookkeeping(monkeypatch):
    """Grid-backed bodies with no wall velocities are re-voxelized wholly on
    the device (diff, ops, projections, ride height); the same moves with
    the host bookkeeping (WT_VOXELIZE_HOST=1) must give the same lattice
    bit for bit: statistics, solid and near maps, populations after steps."""
    from windtunnel_amd.lbm import Voxel
    from windtunnel_amd.synth import ellipsoid_box
    vox = Voxel.build(ellipsoid_box(), 200)
    b = np.array(vox["bbox"])
    dims = (112, 48, 40)

    def place(L, pos, scale):
        pos = np.asarray(pos, dtype=np.float64)
        lo, hi = pos + scale * b[:3] - 2, pos + scale * b[3:] + 2
        L.voxelize_grid(vox, np.eye(3).reshape(-1), pos, scale,
                        dict(x0=lo[0], y0=lo[1], z0=lo[2], x1=hi[0], y1=hi[1], z1=hi[2]))

    lat = []
    for _ in range(2):
        L = LBM(*dims, arith="f64")
        L.set_params(tau=0.6, les=True, trt=True, csm=0.14, u0=0.05, **{"lambda": 3 / 16})
        L.reset()
        lat.append(L)
    D, H = lat
    moves = [((40, 24, 20), 0.8), ((41, 24, 20), 0.8), ((44, 22, 21), 0.8), ((44, 18, 21), 0.8),
             ((60, 25, 19), 0.6), ((58, 24, 20), 0.9)]
    for k, (pos, scale) in enumerate(moves):
        print(f"[device revox] move {k}: {pos} x{scale}", flush=True)
        monkeypatch.delenv("WT_VOXELIZE_HOST", raising=False)
        place(D, pos, scale)
        monkeypatch.setenv("WT_VOXELIZE_HOST", "1")
        place(H, pos, scale)
        monkeypatch.delenv("WT_VOXELIZE_HOST", raising=False)
        assert (D.frontalArea, D.planformArea, D.solidCells, D.rideHeight, bool(D.movingWalls)) == (
            H.frontalArea, H.planformArea, H.solidCells, H.rideHeight, bool(H.movingWalls)), k
        sd, nd = D.solid_near()
        sh, nh = H.solid_near()
        assert sha(sd) == sha(sh) and sha(nd) == sha(nh), k
        np.testing.assert_array_equal(D.step(3), H.step(3))
        assert sha(D.f) == sha(H.f) and sha(D.g) == sha(H.g), k
    D.close()
    H.close()
