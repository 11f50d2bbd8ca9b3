"""Analytic bodies of the app's catalogue rasterised on the GPU
(wt_voxelize_parts) against the reference's own Body.makeTest /
makeWheelVel + LBM.voxelize on the same placements: solid map, near-wall
map, wall velocities (rolling wheels and road), populations and body
statistics bit for bit."""
import numpy as np
import pytest

from _util import load, sha
from windtunnel_amd.lbm import LBM

pytestmark = pytest.mark.gpu

FX = load("lbm_bodies")
KEYS = [c["key"] for c in FX["cases"]]


@pytest.mark.parametrize("key", KEYS)
def test_analytic_body_bit_exact(key):
    c = next(c for c in FX["cases"] if c["key"] == key)
    b = c["body"]
    L = LBM(*FX["dims"], arith="f64")
    L.u0 = 0.06
    L.reset()
    print(f"[body] {key} parts={len(b['parts'])} rolling={c['rolling_u0']}", flush=True)
    L.voxelize_parts(b["parts"], np.array(b["M"]), np.array(b["pos"]), b["scale"], c["box"], c["ground"],
                     c["roadVel"], c["rolling_u0"])
    want = c["after"]
    s, nr = L.solid_near()
    assert sha(s) == want["solid"] and sha(nr) == want["near"]
    uw = L.walls()
    assert [sha(uw[k]) for k in range(3)] == [want["uwx"], want["uwy"], want["uwz"]]
    assert sha(L.f) == want["f"] and sha(L.g) == want["g"]
    F = L.fields()
    for k in ("rho", "ux", "uy", "uz"):
        assert sha(F[k]) == want[k], k
    r = c["result"]
    assert (L.frontalArea, L.planformArea, L.solidCells, L.rideHeight, L.movingWalls) == (
        r["frontalArea"], r["planformArea"], r["solidCells"], r["rideHeight"], r["movingWalls"])


def test_parts_argument_errors():
    L = LBM(24, 12, 12)
    part = dict(t="box", c=[0, 0, 0], s=[0.1, 0.1, 0.1])
    with pytest.raises(ValueError):
        L.voxelize_parts([part] * 65, np.eye(3).reshape(-1), [12, 6, 6], 10, [1, 1, 1, 22, 10, 10])
    with pytest.raises(ValueError):
        L.voxelize_parts([part], np.eye(3).reshape(-1), [12, 6, 6], 0, [1, 1, 1, 22, 10, 10])


WHEELED = [c["key"] for c in FX["cases"] if c["rolling_u0"] and c["ground"]]


@pytest.mark.parametrize("key", WHEELED[:2])
def test_dragged_wheeled_body_device_equals_host(key, monkeypatch):
    """A ground vehicle with spinning wheels on a rolling road, dragged and
    then stopped (wheels and road at rest): after the first placement every
    move keeps the road's solid state, so it runs wholly on the device (the
    diff of wall velocities included); the same moves with the host
    bookkeeping (WT_VOXELIZE_HOST=1) must give the same lattice bit for bit."""
    c = next(c for c in FX["cases"] if c["key"] == key)
    b = c["body"]
    pos0 = np.array(b["pos"], dtype=np.float64)
    moves = [(0.0, 0.0, c["roadVel"], c["rolling_u0"]), (1.0, 0.0, c["roadVel"], c["rolling_u0"]),
             (2.0, 1.0, c["roadVel"], c["rolling_u0"]), (2.0, 1.0, 0.0, 0.0), (-1.0, 0.0, c["roadVel"], c["rolling_u0"])]
    lat = []
    for _ in range(2):
        L = LBM(*FX["dims"], arith="f64")
        L.u0 = 0.06
        L.reset()
        lat.append(L)
    D, H = lat
    for k, (dx, dy, road, rolling) in enumerate(moves):
        print(f"[wheeled drag] {key} move {k}: dx={dx} dy={dy} road={road}", flush=True)
        pos = pos0 + np.array([dx, dy, 0.0])
        box = dict(c["box"]) if isinstance(c["box"], dict) else list(c["box"])
        if isinstance(box, dict):
            box = {kk: v + (dx if kk.startswith("x") else (dy if kk.startswith("y") else 0)) for kk, v in box.items()}
        else:
            box = [box[0] + dx, box[1] + dy, box[2], box[3] + dx, box[4] + dy, box[5]]
        monkeypatch.delenv("WT_VOXELIZE_HOST", raising=False)
        D.voxelize_parts(b["parts"], np.array(b["M"]), pos, b["scale"], box, c["ground"], road, rolling)
        monkeypatch.setenv("WT_VOXELIZE_HOST", "1")
        H.voxelize_parts(b["parts"], np.array(b["M"]), pos, b["scale"], box, c["ground"], road, rolling)
        monkeypatch.delenv("WT_VOXELIZE_HOST", raising=False)
        assert (D.frontalArea, D.planformArea, D.solidCells, D.rideHeight, D.movingWalls) == (
            H.frontalArea, H.planformArea, H.solidCells, H.rideHeight, H.movingWalls), k
        sd, nd = D.solid_near()
        sh, nh = H.solid_near()
        assert sha(sd) == sha(sh) and sha(nd) == sha(nh), k
        wd, wh = D.walls(), H.walls()
        assert [sha(w) for w in wd] == [sha(w) for w in wh], k
        np.testing.assert_array_equal(D.step(2), H.step(2))
        assert sha(D.f) == sha(H.f) and sha(D.g) == sha(H.g), k
    D.close()
    H.close()
