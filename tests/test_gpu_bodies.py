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
