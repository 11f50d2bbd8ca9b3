"""Time-averaged drag of an LES run on the float paths against binary64.

The wake of the synthetic body (ellipsoid unioned with a box, ~20k
triangles, half scale) at 128x64x64 with TRT (Lambda 3/16) and Smagorinsky
LES at Cs = 0.1, tau = 0.55, u0 = 0.05: the eddy viscosity is active and the
wake settles, so the time average compares directly. C_d is the app's
statistic (js/main.js:113-127): pair-averaged coefficients over the last
2000 of 4000 steps. Float32 storage with float32 arithmetic and binary16
offsets with float32 arithmetic must be within 0.5 % of binary64 storage
with binary64 arithmetic (the oracle's numerics).
"""
import numpy as np
import pytest

from windtunnel_amd.lbm import LBM, Voxel
from windtunnel_amd.stats import block_stats, pair_coefficients
from windtunnel_amd.synth import ellipsoid_box

pytestmark = pytest.mark.gpu

DIMS = (128, 64, 64)
STEPS, WINDOW = 4000, 2000


def place_body(L, scale=0.5):
    vox = Voxel.build(ellipsoid_box(), 200)
    pos = np.array([0.3 * L.nx, L.ny / 2, L.nz / 2])
    b = np.array(vox["bbox"])
    lo, hi = pos + scale * b[:3] - 2, pos + scale * b[3:] + 2
    L.voxelize_grid(vox, np.eye(3).reshape(-1), pos, scale,
                    dict(x0=lo[0], y0=lo[1], z0=lo[2], x1=hi[0], y1=hi[1], z1=hi[2]))


def mean_cd(storage, arith):
    L = LBM(*DIMS, storage=storage, arith=arith)
    L.set_params(tau=0.55, les=True, csm=0.1, trt=True, u0=0.05, **{"lambda": 3 / 16})
    L.reset()
    place_body(L)
    F = L.step(STEPS)
    assert np.isfinite(F).all()
    bs = block_stats(pair_coefficients(F[-WINDOW:], 0.05, L.frontalArea)[0])
    L.close()
    return bs["mean"]


def test_les_time_averaged_cd_float_paths_within_half_percent():
    print("[les drag] binary64", flush=True)
    ref = mean_cd("f64", "f64")
    assert 0.2 < ref < 5
    for storage in ("f32", "f16"):
        print(f"[les drag] {storage}", flush=True)
        cd = mean_cd(storage, "f32")
        err = abs(cd - ref) / abs(ref)
        print(f"[les drag] {storage}: C_d {cd:.6f} vs {ref:.6f}, rel {err:.2e}", flush=True)
        assert err < 5e-3, (storage, cd, ref)
