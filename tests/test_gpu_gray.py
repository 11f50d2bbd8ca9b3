"""Gray (partial) bounce-back on the GPU, checked against what the scheme's
mathematics fixes and against the oracle.

The link (i <- j = i - c_q) transmits t = (1 - w_i)(1 - w_j):
* a closed disturbance (fluid at rest around it, random gray field, a solid
  block) keeps its mass, and its momentum changes by minus the reported
  force every step;
* in a lattice gray everywhere with weight w, a uniform flow decays by the
  factor 2t - 1, t = (1 - w)^2, per step;
* in binary64 arithmetic every population equals the oracle's bit for bit.
"""
import numpy as np
import pytest

from oracle.oracle import OracleLBM
from windtunnel_amd.lbm import LBM

pytestmark = pytest.mark.gpu

EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
DIMS = (24, 22, 20)
T = 6


def closed(L, seed=11):
    """Fluid at rest with a random gray field, random population offsets and a
    SOLID_OBJ block, all in a central box that nothing leaves within T steps.
    Works on the oracle and on the GPU lattice alike; returns the offsets."""
    nx, ny, nz = DIMS
    L.set_params(tau=0.58, les=True, trt=True, csm=0.14, u0=0.0, **{"lambda": 3 / 16})
    L.reset()
    lo = np.array([T + 2] * 3)
    hi = np.array(DIMS) - T - 2
    c = (lo + hi) // 2
    L.voxelize_mask([int(c[0]), int(c[1]), int(c[2]), int(c[0]) + 1, int(c[1]) + 2, int(c[2]) + 1],
                    np.ones(12, np.uint8))
    rng = np.random.default_rng(seed)
    w = np.zeros((nz, ny, nx), np.float32)
    sub = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
    g = rng.uniform(0.0, 0.9, size=w[sub].shape).astype(np.float32)
    g[rng.random(g.shape) < 0.3] = 0.0
    w[sub] = g
    L.set_gray(w.reshape(-1))
    f = np.zeros((19, nz, ny, nx))
    for q in range(19):
        f[q][sub] = rng.uniform(-1e-3, 1e-3, size=g.shape)
    return f


def totals(f, solid):
    f = f.reshape(19, -1).astype(np.float64)[:, solid.reshape(-1) == 0]
    s = f.sum(1)
    return f.sum(), np.array([EX @ s, EY @ s, EZ @ s])


@pytest.mark.parametrize("storage,arith,tol", [("f64", "f64", 1e-15), ("f32", "f64", 1e-7), ("f32", "f32", 1e-7)])
def test_closed_gray_disturbance_conserves(storage, arith, tol):
    print(f"[gray conservation] {storage}/{arith}", flush=True)
    G = LBM(*DIMS, storage=storage, arith=arith)
    f = closed(G)
    solid, _ = G.solid_near()
    f = np.where(solid.reshape(1, -1) == 0, f.reshape(19, -1), G.pops(0).reshape(19, -1))
    G.set_pops(f.astype(np.float64 if storage == "f64" else np.float32))
    m0, p0 = totals(G.pops(0), solid)
    for _ in range(T):
        F = G.step(1)[0]
        m1, p1 = totals(G.pops(0), solid)
        assert abs(m1 - m0) < tol and np.abs((p1 - p0) + F).max() < tol, (m1 - m0, (p1 - p0) + F)
        m0, p0 = m1, p1
    G.close()


def test_closed_gray_disturbance_equals_oracle_bitwise():
    O = OracleLBM(*DIMS, prec=8)
    G = LBM(*DIMS, storage="f64", arith="f64")
    f = closed(O)
    closed(G)
    fo = O.array("f").reshape(19, -1)
    fluid = O.array("solid") == 0
    fo[:, fluid] = f.reshape(19, -1)[:, fluid]
    G.set_pops(O.array("f").copy())
    assert np.array_equal(G.solid_near()[1], O.array("near"))
    fF, oF = G.step(T), O.step(T)
    assert np.array_equal(G.pops(0), O.array("f"))
    np.testing.assert_allclose(fF, oF, rtol=1e-12, atol=1e-18)
    G.close()


@pytest.mark.parametrize("storage,arith,tol", [("f64", "f64", 1e-15), ("f32", "f32", 1e-7), ("f16", "f32", 2e-4)])
@pytest.mark.parametrize("w", [0.1, 0.37])
def test_uniform_gray_medium_decay(storage, arith, tol, w):
    print(f"[gray decay] {storage}/{arith} w={w}", flush=True)
    nx = ny = nz = 20
    u0 = 0.05
    G = LBM(nx, ny, nz, storage=storage, arith=arith)
    G.set_params(tau=0.6, les=False, trt=True, csm=0.14, u0=u0, **{"lambda": 3 / 16})
    G.reset()
    G.set_gray(np.full(G.n, w, np.float32))
    t = (1 - np.float64(np.float32(w))) ** 2
    c = nx // 2 + nx * (ny // 2 + ny * (nz // 2))
    for k in range(1, 6):
        G.step(1)
        fl = G.fields()
        assert abs(fl["ux"][c] - u0 * (2 * t - 1) ** k) < tol, (k, fl["ux"][c], u0 * (2 * t - 1) ** k)
        assert abs(fl["uy"][c]) < tol and abs(fl["uz"][c]) < tol and abs(fl["rho"][c] - 1) < tol
    G.close()
