"""The oracle's gray (partial) bounce-back, pinned by what the mathematics of
the scheme fixes, not by the reference (which has no gray cells).

A link between fluid cells i and j = i - c_q transmits t = (1 - w_i)(1 - w_j)
of what streams along it and bounces the rest back halfway:
    f_q(i) = t g_q(j) + (1 - t) g_qbar(i),   F -= (1 - t) 2 g_qbar(i) c_q.
Three consequences hold exactly and owe nothing to the code under test:

* mass: cell j takes the same t for the opposite link, so every pair of links
  conserves mass, and a closed disturbance keeps its total mass;
* momentum: the reported force is the momentum the fluid loses, so the fluid
  momentum changes by -F every step (solid links included);
* uniform medium: in a lattice gray everywhere with one weight w, a uniform
  flow stays uniform and its momentum is multiplied by 2t - 1, t = (1 - w)^2,
  every step (the collision conserves momentum; streaming returns t of each
  population and bounces 1 - t of its opposite).

Each check is also run against variants of the oracle with a mistake planted
(force sign, one-sided link weight, swapped mixing weights), which must fail
at least one of them.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleLBM

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(HERE, "..", "oracle")
EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])


def closed_lattice(libpath=None, dims=(24, 22, 20), T=6, seed=11, solid=True, prec=8):
    """Fluid at rest (u0 = 0) with a random gray field and random population
    offsets confined to a central box, and a small SOLID_OBJ block inside
    it. Nothing reaches the boundary planes within T steps, so the
    disturbance is a closed system."""
    nx, ny, nz = dims
    L = OracleLBM(nx, ny, nz, prec=prec, libpath=libpath)
    L.set_params(tau=0.58, les=True, trt=True, csm=0.14, u0=0.0, **{"lambda": 3 / 16})
    L.reset()
    lo = np.array([T + 2, T + 2, T + 2])
    hi = np.array([nx, ny, nz]) - T - 2
    if solid:
        c = (lo + hi) // 2
        box = [int(c[0]), int(c[1]), int(c[2]), int(c[0]) + 1, int(c[1]) + 2, int(c[2]) + 1]
        L.voxelize_mask(box, np.ones(2 * 3 * 2, np.uint8))
    rng = np.random.default_rng(seed)
    w = np.zeros((nz, ny, nx), np.float32)
    sub = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
    g = rng.uniform(0.0, 0.9, size=w[sub].shape).astype(np.float32)
    g[rng.random(g.shape) < 0.3] = 0.0          # fluid cells among the gray ones
    w[sub] = g
    L.set_gray(w.reshape(-1))
    f = L.array("f").reshape(19, nz, ny, nx)
    fluid = L.array("solid").reshape(nz, ny, nx) == 0
    for q in range(19):
        d = np.zeros((nz, ny, nx))
        d[sub] = rng.uniform(-1e-3, 1e-3, size=g.shape)
        f[q] = np.where(fluid, d, f[q])
    return L


def totals(L):
    """Mass and momentum of the fluid's stored populations (offsets f - w:
    the weights cancel in both, since w_q = w_qbar)."""
    f = L.array("f").reshape(19, -1).astype(np.float64)
    fluid = L.array("solid") == 0
    f = f[:, fluid]
    return f.sum(), np.array([EX @ f.sum(1), EY @ f.sum(1), EZ @ f.sum(1)])


def conservation_errors(libpath=None, T=6, prec=8):
    L = closed_lattice(libpath, T=T, prec=prec)
    m0, p0 = totals(L)
    dm, dp = 0.0, 0.0
    for _ in range(T):
        F = L.step(1)[0]
        m1, p1 = totals(L)
        dm = max(dm, abs(m1 - m0))
        dp = max(dp, float(np.abs((p1 - p0) + F).max()))
        m0, p0 = m1, p1
    return dm, dp


def uniform_decay_error(libpath=None, w=0.1, T=5, dims=(20, 20, 20)):
    """max |u(centre) - u0 (2t - 1)^k| over steps k, t = (1 - w)^2."""
    nx, ny, nz = dims
    L = OracleLBM(nx, ny, nz, prec=8, libpath=libpath)
    u0 = 0.05
    L.set_params(tau=0.6, les=False, trt=True, csm=0.14, u0=u0, **{"lambda": 3 / 16})
    L.reset()
    L.set_gray(np.full(L.n, w, np.float32))   # the boundary planes keep weight 0
    t = (1 - np.float64(np.float32(w))) ** 2
    c = nx // 2 + nx * (ny // 2 + ny * (nz // 2))
    err = 0.0
    for k in range(1, T + 1):
        L.step(1)
        ux, uy, uz, rho = (L.array(a)[c] for a in ("ux", "uy", "uz", "rho"))
        err = max(err, abs(ux - u0 * (2 * t - 1) ** k), abs(uy), abs(uz), abs(rho - 1))
    return err


def test_gray_links_conserve_mass_and_momentum():
    dm, dp = conservation_errors()
    assert dm < 1e-15 and dp < 1e-15, (dm, dp)


def test_gray_links_float_storage_within_rounding():
    """Float32 storage: the same balance to the rounding of the stores."""
    dm, dp = conservation_errors(prec=4)
    assert dm < 1e-7 and dp < 1e-7, (dm, dp)


@pytest.mark.parametrize("w", [0.1, 0.05, 0.37])
def test_uniform_gray_medium_decays_by_2t_minus_1(w):
    assert uniform_decay_error(w=w) < 1e-15


def test_gray_cells_are_near_cells():
    """Every link of a gray cell is gray: the cell itself takes the near-wall
    path, as do the fluid cells next to it."""
    L = OracleLBM(10, 10, 10, prec=8)
    L.reset()
    w = np.zeros((10, 10, 10), np.float32)
    w[5, 5, 5] = 0.5
    L.set_gray(w.reshape(-1))
    near = L.array("near").reshape(10, 10, 10)
    exp = np.zeros_like(near)
    for q in range(19):
        exp[5 + EZ[q], 5 + EY[q], 5 + EX[q]] = 1
    assert np.array_equal(near, exp)


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    out = tmp_path_factory.mktemp("planted")
    libs = {}
    for k in (1, 2, 3):
        so = str(out / f"libwt_oracle_plant{k}.so")
        subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                        f"-DWTO_PLANT_GRAY={k}", "-o", so, os.path.join(ORACLE_DIR, "wt_oracle.c"), "-lm",
                        "-pthread"], check=True)
        libs[k] = so
    return libs


@pytest.mark.parametrize("k", [1, 2, 3])
def test_planted_mistakes_fail_a_check(planted, k):
    dm, dp = conservation_errors(planted[k])
    decay = uniform_decay_error(planted[k])
    assert max(dm, dp) > 1e-8 or decay > 1e-8, (k, dm, dp, decay)
