"""wt_sample_noalloc, the batched sample that allocates no device memory,
called through the C ABI and checked bit for bit against wt_sample on the
same lattice: two calls in a row, the second with fewer points, in float
and double storage, and the size limit refused."""
import ctypes as C

import numpy as np
import pytest

from windtunnel_amd import lbm as wl
from windtunnel_amd.lbm import LBM

pytestmark = pytest.mark.gpu


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def lattice(storage):
    dims = (64, 24, 20)
    L = LBM(*dims, storage=storage, arith="f64" if storage == "f64" else "f32")
    L.set_params(tau=0.56, les=True, csm=0.1, trt=True, u0=0.05, **{"lambda": 3 / 16})
    L.reset()
    z, y, x = np.meshgrid(np.arange(5, 15), np.arange(7, 17), np.arange(15, 25), indexing="ij")
    mask = (((x + 0.5 - 20) ** 2 + (y + 0.5 - 12) ** 2 + (z + 0.5 - 10) ** 2) <= 16).astype(np.uint8)
    L.voxelize_mask([15, 7, 5, 24, 16, 14], mask.reshape(-1))
    L.step(25)
    return L


def points(n, seed):
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(-2, 66, n), rng.uniform(-2, 26, n), rng.uniform(-2, 22, n)], 1)
    p[::97] = np.nan
    return np.ascontiguousarray(p)


def both(L, p):
    lib = wl.lib()
    a = np.full_like(p, -7.0)
    b = np.full_like(p, -9.0)
    assert lib.wt_sample(L._h, _ptr(p), p.shape[0], _ptr(a)) == 0, lib.wt_last_error()
    rc = lib.wt_sample_noalloc(L._h, _ptr(p), p.shape[0], _ptr(b))
    assert rc == 0, lib.wt_last_error()
    return a, b


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_noalloc_sample_equals_sample(storage):
    print(f"[noalloc sample] {storage}", flush=True)
    L = lattice(storage)
    for n, seed in ((50000, 1), (777, 2)):
        p = points(n, seed)
        a, b = both(L, p)
        assert np.array_equal(a, b, equal_nan=True), n
        ok = np.isfinite(p).all(1)
        assert np.isfinite(a[ok]).all() and np.abs(a[ok]).max() > 0
    L.close()


def test_noalloc_sample_limit_and_arguments():
    L = lattice("f32")
    lib = wl.lib()
    big = np.zeros((262144 + 1, 3))
    out = np.zeros_like(big)
    assert lib.wt_sample_noalloc(L._h, _ptr(big), big.shape[0], _ptr(out)) == wl.E_ARG
    assert b"exceed" in lib.wt_last_error()
    # the largest allowed batch works and matches wt_sample
    p = points(262144, 3)
    a, b = both(L, p)
    assert np.array_equal(a, b, equal_nan=True)
    assert lib.wt_sample_noalloc(L._h, None, 0, None) == 0
    L.close()
