"""Slice view and batched sample on the device against the oracle: the
slice restatement of js/main.js:925-965 over the downloaded fields (bit
for bit), and the oracle's sample() (js/lbm.js:515-539) at random points
of a binary64-arithmetic lattice whose fields equal the oracle's."""
import numpy as np
import pytest

from _util import load, replay_voxelize, set_params
from oracle.oracle import OracleLBM, slice_view
from windtunnel_amd.lbm import LBM

pytestmark = pytest.mark.gpu


def sphere_run(storage="f32", arith="f32", steps=150):
    fx = load("lbm_bgk_sphere")
    L = LBM(*fx["dims"], storage=storage, arith=arith)
    set_params(L, fx["params"])
    L.reset()
    replay_voxelize(L, fx["voxelize"])
    L.step(steps)
    return fx, L


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_slice_equals_restatement(storage):
    fx, L = sphere_run(storage, "f64" if storage == "f64" else "f32")
    dims = tuple(fx["dims"])
    F = L.fields()
    solid = L.solid
    for axis, k in (("xy", dims[2] // 2), ("xz", dims[1] // 2), ("xy", 0), ("xz", dims[1] - 1)):
        for field in ("speed", "pressure", "vorticity"):
            got, sol = L.slice(axis, k, field)
            want = slice_view(F, solid, dims, axis, k, field, L.u0)
            assert np.array_equal(got, want), (axis, k, field, np.abs(got - want).max())
            plane = solid.reshape(dims[::-1])[k] if axis == "xy" else solid.reshape(dims[::-1])[:, k]
            assert np.array_equal(sol, plane)
    with pytest.raises(ValueError):
        L.slice("xy", dims[2], "speed")


def test_sample_many_equals_oracle():
    fx = load("lbm_bgk_sphere")
    dims = tuple(fx["dims"])
    O = OracleLBM(*dims)
    set_params(O, fx["params"])
    O.reset()
    replay_voxelize(O, fx["voxelize"])
    O.step(60)
    _, L = sphere_run("f32", "f64", 60)
    rng = np.random.default_rng(11)
    P = rng.uniform(-1.0, 1.0, (2000, 3)) * (np.array(dims) + 2) / 2 + np.array(dims) / 2
    P[:5] = [[0.5, 0.5, 0.5], [dims[0] - 1.5, dims[1] - 1.5, dims[2] - 1.5], [0.49, 3, 3], [np.nan, 3, 3],
             [3, 3, dims[2] - 1.49]]
    got = L.sample_many(P)
    want = np.array([O.sample(*p) for p in P])
    assert np.array_equal(got, want, equal_nan=True)
    assert np.array_equal(L.sample_many(np.zeros((0, 3))), np.zeros((0, 3)))
    # the one-point sample() of the binding is the same device sample
    for p in P[:8]:
        assert np.array_equal(np.array(L.sample(*p)), np.array(O.sample(*p)), equal_nan=True)
