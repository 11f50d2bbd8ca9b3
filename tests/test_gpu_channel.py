"""Channel boundaries on the GPU (free-slip side walls, equilibrium outflow:
the 96x48x48 configuration's boundaries; no reference function) against
the oracle's channel mode, which is pinned by the mirror and uniform-flow
tests of tests/test_oracle_channel.py: bit for bit in binary64 arithmetic
(both population layouts, float and double storage), within 1e-4 in float
arithmetic, and the calls that refuse the mode."""
import numpy as np
import pytest

from _util import arr, load, rel_l2, replay_voxelize, set_params, sha
from oracle.oracle import OracleLBM
from windtunnel_amd.lbm import LBM, WindtunnelError

pytestmark = pytest.mark.gpu


@pytest.fixture
def block_shift(request, monkeypatch):
    monkeypatch.setenv("WT_BLOCK_SHIFT", str(request.param))
    return request.param


def c1(L, fx):
    set_params(L, fx["params"])
    L.set_boundary_mode("channel")
    L.reset()
    replay_voxelize(L, fx["voxelize"])
    return L


@pytest.mark.parametrize("block_shift", [0, 4], indirect=True)
def test_exact_mode_equals_oracle(block_shift):
    fx = load("lbm_c1")
    print(f"[channel] exact mode, block shift {block_shift}", flush=True)
    O = c1(OracleLBM(*fx["dims"]), fx)
    G = c1(LBM(*fx["dims"], arith="f64"), fx)
    for n in (1, 199):
        fo, fg = O.step(n), G.step(n)
        np.testing.assert_allclose(fg, fo, rtol=1e-11, atol=1e-14)
        assert sha(G.f) == sha(O.array("f")) and sha(G.g) == sha(O.array("g"))
        F = G.fields()
        for k in ("rho", "ux", "uy", "uz"):
            assert sha(F[k]) == sha(O.array(k)), k
    # the channel run differs from the reference boundaries' run
    R = LBM(*fx["dims"], arith="f64")
    set_params(R, fx["params"])
    R.reset()
    replay_voxelize(R, fx["voxelize"])
    R.step(200)
    assert sha(R.fields()["ux"]) != sha(G.fields()["ux"])


def test_double_storage_equals_oracle():
    fx = load("lbm_c1")
    O = c1(OracleLBM(*fx["dims"], prec=8), fx)
    G = c1(LBM(*fx["dims"], storage="f64", arith="f64"), fx)
    fo, fg = O.step(150), G.step(150)
    np.testing.assert_allclose(fg, fo, rtol=1e-11, atol=1e-14)
    assert np.array_equal(G.f, O.array("f"))


@pytest.mark.parametrize("storage,tol", [("f32", 1e-4), ("f16", 3e-2)])
def test_float_arithmetic_tracks_oracle(storage, tol):
    fx = load("lbm_c1")
    O = c1(OracleLBM(*fx["dims"]), fx)
    G = c1(LBM(*fx["dims"], storage=storage), fx)
    fo, fg = O.step(400), G.step(400)
    F = G.fields()
    errs = {k: rel_l2(F[k], O.array(k)) for k in ("ux", "uy", "uz")}
    errs["cd"] = abs(np.mean(fg[-100:, 0]) / np.mean(fo[-100:, 0]) - 1)
    print(f"[channel] {storage}: {errs}", flush=True)
    assert all(np.isfinite(fg).ravel())
    assert max(errs.values()) < tol


def test_refusals():
    L = LBM(24, 10, 10, layout="esoteric")
    with pytest.raises(ValueError, match="AB layout"):
        L.set_boundary_mode("channel")
    with pytest.raises(ValueError):
        L.set_boundary_mode("periodic")
    A = LBM(24, 10, 10)
    A.set_boundary_mode("channel")
    with pytest.raises(WindtunnelError) as e:
        A.step_slab(1, 9)
    assert e.value.code == -4
    with pytest.raises(ValueError, match="reference boundaries"):
        A.adjoint(4, np.zeros((4, 3)))
    A.set_boundary_mode("reference")
    A.step_slab(1, 9)
