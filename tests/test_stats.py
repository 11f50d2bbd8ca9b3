"""Force statistics of the app (js/main.js:60-213), pinned by closed forms:
the pair average cancels a step-parity oscillation exactly, the blocking
method on block-constant data gives the textbook SEM of the block values,
short series fall back to sigma/sqrt(n), the ring keeps the last 6000
samples in order, and the convergence states follow their thresholds."""
import numpy as np
import pytest

from windtunnel_amd.stats import HIST, SETTLE, ForceStats, block_stats, coefficient_stats, pair_coefficients


# every test here also runs in the GPU suite (-m gpu), against the shipped build on the GPU machine
@pytest.fixture(params=[pytest.param("cpu"), pytest.param("gpu", marks=pytest.mark.gpu)], autouse=True)
def suite(request):
    return request.param



def test_pair_average_cancels_parity_oscillation():
    k = np.arange(400)
    F = np.stack([0.3 + 0.05 * (-1.0) ** k, -0.1 + 0.02 * (-1.0) ** k, 0 * k], 1)
    cd, cl, cy = pair_coefficients(F, 0.08, 120)
    q = 0.5 * 0.08 * 0.08 * 120
    assert cd.dtype == np.float32 and cd.size == 399
    assert np.all(cd == np.float32(0.3 / q)) and np.all(cl == np.float32(-0.1 / q))
    st = coefficient_stats(F, 0.08, 120, settle=0)
    assert st["cd"]["sem"] == 0 and st["state"] == "averaging"


def test_settle_and_first_sample_prime_the_pair():
    st = ForceStats()
    F = np.random.default_rng(1).normal(size=(SETTLE + 11, 3))
    st.accumulate_series(F, 0.05, 10)
    assert st.n == 10
    want = pair_coefficients(F[SETTLE:], 0.05, 10)[0]
    assert np.array_equal(st.ordered("cd"), want.astype(np.float64))


def test_ring_keeps_last_samples_in_order():
    st = ForceStats(settle=0)
    F = np.zeros((HIST + 501, 3))
    F[:, 0] = np.arange(HIST + 501)
    st.accumulate_series(F, 1.0, 2.0)       # q = 1: sample k = k + 0.5
    x = st.ordered("cd")
    assert x.size == HIST and st.n == HIST + 500
    assert np.array_equal(x, np.arange(500, HIST + 500, dtype=np.float32) + np.float32(0.5))


def test_block_stats_closed_forms():
    rng = np.random.default_rng(3)
    b = rng.normal(size=16)
    x = np.repeat(b, 10)
    r = block_stats(x)
    assert r["blocked"] and r["used"] == 160
    assert r["mean"] == pytest.approx(b.mean(), rel=1e-14)
    assert r["sem"] == pytest.approx(b.std(ddof=1) / 4, rel=1e-12)
    x2 = np.concatenate([x, rng.normal(size=9)])   # the tail beyond 16 blocks is dropped
    assert block_stats(x2)["mean"] == pytest.approx(b.mean(), rel=1e-14)
    y = rng.normal(size=50)
    r = block_stats(y)
    assert not r["blocked"]
    assert r["mean"] == pytest.approx(y.mean(), rel=1e-13)
    assert r["sem"] == pytest.approx(y.std(ddof=1) / np.sqrt(50), rel=1e-12)


def test_convergence_states():
    def run(cd):
        F = np.zeros((cd.size + 1, 3))
        F[:, 0] = np.concatenate([[cd[0]], cd])    # pairs (F_k + F_{k-1})/2 track cd
        return coefficient_stats(F, 1.0, 2.0, settle=0)
    assert run(np.ones(58)) is None
    assert run(np.ones(300))["state"] == "averaging"
    assert run(np.ones(1000))["state"] == "converged"
    assert run(np.linspace(1, 2, 1000))["state"] == "drifting"
    rng = np.random.default_rng(5)
    assert run(1 + 0.5 * rng.normal(size=1000))["state"] == "noisy"
