"""Force-coefficient statistics of the wind-tunnel app (js/main.js:60-213).

The momentum-exchange force alternates every step (lag-1 autocorrelation
near -1), so each completed step after a settling period contributes the
average of its force and the previous step's (a sliding pair average,
js/main.js:114-129), divided by the dynamic pressure 0.5 u0^2 A; the
history keeps the last HIST = 6000 such samples in float32
(js/main.js:65-68). The mean and its standard error come from 16 block
means (js/main.js:153-179), and the run is classified as averaging,
drifting, noisy or converged (js/main.js:186-207).
"""
from __future__ import annotations

import numpy as np

HIST = 6000      # samples kept (js/main.js:65)
BLOCKS = 16      # blocking-method block count (js/main.js:163)
SETTLE = 150     # steps skipped after a reset of the statistics (js/main.js:83)


def block_stats(x) -> dict:
    """Mean and standard error by the blocking method (js/main.js:153-179):
    16 block means of floor(n/16) samples each (the tail beyond 16 blocks is
    dropped); below 4 samples per block, the plain sample mean and
    sigma/sqrt(n)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    bl = n // BLOCKS
    if bl < 4:
        m = 0.0
        for v in x:
            m += v
        m /= n
        var = 0.0
        for v in x:
            d = v - m
            var += d * d
        return dict(mean=m, sem=float(np.sqrt(var / max(1, n - 1) / n)), blocked=False)
    bm = np.empty(BLOCKS)
    mean = 0.0
    for b in range(BLOCKS):
        s = 0.0
        for v in x[b * bl:(b + 1) * bl]:
            s += v
        bm[b] = s / bl
        mean += bm[b]
    mean /= BLOCKS
    var = 0.0
    for b in range(BLOCKS):
        d = bm[b] - mean
        var += d * d
    return dict(mean=mean, sem=float(np.sqrt(var / (BLOCKS - 1) / BLOCKS)), blocked=True, used=BLOCKS * bl)


def pair_coefficients(forces, u0: float, area: float, prev=None):
    """Sliding pair-averaged coefficients of consecutive step forces:
    sample k = 0.5 (F_k + F_{k-1}) / (0.5 u0^2 A), stored as float32
    (js/main.js:114-129). Without `prev` the first force only primes the
    pair, as after a reset; returns (cd, cl, cy)."""
    F = np.asarray(forces, dtype=np.float64).reshape(-1, 3)
    if prev is not None:
        F = np.vstack([np.asarray(prev, dtype=np.float64).reshape(1, 3), F])
    q = 0.5 * u0 * u0 * area
    s = 0.5 * (F[1:] + F[:-1]) / q
    s = s.astype(np.float32)
    return s[:, 0], s[:, 1], s[:, 2]


class ForceStats:
    """The app's running statistics for one set of test conditions."""

    def __init__(self, settle: int = SETTLE):
        self.reset(settle)

    def reset(self, settle: int = SETTLE) -> None:
        """Forget everything and skip the next `settle` steps (js/main.js:76-84)."""
        self.settle = int(settle)
        self.count = 0
        self.n = 0                     # samples taken (may exceed HIST)
        self.ema = np.zeros(3)
        self.prev = None
        self.cd = np.zeros(HIST, dtype=np.float32)
        self.cl = np.zeros(HIST, dtype=np.float32)
        self.cy = np.zeros(HIST, dtype=np.float32)

    def accumulate(self, force, u0: float, area: float) -> None:
        """Fold one completed step's force in (js/main.js:105-110)."""
        if self.settle > 0:
            self.settle -= 1
            return
        self.count += 1
        a = 1.0 / min(self.count, 500)
        f = np.asarray(force, dtype=np.float64)
        self.ema += (f - self.ema) * a
        if self.prev is None:
            self.prev = f.copy()
            return
        cd, cl, cy = pair_coefficients(f, u0, area, prev=self.prev)
        i = self.n % HIST
        self.cd[i], self.cl[i], self.cy[i] = cd[0], cl[0], cy[0]
        self.prev = f.copy()
        self.n += 1

    def accumulate_series(self, forces, u0: float, area: float) -> None:
        for f in np.asarray(forces, dtype=np.float64).reshape(-1, 3):
            self.accumulate(f, u0, area)

    def ordered(self, which: str = "cd") -> np.ndarray:
        """The kept samples in chronological order (js/main.js:147-153)."""
        src = getattr(self, which)
        n = min(self.n, HIST)
        s = 0 if self.n <= HIST else self.n % HIST
        return np.array([src[(s + i) % HIST] for i in range(n)], dtype=np.float64)

    def run_stats(self):
        """Mean/SEM of C_d, C_l, C_y, drift between the halves and the
        convergence state, or None below 60 samples (js/main.js:186-207)."""
        x = self.ordered("cd")
        n = x.size
        if n < 60:
            return None
        cd = block_stats(x)
        h = n >> 1
        m1 = 0.0
        for v in x[:h]:
            m1 += v
        m2 = 0.0
        for v in x[h:]:
            m2 += v
        m1 /= h
        m2 /= n - h
        drift = abs(m1 - m2)
        cl = block_stats(self.ordered("cl"))
        cy = block_stats(self.ordered("cy"))
        rel = 1.96 * cd["sem"] / max(1e-9, abs(cd["mean"]))
        drifting = drift > 2.5 * cd["sem"]
        state = "averaging" if n < 400 else ("drifting" if drifting else ("converged" if rel < 0.03 else "noisy"))
        return dict(n=n, cd=cd, cl=cl, cy=cy, drift=drift, rel=rel, state=state)


def coefficient_stats(forces, u0: float, area: float, settle: int = SETTLE):
    """run_stats() of a whole force series folded in step by step."""
    st = ForceStats(settle)
    st.accumulate_series(forces, u0, area)
    return st.run_stats()


def pressure_coefficient(rho, u0: float):
    """C_p = ((rho - 1) / 3) / (0.5 u0^2) of a density field (js/main.js:950),
    without the display's neighbour smoothing (js/main.js:944-949)."""
    r = np.asarray(rho, dtype=np.float64)
    return ((r - 1) / 3) / (0.5 * u0 * u0)
