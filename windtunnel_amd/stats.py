"""Force-coefficient statistics and the run report of the wind-tunnel app
(js/main.js:60-213, 1228-1322).

The momentum-exchange force alternates every step (lag-1 autocorrelation
near -1), so each completed step after a settling period contributes the
average of its force and the previous step's (a sliding pair average,
js/main.js:114-129), divided by the dynamic pressure 0.5 u0^2 A; the
history keeps the last HIST = 6000 such samples in float32
(js/main.js:65-68). The mean and its standard error come from 16 block
means (js/main.js:153-179), and the run is classified as averaging,
drifting, noisy or converged (js/main.js:186-207). An explicit capture
window records up to CAP_MAX samples, and run_csv writes the app's CSV run
report (its rows, order and number formats: JavaScript's Number to string,
toFixed and toExponential) from the capture or the ring. The Node drop-in
has the same layer in js/stats.js.
"""
from __future__ import annotations

import json
import math
from datetime import datetime, timezone
from decimal import ROUND_HALF_UP, Decimal

import numpy as np

HIST = 6000      # samples kept (js/main.js:65)
BLOCKS = 16      # blocking-method block count (js/main.js:163)
SETTLE = 150     # steps skipped after a reset of the statistics (js/main.js:83)
CAP_MAX = 250000  # samples a capture window keeps (js/main.js:74)
RHO_AIR = 1.225   # kg/m^3 (js/main.js:12)
NU_AIR = 1.48e-5  # m^2/s (js/main.js:13)


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

    def reset(self, settle: int = SETTLE, step: int = 0) -> None:
        """Forget everything and skip the next `settle` steps; a running
        capture ends at solver step `step` (js/main.js:76-84)."""
        cap = getattr(self, "capture", None)
        if cap is None:
            self.capture = dict(on=False, cd=[], cl=[], cy=[], start=0, end=0)
        elif cap["on"]:
            cap["on"] = False
            cap["end"] = int(step)
        self.settle = int(settle)
        self.count = 0
        self.n = 0                     # samples taken (may exceed HIST)
        self.ema = np.zeros(3)
        self.prev = None
        self.cd = np.zeros(HIST, dtype=np.float32)
        self.cl = np.zeros(HIST, dtype=np.float32)
        self.cy = np.zeros(HIST, dtype=np.float32)

    def set_capture(self, on: bool, step: int) -> None:
        """Start a new capture window, or stop the running one, at solver step
        `step` (js/main.js:131-135)."""
        if on:
            self.capture = dict(on=True, cd=[], cl=[], cy=[], start=int(step), end=int(step))
        else:
            self.capture["on"] = False
            self.capture["end"] = int(step)

    def accumulate(self, force, u0: float, area: float, step: int = 0) -> None:
        """Fold one completed step's force in (js/main.js:105-129); `step` is
        the solver's step count after it (the capture's end)."""
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
        q = 0.5 * u0 * u0 * area
        c = [float(0.5 * (float(f[k]) + float(self.prev[k])) / q) for k in range(3)]
        i = self.n % HIST
        self.cd[i], self.cl[i], self.cy[i] = c      # float32 ring
        self.prev = f.copy()
        self.n += 1
        cap = self.capture
        if cap["on"] and len(cap["cd"]) < CAP_MAX:  # the capture keeps doubles
            cap["cd"].append(c[0]); cap["cl"].append(c[1]); cap["cy"].append(c[2])
            cap["end"] = int(step)

    def accumulate_series(self, forces, u0: float, area: float, step0: int = 0) -> None:
        """Fold in consecutive step forces; the first completes step step0 + 1."""
        for k, f in enumerate(np.asarray(forces, dtype=np.float64).reshape(-1, 3)):
            self.accumulate(f, u0, area, step0 + k + 1)

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


    def run_csv(self, sim, cond: dict, generated: str | None = None) -> str:
        """The app's CSV run report (js/main.js:1228-1322): conditions, results
        with 95 % intervals and the raw series, from the capture when it
        holds samples, else from the ring. `sim` is the solver (attributes nx,
        ny, nz, n, solidCells, frontalArea, planformArea, u0, tau, nu, les,
        csm, steps); `cond` holds the page's conditions: name, custom, credit
        (dict of title/author/license/source, or None), ref ('frontal' or
        'planform'), scale (cells per body length), ground, lengthM, windMs,
        aoa, yaw. `generated` is the ISO time stamp (now when None)."""
        cap = self.capture
        captured = len(cap["cd"]) > 0
        if captured:
            ser = dict(cd=np.array(cap["cd"]), cl=np.array(cap["cl"]), cy=np.array(cap["cy"]))
        else:
            ser = dict(cd=self.ordered("cd"), cl=self.ordered("cl"), cy=self.ordered("cy"))
        n = ser["cd"].size
        st = None
        if n >= 60:
            rs = None if captured else self.run_stats()
            st = dict(n=n, cd=block_stats(ser["cd"]), cl=block_stats(ser["cl"]), cy=block_stats(ser["cy"]),
                      state="captured window" if captured else ((rs or {}).get("state") or "live window"))
        planform = cond.get("ref") == "planform"
        cells = sim.planformArea if planform else sim.frontalArea
        m_per_cell = cond["lengthM"] / cond["scale"]
        area = cells * m_per_cell * m_per_cell
        U = cond["windMs"]
        q = 0.5 * RHO_AIR * U * U
        lines = ["Wind Tunnel run report,,", "parameter,value,unit"]

        def row(k, v, unit=""):
            lines.append(",".join([k, v if isinstance(v, str) else js_str(v), unit]))
        row("generated", generated or js_iso_now())
        row("body", json.dumps(cond["name"], ensure_ascii=False))
        row("body_source", "uploaded" if cond.get("custom") else "bundled")
        if cond.get("credit"):
            for k in ("title", "author", "license", "source"):
                row("model_" + k, json.dumps(cond["credit"].get(k) or "", ensure_ascii=False))
        row("grid_nx", sim.nx, "cells"); row("grid_ny", sim.ny, "cells"); row("grid_nz", sim.nz, "cells")
        row("grid_total", sim.n, "cells")
        row("solid_cells", sim.solidCells, "cells")
        row("reference_area_type", "planform" if planform else "frontal")
        row("reference_area", cells, "cells")
        row("reference_area_real", js_fixed(area, 4), "m^2")
        row("blockage", js_fixed(100 * sim.frontalArea / (sim.ny * sim.nz), 2), "%")
        row("body_length", cond["lengthM"], "m")
        row("cells_per_body_length", js_fixed(cond["scale"], 2), "cells")
        row("wind_speed", U, "m/s")
        row("wind_speed_kmh", js_fixed(U * 3.6, 1), "km/h")
        row("mach", js_fixed(U / 343, 4))
        row("dynamic_pressure", js_fixed(q, 2), "Pa")
        row("angle_of_attack", cond["aoa"], "deg")
        row("yaw", cond["yaw"], "deg")
        row("ground_plane", "yes" if cond.get("ground") else "no")
        row("lattice_u0", sim.u0)
        row("lattice_tau", js_fixed(sim.tau, 6))
        row("lattice_nu", js_exponential(sim.nu, 4))
        row("les_smagorinsky", "on" if sim.les else "off")
        row("smagorinsky_C", sim.csm)
        row("reynolds_simulated", js_round(sim.u0 * cond["scale"] / sim.nu))
        row("reynolds_full_scale", js_round(U * cond["lengthM"] / NU_AIR))
        row("timesteps_total", sim.steps)
        row("sample_source", "explicit capture" if captured else "rolling live window")
        if captured:
            row("capture_from_step", cap["start"])
            row("capture_to_step", cap["end"])
            if len(cap["cd"]) >= CAP_MAX:
                row("capture_truncated_at", CAP_MAX, "samples")
        row("samples", st["n"] if st else 0)
        row("convergence", st["state"] if st else "not started")
        lines += [",,", "result,mean,ci95,unit"]
        if st:
            def res(k, m, sem, unit):
                lines.append(",".join([k, js_fixed(m, 5), js_fixed(1.96 * sem, 5), unit]))
            res("Cd", st["cd"]["mean"], st["cd"]["sem"], "-")
            res("Cl", st["cl"]["mean"], st["cl"]["sem"], "-")
            res("Cy", st["cy"]["mean"], st["cy"]["sem"], "-")
            res("drag", st["cd"]["mean"] * q * area, st["cd"]["sem"] * q * area, "N")
            res("lift", st["cl"]["mean"] * q * area, st["cl"]["sem"] * q * area, "N")
            res("side", st["cy"]["mean"] * q * area, st["cy"]["sem"] * q * area, "N")
            res("power", st["cd"]["mean"] * q * area * U, st["cd"]["sem"] * q * area * U, "W")
        lines += [",,",
                  "# ci95 is from the blocking method (16 blocks) which accounts for,,",
                  "# correlation between consecutive timesteps; sigma/sqrt(N) would,,",
                  "# understate it badly in a shedding wake.,,",
                  ",,", "sample,Cd,Cl,Cy"]
        for i in range(n):
            lines.append(f"{i},{js_fixed(ser['cd'][i], 5)},{js_fixed(ser['cl'][i], 5)},{js_fixed(ser['cy'][i], 5)}")
        return "\n".join(lines)


# ---------------------------------------------------- JavaScript number text
def _digits(x: float):
    """Shortest round-trip decimal digits of |x| and the exponent n with
    |x| = 0.d1d2... x 10^n (the digits both Python's repr and ECMAScript's
    Number to string produce)."""
    t = Decimal(repr(abs(x))).normalize().as_tuple()
    d = "".join(map(str, t.digits))
    return d, len(d) + t.exponent


def js_str(v) -> str:
    """ECMAScript Number::toString of a number (ints as they are)."""
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    x = float(v)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0:
        return "0"
    sign = "-" if x < 0 else ""
    d, n = _digits(x)
    k = len(d)
    if k <= n <= 21:
        body = d + "0" * (n - k)
    elif 0 < n <= 21:
        body = d[:n] + "." + d[n:]
    elif -6 < n <= 0:
        body = "0." + "0" * (-n) + d
    else:
        e = n - 1
        body = (d if k == 1 else d[0] + "." + d[1:]) + "e" + ("+" if e >= 0 else "-") + str(abs(e))
    return sign + body


def js_fixed(v, f: int) -> str:
    """ECMAScript Number.prototype.toFixed(f): the exact binary value rounded
    to f decimals, a tie going to the larger magnitude."""
    x = float(v)
    if math.isnan(x):
        return "NaN"
    if abs(x) >= 1e21:
        return js_str(x)
    q = Decimal(abs(x)).quantize(Decimal(1).scaleb(-f), rounding=ROUND_HALF_UP)
    return ("-" if x < 0 else "") + f"{q:.{f}f}"


def js_exponential(v, f: int) -> str:
    """ECMAScript Number.prototype.toExponential(f)."""
    x = float(v)
    if x == 0:
        return "0" + ("." + "0" * f if f else "") + "e+0"
    a = Decimal(abs(x))
    e = a.adjusted()
    m = (a.scaleb(-e)).quantize(Decimal(1).scaleb(-f), rounding=ROUND_HALF_UP)
    if m >= 10:
        e += 1
        m = (a.scaleb(-e)).quantize(Decimal(1).scaleb(-f), rounding=ROUND_HALF_UP)
    return ("-" if x < 0 else "") + f"{m:.{f}f}" + "e" + ("+" if e >= 0 else "-") + str(abs(e))


def js_round(v) -> int:
    """ECMAScript Math.round: the nearest integer, a tie toward +infinity."""
    return int(math.floor(Decimal(float(v)) + Decimal("0.5")))


def js_iso_now() -> str:
    """new Date().toISOString()."""
    t = datetime.now(timezone.utc)
    return t.strftime("%Y-%m-%dT%H:%M:%S.") + f"{t.microsecond // 1000:03d}Z"


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
