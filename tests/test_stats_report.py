"""The app's capture window and CSV run report (js/main.js:70-136,
1228-1322) in stats.py and in the Node drop-in's js/stats.js.

Pinned by closed forms on known force series (a constant force with a
step-parity oscillation gives every pair-averaged sample the same value, so
mean, interval and rows are known exactly) and by the two implementations
writing the same text for the same seeded series, with and without a
capture window."""
import json
import os
import shutil
import subprocess
from types import SimpleNamespace

import numpy as np
import pytest

from _util import ROOT
from windtunnel_amd.stats import CAP_MAX, ForceStats, js_fixed

NODE = shutil.which("node")
SUITES = [pytest.param("cpu"), pytest.param("gpu", marks=pytest.mark.gpu)]

# rows of the report's parameter table, in the app's order (js/main.js:1257-1297)
KEYS = ["generated", "body", "body_source", "grid_nx", "grid_ny", "grid_nz", "grid_total", "solid_cells",
        "reference_area_type", "reference_area", "reference_area_real", "blockage", "body_length",
        "cells_per_body_length", "wind_speed", "wind_speed_kmh", "mach", "dynamic_pressure", "angle_of_attack",
        "yaw", "ground_plane", "lattice_u0", "lattice_tau", "lattice_nu", "les_smagorinsky", "smagorinsky_C",
        "reynolds_simulated", "reynolds_full_scale", "timesteps_total", "sample_source"]

SIM = dict(nx=96, ny=44, nz=44, n=96 * 44 * 44, solidCells=1234, frontalArea=144, planformArea=210, u0=0.06,
           tau=0.503, nu=(0.503 - 0.5) / 3, les=True, csm=0.14, steps=900)
COND = dict(name="sedan \"S\"", custom=False, credit=None, ref="frontal", scale=40.0, ground=True, lengthM=4.6,
            windMs=30, aoa=0, yaw=5)


def parity_series(n, mean=(0.3, -0.1, 0.02), amp=0.05):
    k = np.arange(n)
    return np.stack([m + amp * (-1.0) ** k for m in mean], 1)


def report(forces, capture=None, settle=150, sim=SIM, cond=COND):
    st = ForceStats(settle)
    for k, f in enumerate(forces):
        if capture and k == capture[0]:
            st.set_capture(True, k)
        st.accumulate(f, sim["u0"], sim["frontalArea"], k + 1)
        if capture and k + 1 == capture[1]:
            st.set_capture(False, k + 1)
    return st.run_csv(SimpleNamespace(**sim), cond, "2026-01-02T03:04:05.678Z")


def table(csv):
    lines = csv.split("\n")
    head = lines[2:lines.index(",,")]
    return {l.split(",")[0]: l.split(",")[1:] for l in head}, [l.split(",")[0] for l in head], lines


@pytest.mark.parametrize("suite", SUITES)
def test_constant_force_report_is_exact(suite):
    F = parity_series(150 + 1 + 700)
    csv = report(F)
    t, keys, lines = table(csv)
    assert keys == KEYS + ["samples", "convergence"]
    q = 0.5 * 0.06 ** 2 * 144
    assert t["samples"] == ["700", ""] and t["sample_source"][0] == "rolling live window"
    cd = lines[lines.index("result,mean,ci95,unit") + 1].split(",")
    assert cd == ["Cd", js_fixed(np.float32(0.3 / q), 5), "0.00000", "-"]
    assert t["convergence"][0] == "converged"
    assert t["body"][0] == '"sedan \\"S\\""'
    rows = lines[lines.index("sample,Cd,Cl,Cy") + 1:]
    assert len(rows) == 700 and rows[0].startswith("0,") and rows[-1].startswith("699,")
    assert t["reynolds_full_scale"][0] == str(round(30 * 4.6 / 1.48e-5))


@pytest.mark.parametrize("suite", SUITES)
def test_capture_window_report(suite):
    F = parity_series(1500)
    csv = report(F, capture=(400, 1100))
    t, keys, lines = table(csv)
    assert keys == KEYS + ["capture_from_step", "capture_to_step", "samples", "convergence"]
    assert t["capture_from_step"][0] == "400" and t["capture_to_step"][0] == "1100"
    assert t["samples"][0] == "700" and t["convergence"][0] == "captured window"
    assert t["sample_source"][0] == "explicit capture"
    assert len(lines) - lines.index("sample,Cd,Cl,Cy") - 1 == 700


def test_capture_stops_at_its_maximum():
    st = ForceStats(0)
    st.set_capture(True, 0)
    f = np.array([0.3, 0.0, 0.0])
    for k in range(CAP_MAX + 3):
        st.accumulate(f, 0.06, 144, k + 1)
    assert len(st.capture["cd"]) == CAP_MAX and st.capture["end"] == CAP_MAX + 1
    t, keys, _ = table(st.run_csv(SimpleNamespace(**SIM), COND, "x"))
    assert t["capture_truncated_at"] == [str(CAP_MAX), "samples"]


@pytest.mark.skipif(NODE is None, reason="node not installed")
@pytest.mark.parametrize("suite", SUITES)
@pytest.mark.parametrize("capture,planform", [(None, False), ((300, 1700), True)])
def test_node_drop_in_writes_the_same_report(tmp_path, suite, capture, planform):
    rng = np.random.default_rng(9)
    F = parity_series(2400) + rng.normal(scale=0.01, size=(2400, 3))
    cond = dict(COND, ref="planform" if planform else "frontal",
                credit=dict(title="Car", author="A. B.", license="CC-BY", source="x") if planform else None)
    area = SIM["planformArea"] if planform else SIM["frontalArea"]
    inp = dict(forces=F.tolist(), u0=SIM["u0"], area=area, settle=150, capture=list(capture) if capture else None,
               sim=SIM, cond=cond, generated="2026-01-02T03:04:05.678Z")
    p = tmp_path / "in.json"
    p.write_text(json.dumps(inp))
    js = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "stats_report.js"), str(p)], capture_output=True,
                        text=True, timeout=120, check=True).stdout
    st = ForceStats(150)
    for k, f in enumerate(F):
        if capture and k == capture[0]:
            st.set_capture(True, k)
        st.accumulate(f, SIM["u0"], area, k + 1)
        if capture and k + 1 == capture[1]:
            st.set_capture(False, k + 1)
    py = st.run_csv(SimpleNamespace(**SIM), cond, "2026-01-02T03:04:05.678Z")
    assert js == py
