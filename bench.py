"""Throughput of the D3Q19 step on one MI355X: MLUPS and fraction of the HBM
roofline, fp32 against binary16-offset storage, plus the adjoint/forward cost.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2]
                    [--no-cpu] [--no-f16] [--no-precision] [--no-adjoint]
                    [--no-exact] [--no-micro] [--no-c1] [--no-c2] [--no-drag]
                    [--no-scale]
                    [--fd-cells 64]

Throughput workload (default): D3Q19 TRT (Lambda = 3/16) with Smagorinsky LES
(Cs = 0.1), tau = 0.505, inflow u0 = 0.08, on a 512x256x256 lattice (33.5 M
cells) around the synthetic bluff-body mesh scaled 2x (ellipsoid 40x14x12
unioned with a 30x16x10 box, ~20k jittered triangles), voxelized on the GPU
(Voxel.build at res 200, then the grid-backed body lookup). One step is one
fused stream/collide/bounce-back/force launch over the whole lattice plus the
per-step force reduction; the populations are resident in HBM before timing
starts. The headline value is the float32 storage path; the binary16-offset
storage path runs on the same lattice and is reported beside it, and so is
float32 storage with binary64 arithmetic (the reference's own numbers).

Roofline: HBM bound. Algorithmic bytes per cell update = 19 populations read
+ 19 written = 152 B (float32) or 76 B (binary16); MLUPS counts all nx*ny*nz
cells as the reference's benchmark does (js/main.js:621). `achieved` divides
the algorithmic bytes of one launch by the average step-kernel duration,
measured with HIP events on the lattice's stream right after the timed
steps: the device time of N + 1 back-to-back launches less that of one
launch, over N = max(K, 100) (each region ends with the launch that writes
the fields; `kernel_launches` is N).
`kernel_ms_per_launch_events` is the same average with an event pair around
every launch, which adds the records' own waits. `value` is the wall clock
of the timed steps, which carry no events.

Precision leg: the same body at half scale on 128x64x64 in binary16, float32
and binary64 storage (the binary64 GPU path equals the oracle's binary64 run
bit for bit), 4000 steps, laminar (tau 0.6) and in the TRT+LES form; and
binary16 against float32 on 512x256x256 over 4000 steps. C_d is the
pair-averaged mean over the last half (16-block SEM); fields are compared
at the last step.

Adjoint leg: 192x96x96 TRT (tau 0.6, u0 0.04) around a gray-walled ellipsoid
(seeded smooth weights, 1 inside, 0 outside), 3000 steps with sqrt(N)
checkpointing, objective J = mean C_d over the last 500 steps; cost ratio =
(recompute + adjoint) device time per step / forward device time per step;
central differences (eps 1e-3, binary64) on seeded gray cells, and their
Richardson extrapolation from eps and eps/2 (which removes the O(eps^2)
truncation error of the prescribed eps).

C2 leg: the 256x128x128 TRT+LES wake of the same body (tau 0.505, u0 0.08)
placed as the app places it, 5000 steps in binary64 arithmetic (equal to the
reference's recorded run step for step), float32 and binary16 storage, with
the app's time-averaged C_d/C_l statistics; at the configured Cs = 0.1 (the
flow's stability edge) and at the solver's default Cs = 0.14.
"""
from __future__ import annotations

import argparse
import base64
import gzip
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "D3Q19 MLUPS & % HBM roofline (fp32 vs fp16-store); adjoint/forward cost"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_kernel_hbm_traffic_512x256x256.md")

WORKLOADS = {
    # TRT+LES around the synthetic body; 512x256x256 is where the throughput
    # targets are quoted, 256x128x128 is its half-scale sibling.
    "c3": dict(dims=(512, 256, 256), scale=2, tau=0.505, u0=0.08, lam=3 / 16, csm=0.1),
    "c2": dict(dims=(256, 128, 128), scale=1, tau=0.505, u0=0.08, lam=3 / 16, csm=0.1),
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------------ bodies
_MESH = {}


def body_mesh():
    """The ~20k-triangle synthetic bluff body, in cells around its centre."""
    if "P" not in _MESH:
        from windtunnel_amd.synth import ellipsoid_box
        _MESH["P"] = ellipsoid_box()
    return _MESH["P"]


def place_mesh_body(L, scale: float, res: int = 200) -> dict:
    """Voxelize the body mesh on the GPU (Voxel.build) and put it into the
    lattice at (0.3 nx, ny/2, nz/2), `scale` cells per mesh unit, through the
    grid-backed body lookup over its box grown by 2 cells."""
    from windtunnel_amd.lbm import Voxel
    t0 = time.perf_counter()
    vox = Voxel.build(body_mesh(), res)
    pos = np.array([0.3 * L.nx, L.ny / 2, L.nz / 2])
    b = np.array(vox["bbox"])
    lo, hi = pos + scale * b[:3] - 2, pos + scale * b[3:] + 2
    box = dict(x0=lo[0], y0=lo[1], z0=lo[2], x1=hi[0], y1=hi[1], z1=hi[2])
    L.voxelize_grid(vox, np.eye(3).reshape(-1), pos, scale, box)
    return dict(voxelize_s=time.perf_counter() - t0, grid=[vox["nx"], vox["ny"], vox["nz"]],
                tris=vox["stats"]["tris"], openness=vox["stats"]["openness"])


def set_flow(L, wl) -> None:
    L.set_params(tau=wl["tau"], les=True, csm=wl["csm"], trt=True, u0=wl["u0"], **{"lambda": wl["lam"]})


# -------------------------------------------------------------- throughput
def run_gpu(wl, storage, steps, warmup, arith="f32"):
    from windtunnel_amd.lbm import LBM
    L = LBM(*wl["dims"], storage=storage, arith=arith)
    set_flow(L, wl)
    L.reset()
    vinfo = place_mesh_body(L, wl["scale"])
    L.step(warmup)
    # wall clock over the K steps as a caller runs them (no per-step events)
    t0 = time.perf_counter()
    forces = L.step(steps)          # returns after the stream has drained
    wall = time.perf_counter() - t0
    # then the step kernel's device time back to back: HIP events on the
    # lattice's stream around N + 1 launches and around one launch (median of
    # three); each region ends with the field-writing launch, so the
    # difference is N launches of the timed steps' kernel. N = max(K, 100):
    # the binary64-arithmetic and binary16 steps run several % slower for the
    # first ~20 ms of a lattice's run (the trace shows it in every leg),
    # and 100 launches keep that transient from setting the figure when K
    # is small
    nk = max(steps, 100)
    t_long = L.time_steps(nk + 1)
    t_one = float(np.median([L.time_steps(1) for _ in range(3)]))
    kavg = (t_long - t_one) / nk
    # and with an event pair around every launch (each pair adds its own
    # records and waits to the launch it brackets)
    L.set_timing(True)
    forces2 = L.step(steps)
    kms, launches = L.timing()
    L.set_timing(False)
    n = L.n
    bpc = 76 if storage == "f16" else 152
    res = dict(
        mlups=n * steps / wall / 1e6, ms_per_step=wall * 1e3 / steps, kernel_ms=kavg, kernel_launches=nk,
        kernel_ms_per_launch_events=kms / launches,
        achieved_gbs=n * bpc / (kavg * 1e-3) / 1e9, bytes_per_cell=bpc, block_cells=L.block_cells,
        solid_cells=L.solidCells, frontal_area=L.frontalArea, body=vinfo,
        finite=bool(np.all(np.isfinite(forces)) and np.all(np.isfinite(forces2))),
    )
    L.close()
    return res


def measured_traffic():
    """HBM bytes per step launch from the committed rocprofv3 PMC summary
    (lines `traffic[f32] = <bytes>`), or None per storage type."""
    out = {}
    try:
        with open(TRAFFIC_FILE) as fh:
            for m in re.finditer(r"traffic\[(\w+)\]\s*=\s*([0-9.eE+]+)", fh.read()):
                out[m.group(1)] = float(m.group(2))
    except OSError:
        pass
    return out


# --------------------------------------------------------------- precision
def _storage_runs(dims, scale, tau, u0, les, steps, avg, storages, csm=0.1):
    """Run the mesh-body lattice in each (storage, arith); keep C_d statistics
    over the last `avg` steps (pair-averaged, 16-block SEM) and the fields."""
    from windtunnel_amd.lbm import LBM
    from windtunnel_amd.stats import block_stats, pair_coefficients
    out = {}
    for storage, arith in storages:
        L = LBM(*dims, storage=storage, arith=arith)
        L.set_params(tau=tau, les=les, csm=csm, trt=True, u0=u0, **{"lambda": 3 / 16})
        L.reset()
        place_mesh_body(L, scale)
        F = L.step(steps)
        bs = block_stats(pair_coefficients(F[-avg:], u0, L.frontalArea)[0])
        out[storage] = dict(cd=bs["mean"], cd_sem=bs["sem"], finite=bool(np.all(np.isfinite(F))),
                            fields={k: np.asarray(v, dtype=np.float64).copy() for k, v in L.fields().items()})
        L.close()
    return out


def _compare(runs, ref_key):
    ref = runs[ref_key]

    def rl2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out = {"cd_" + ref_key: ref["cd"], "cd_sem_" + ref_key: ref["cd_sem"]}
    for k, r in runs.items():
        if k == ref_key:
            continue
        f, g = r["fields"], ref["fields"]
        out[k] = dict(cd=r["cd"], cd_sem=r["cd_sem"], cd_rel_err=abs(r["cd"] - ref["cd"]) / abs(ref["cd"]),
                      ux_rel_l2=rl2(f["ux"], g["ux"]), rho_dev_rel_l2=rl2(f["rho"] - 1, g["rho"] - 1),
                      finite=r["finite"])
    return out


def run_precision(c3_steps=4000):
    """Binary16-offset storage against float32 and binary64 storage.

    * laminar: 128x64x64, TRT tau 0.6, u0 0.05 (steady wake), 4000 steps;
      deterministic, so fields and C_d compare directly against binary64.
    * les_steady: TRT + Smagorinsky LES with the configured Cs = 0.1 on
      128x64x64 at tau 0.55, u0 0.05: the eddy viscosity is active but the
      wake stays steady, so the time-averaged C_d compares directly.
    * les_128: the TRT+LES form (tau 0.505, u0 0.08) on 128x64x64, 4000
      steps: a chaotic wake, so instantaneous fields decorrelate and a C_d
      mean over 2000 steps carries its SEM (which understates the error at
      ~2 shedding periods per window).
    * c3: 512x256x256 TRT+LES, binary16 against float32 on the same GPU.
    The LES legs use Cs = 0.14 (the solver's default, js/lbm.js:109): with
    Cs = 0.1 the impulsively started 512x256x256 and 256x128x128 wakes
    diverge within 1000 steps, in binary64 arithmetic (the reference's own)
    as in float.
    """
    f64 = [("f64", "f64"), ("f32", "f32"), ("f16", "f32")]
    res = dict(laminar=dict(grid="128x64x64", tau=0.6, u0=0.05, les=False, steps=4000, cd_window=2000),
               les_steady=dict(grid="128x64x64", tau=0.55, u0=0.05, les=True, csm=0.1, steps=4000, cd_window=2000),
               les_128=dict(grid="128x64x64", tau=0.505, u0=0.08, les=True, csm=0.14, steps=4000, cd_window=2000,
                            chaotic=True),
               c3=dict(grid="512x256x256", tau=0.505, u0=0.08, les=True, csm=0.14, steps=c3_steps,
                       cd_window=c3_steps // 2, chaotic=True,
                       note="Cs=0.1 diverges within 1000 steps here, in binary64 arithmetic as in float"))
    res["laminar"].update(_compare(_storage_runs((128, 64, 64), 0.5, 0.6, 0.05, False, 4000, 2000, f64), "f64"))
    res["les_steady"].update(_compare(_storage_runs((128, 64, 64), 0.5, 0.55, 0.05, True, 4000, 2000, f64, 0.1),
                                      "f64"))
    res["les_128"].update(_compare(_storage_runs((128, 64, 64), 0.5, 0.505, 0.08, True, 4000, 2000, f64, 0.14), "f64"))
    if c3_steps > 0:
        res["c3"].update(_compare(_storage_runs((512, 256, 256), 2, 0.505, 0.08, True, c3_steps, c3_steps // 2,
                                                [("f32", "f32"), ("f16", "f32")], 0.14), "f32"))
    return res


# ----------------------------------------------------------------- adjoint
def gray_ellipsoid(dims, seed=4):
    """Gray weights of the differentiation workload: an ellipsoid of semi-axes
    (24, 9, 8) at (0.3 nx, ny/2, nz/2) whose wall is a ~3-cell band where the
    weight falls from 1 to 0, displaced by a seeded smooth field."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    c = (0.3 * nx, ny / 2, nz / 2)
    s = np.sqrt(((x - c[0]) / 24) ** 2 + ((y - c[1]) / 9) ** 2 + ((z - c[2]) / 8) ** 2)
    phi = np.zeros_like(s)
    for _ in range(4):
        k = rng.normal(scale=0.25, size=3)
        phi += np.sin(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 2 * np.pi)) / 4
    d = (s - 1) * 8 + 0.8 * phi
    return np.clip(0.5 - d / 3, 0, 1).reshape(-1).astype(np.float32)


def run_adjoint(fd_cells=64, N=3000, K=500, eps=1e-3):
    from windtunnel_amd.lbm import LBM
    dims = (192, 96, 96)
    w = gray_ellipsoid(dims)
    solid = (w >= 1).reshape(dims[::-1])
    idx = np.argwhere(solid)
    lo, hi = idx.min(0), idx.max(0)
    box = [int(lo[2]), int(lo[1]), int(lo[0]), int(hi[2]), int(hi[1]), int(hi[0])]
    mask = solid[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1].astype(np.uint8).reshape(-1)
    wg = np.where(w >= 1, 0, w).astype(np.float32)

    def lattice(storage, arith):
        L = LBM(*dims, storage=storage, arith=arith)
        L.set_params(tau=0.6, les=False, trt=True, u0=0.04, **{"lambda": 3 / 16})
        L.reset()
        L.voxelize_mask(box, mask)
        return L

    def prepare(L, weights):
        L.reset()
        L.set_gray(weights)

    res = dict(grid="192x96x96", steps=N, objective=f"mean C_d over the last {K} steps", gray_cells=int(
        np.count_nonzero((wg > 0) & (wg < 1))))
    # float32: forward device time per step, then recorded forward + reverse
    L = lattice("f32", "f32")
    q = 0.5 * 0.04 ** 2 * L.frontalArea
    Fbar = np.zeros((N, 3))
    Fbar[N - K:, 0] = 1.0 / (K * q)
    prepare(L, wg)
    fwd_ms = L.time_steps(200) / 200
    prepare(L, wg)
    t0 = time.perf_counter()
    F32 = L.adjoint_forward(N)
    fwd_wall = time.perf_counter() - t0
    g32, rev_ms = L.adjoint_backward(Fbar)
    L.adjoint_release()
    L.close()
    res.update(forward_ms_per_step=fwd_ms, reverse_ms_per_step=rev_ms / N,
               cost_ratio=(rev_ms / N) / fwd_ms, recorded_forward_ms_per_step=fwd_wall * 1e3 / N,
               J_fp32=float(np.sum(Fbar * F32)))
    log(f"adjoint fp32: forward {fwd_ms:.4f} ms/step, reverse {rev_ms / N:.4f} ms/step, "
        f"ratio {res['cost_ratio']:.2f}")
    # binary64: adjoint and central differences
    D = lattice("f64", "f64")
    prepare(D, wg)
    F64 = D.adjoint_forward(N)
    g64, _ = D.adjoint_backward(Fbar)
    D.adjoint_release()
    gray = np.flatnonzero((wg > 0) & (wg < 1))
    res["J_fp64"] = float(np.sum(Fbar * F64))
    res["grad_fp32_vs_fp64_rel_l2"] = float(np.linalg.norm(g32[gray] - g64[gray]) / np.linalg.norm(g64[gray]))
    if fd_cells > 0:
        ok = np.flatnonzero((wg > 2 * eps) & (wg < 1 - 2 * eps))
        cells = np.random.default_rng(4).choice(ok, min(fd_cells, ok.size), replace=False)

        def central(c, e):
            # central difference over the float32 weights actually stored;
            # returns (difference quotient, half-width)
            J, wv = [], []
            for sgn in (1, -1):
                wp = wg.copy()
                wp[c] = np.float32(wg[c] + sgn * e)
                wv.append(float(wp[c]))
                prepare(D, wp)
                J.append(float(np.sum(D.step(N) * Fbar)))
            return (J[0] - J[1]) / (wv[0] - wv[1]), 0.5 * (wv[0] - wv[1])

        fds, fdr = [], []
        t0 = time.perf_counter()
        for c in cells:
            d1, h1 = central(c, eps)
            d2, h2 = central(c, eps / 2)
            fds.append(d1)
            # Richardson: the central difference errs by a h^2 + O(h^4)
            fdr.append((h1 * h1 * d2 - h2 * h2 * d1) / (h1 * h1 - h2 * h2))
        fds, fdr = np.array(fds), np.array(fdr)
        g = g64[cells]

        def rel(ref):
            return np.abs(g - ref) / np.maximum(np.abs(ref), 1e-3 * np.max(np.abs(ref)))

        r1, r2 = rel(fds), rel(fdr)
        res["fd"] = dict(cells=int(cells.size), eps=eps, max_rel_err=float(r1.max()),
                         median_rel_err=float(np.median(r1)), within_1e3=int(np.sum(r1 <= 1e-3)),
                         richardson=dict(eps=[eps, eps / 2], max_rel_err=float(r2.max()),
                                         median_rel_err=float(np.median(r2)), within_1e3=int(np.sum(r2 <= 1e-3))),
                         rel_floor="errors relative to max(|fd|, 1e-3 max|fd|)",
                         seconds=round(time.perf_counter() - t0, 1))
        log(f"adjoint FD: {res['fd']}")
    D.close()
    return res


# --------------------------------------------------------------------- CPU
def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu(wl, threads):
    """The oracle (C restatement of js/lbm.js with the worker-pool slab split
    of js/lbmWorker.js) on the host cores with the reference benchmark's
    protocol (js/main.js:614-620): 12 warm-up steps, then the best of 3 x 20
    steps on the full lattice; then one thread on the same lattice, as a
    bounded sample (1 warm-up step, best of 2 x 1 step: 20-step timings at
    one thread would take minutes). The body is the same mesh body: its
    solid map is taken from the GPU lattice. Returns (MLUPS at `threads`,
    MLUPS at 1 thread)."""
    from oracle.oracle import OracleLBM
    dims = wl["dims"]
    box, mask = _cpu_mask(dims, wl["scale"])
    O = OracleLBM(*dims)
    O.set_params(tau=wl["tau"], les=True, csm=wl["csm"], trt=True, u0=wl["u0"], **{"lambda": wl["lam"]})
    O.reset()
    O.voxelize_mask(box, mask)
    O.step_pool(12, threads, threaded=True)
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        O.step_pool(20, threads, threaded=True)
        best = min(best, (time.perf_counter() - t0) / 20)
    O.step_pool(1, 1, threaded=False)
    best1 = float("inf")
    for _ in range(2):
        t0 = time.perf_counter()
        O.step_pool(1, 1, threaded=False)
        best1 = min(best1, time.perf_counter() - t0)
    n = O.n
    del O
    return n / best / 1e6, n / best1 / 1e6


def host_cpus() -> dict:
    """Host description: model, logical CPUs, physical cores, and the CPUs this
    process may run on (nproc)."""
    cores = set()
    try:
        with open("/proc/cpuinfo") as fh:
            phys = core = None
            for line in fh:
                if line.startswith("physical id"):
                    phys = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    core = line.split(":")[1].strip()
                    cores.add((phys, core))
    except OSError:
        pass
    return dict(model=cpu_model(), logical=os.cpu_count(), physical_cores=len(cores) or None,
                nproc=len(os.sched_getaffinity(0)))


def run_cpu_detail(threads):
    """The oracle at 1 thread and at `threads` threads on the CPU-runnable
    96x48x48 BGK sphere (12 warm-up steps, best of 3 x 20 steps: the
    reference benchmark's protocol, js/main.js:614-620) and on the
    256x128x128 TRT+LES mesh body (2 warm-up steps, best of 3 x 4 steps)."""
    from oracle.oracle import OracleLBM

    def timed(O, warm, reps, steps, t):
        O.step_pool(warm, t, threaded=True)
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            O.step_pool(steps, t, threaded=True)
            best = min(best, (time.perf_counter() - t0) / steps)
        return round(O.n / best / 1e6, 2)

    out = {"host": host_cpus()}
    lo, hi = [23, 15, 15], [41, 33, 33]
    z, y, x = np.meshgrid(np.arange(lo[2], hi[2] + 1) + 0.5, np.arange(lo[1], hi[1] + 1) + 0.5,
                          np.arange(lo[0], hi[0] + 1) + 0.5, indexing="ij")
    mask = ((x - 32) ** 2 + (y - 24) ** 2 + (z - 24) ** 2 <= 64).astype(np.uint8).reshape(-1)
    c1 = {}
    for t in sorted({1, threads}):
        O = OracleLBM(96, 48, 48)
        O.set_params(tau=0.56, les=False, trt=False, u0=0.05, csm=0.14, **{"lambda": 0.25})
        O.reset()
        O.voxelize_mask(lo + hi, mask)
        c1[f"{t}_threads"] = timed(O, 12, 3, 20, t)
        del O
    out["c1_96x48x48_bgk_mlups"] = c1
    wl = WORKLOADS["c2"]
    box, m2 = _cpu_mask(wl["dims"], wl["scale"])
    c2 = {}
    for t in sorted({1, threads}):
        O = OracleLBM(*wl["dims"])
        O.set_params(tau=wl["tau"], les=True, csm=wl["csm"], trt=True, u0=wl["u0"], **{"lambda": wl["lam"]})
        O.reset()
        O.voxelize_mask(box, m2)
        c2[f"{t}_threads"] = timed(O, 2, 3, 4, t)
        del O
    out["c2_256x128x128_trt_les_mlups"] = c2
    out["reference_lbm_js_node_1_thread_mlups"] = {
        "c1": 7.7, "c2": 2.0,
        "note": "js/lbm.js itself under Node v12, measured in the development container (8 x Xeon @ 2.70 GHz), "
                "not on this host"}
    return out


# ------------------------------------------------------------------- micro
def run_micro(steps=100):
    """Kernel microbenchmark: 512x256x256 TRT+LES with no body, populations
    set to random deviations uniform in +-1e-3 (seed 7): the step kernel
    alone against the HBM roofline."""
    from windtunnel_amd.lbm import LBM
    dims = WORKLOADS["c3"]["dims"]
    n = dims[0] * dims[1] * dims[2]
    r = np.random.default_rng(7).random(19 * n, dtype=np.float32)
    r *= 2e-3
    r -= 1e-3
    out = {"grid": "512x256x256", "body": None, "populations": "uniform +-1e-3 deviations, numpy default_rng(7)",
           "steps": steps}
    for storage, bpc in (("f32", 152), ("f16", 76)):
        L = LBM(*dims, storage=storage)
        set_flow(L, WORKLOADS["c3"])
        L.reset()
        L.set_pops(r)
        L.time_steps(10)
        ms = L.time_steps(steps) / steps
        out[storage] = dict(kernel_ms=round(ms, 5), mlups=round(n / (ms * 1e-3) / 1e6, 1),
                            roofline_frac=round(n * bpc / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4))
        L.close()
    del r
    return out


# ---------------------------------------------------------------------- c1
def run_c1(steps=2000):
    """The CPU-runnable workload on the GPU: 96x48x48 BGK, sphere r = 8 at
    (32, 24, 24) tested at cell centres, tau = 0.56, u0 = 0.05, 2000 steps,
    with the reference's boundaries and with the configuration's own
    (free-slip side walls, equilibrium outflow); mean C_d over the last
    1000 steps."""
    from windtunnel_amd.lbm import LBM
    dims = (96, 48, 48)
    lo, hi = [23, 15, 15], [41, 33, 33]
    z, y, x = np.meshgrid(np.arange(lo[2], hi[2] + 1) + 0.5, np.arange(lo[1], hi[1] + 1) + 0.5,
                          np.arange(lo[0], hi[0] + 1) + 0.5, indexing="ij")
    mask = ((x - 32) ** 2 + (y - 24) ** 2 + (z - 24) ** 2 <= 64).astype(np.uint8).reshape(-1)
    out = {"grid": "96x48x48", "steps": steps, "tau": 0.56, "u0": 0.05, "collision": "BGK",
           "timing": "wall clock of the 2000 steps from reset(), after 20 steps on a throwaway lattice of the "
                     "same configuration (the first launch of a kernel loads its code)"}
    for mode in ("reference", "channel"):
        for warm in (True, False):
            L = LBM(*dims)
            L.set_params(tau=0.56, les=False, trt=False, u0=0.05, csm=0.14, **{"lambda": 0.25})
            L.set_boundary_mode(mode)
            L.reset()
            L.voxelize_mask(lo + hi, mask)
            if warm:
                L.step(20)
                L.close()
        t0 = time.perf_counter()
        F = L.step(steps)
        wall = time.perf_counter() - t0
        q = 0.5 * 0.05 ** 2 * L.frontalArea
        out[mode] = dict(cd=float(np.mean(F[steps // 2:, 0]) / q), mlups=round(L.n * steps / wall / 1e6, 1),
                         finite=bool(np.all(np.isfinite(F))))
        L.close()
    return out


# -------------------------------------------------------------------- drag
def run_drag(moves=12):
    """Dragging the body: the 512x256x256 lattice with the mesh body (x2,
    grid-backed lookup), moved by one cell along x per re-voxelize, as the
    app re-voxelizes on the next frame after a drag moves the body
    (js/main.js:478-485,1387-1394); wall time of each voxelize call, median
    over the moves after the first two."""
    from windtunnel_amd.lbm import LBM, Voxel
    wl = WORKLOADS["c3"]
    vox = Voxel.build(body_mesh(), 200)
    L = LBM(*wl["dims"])
    set_flow(L, wl)
    L.reset()
    b = np.array(vox["bbox"])
    ts = []
    for k in range(moves):
        pos = np.array([0.3 * L.nx + k, L.ny / 2, L.nz / 2])
        lo, hi = pos + wl["scale"] * b[:3] - 2, pos + wl["scale"] * b[3:] + 2
        t0 = time.perf_counter()
        L.voxelize_grid(vox, np.eye(3).reshape(-1), pos, wl["scale"],
                        dict(x0=lo[0], y0=lo[1], z0=lo[2], x1=hi[0], y1=hi[1], z1=hi[2]))
        ts.append(time.perf_counter() - t0)
    out = {"grid": "512x256x256", "moves": moves, "solid_cells": int(L.solidCells),
           "voxelize_ms_median": round(1e3 * float(np.median(ts[2:])), 3),
           "voxelize_ms_first": round(1e3 * ts[0], 3)}
    L.close()
    return out


# ---------------------------------------------------------------------- c2
C2_FIXTURE = os.path.join(ROOT, "tests", "golden", "fixtures", "lbm_c2.json")


def _fixture_array(rec: dict) -> np.ndarray:
    raw = gzip.decompress(base64.b64decode(rec["gz_b64"]))
    return np.frombuffer(raw, dtype={"f4": np.float32, "f8": np.float64}[rec["dtype"]]).reshape(rec["shape"])


def run_c2(steps=5000):
    """The 256x128x128 TRT (Lambda 3/16) + LES (Cs 0.1) wake at tau 0.505,
    u0 0.08, of the ~20k-triangle body placed as the app places it
    (Shapes.orient to unit length, Voxel.build at res 200, the grid-backed
    Body.makeTest): 5000 steps, the app's time-averaged C_d/C_l (settle 150,
    pair averages, 16-block SEM) and the largest final speed. The reference's
    own 5000-step run under Node (its force series, recorded with the golden
    vectors) is the comparison: binary64 arithmetic reproduces it bit for bit.
    This flow sits on its stability edge (its speed passes the lattice sound
    speed), so the same runs are repeated with the solver's default
    Cs = 0.14, where the float paths are compared with binary64 by their
    statistics (the wake is chaotic)."""
    from windtunnel_amd.lbm import LBM, Voxel
    from windtunnel_amd.shapes import orient
    from windtunnel_amd.stats import coefficient_stats
    from windtunnel_amd.synth import ellipsoid_box
    with open(C2_FIXTURE) as fh:
        fx = json.load(fh)
    ref = _fixture_array(fx["forces"])[:steps]
    p = fx["params"]
    P = ellipsoid_box()
    orient(P, 0, 1, True, False)
    t0 = time.perf_counter()
    vox = Voxel.build(P, fx["mesh"]["addModel"]["res"])
    vox_s = time.perf_counter() - t0
    b, bi = fx["voxelize"]["body"], fx["voxelize"]["box_in"]

    def stats(F, area):
        s = coefficient_stats(F, p["u0"], area)
        return None if s is None else dict(cd=s["cd"]["mean"], cd_sem=s["cd"]["sem"], cl=s["cl"]["mean"],
                                           cl_sem=s["cl"]["sem"], samples=s["n"], state=s["state"])

    out = {"grid": "256x128x128", "steps": steps, "tau": p["tau"], "u0": p["u0"], "Lambda": p["lambda"],
           "voxelize_s": round(vox_s, 3), "reference_node_ms_per_step": fx["node_ms_per_step"]}
    area = None
    for csm in (p["csm"], 0.14):
        runs = {}
        for storage, arith in (("f32", "f64"), ("f32", "f32"), ("f16", "f32")):
            L = LBM(*fx["dims"], storage=storage, arith=arith)
            L.set_params(tau=p["tau"], les=p["les"], csm=csm, trt=p["trt"], u0=p["u0"], **{"lambda": p["lambda"]})
            L.reset()
            L.voxelize_grid(vox, b["M"], b["pos"], b["scale"],
                            dict(x0=bi[0], y0=bi[1], z0=bi[2], x1=bi[3], y1=bi[4], z1=bi[5]))
            area = L.frontalArea
            t0 = time.perf_counter()
            F = L.step(steps)
            wall = time.perf_counter() - t0
            fin = np.isfinite(F).all(axis=1)
            first_bad = None if fin.all() else int(np.argmin(fin)) + 1
            r = dict(mlups=round(L.n * steps / wall / 1e6, 1), finite=first_bad is None)
            if first_bad is None:
                r["stats"] = stats(F, area)
                Fl = L.fields()
                r["max_speed"] = float(np.sqrt(Fl["ux"].astype(np.float64) ** 2 + Fl["uy"] ** 2 + Fl["uz"] ** 2).max())
            else:
                r["first_nonfinite_step"] = first_bad
            if arith == "f64" and csm == p["csm"]:
                r["max_force_dev_vs_reference"] = float(np.abs(F - ref).max())
            runs[f"{storage}_store_{arith}_arith"] = r
            L.close()
        base = runs["f32_store_f64_arith"].get("stats")
        for k, r in runs.items():
            if base and r.get("stats") and k != "f32_store_f64_arith":
                r["cd_rel_diff_vs_f64"] = abs(r["stats"]["cd"] - base["cd"]) / abs(base["cd"])
        out["Cs=%g" % csm] = runs
    out["reference"] = stats(ref, area)
    out["ensemble_Cs=%g" % p["csm"]] = c2_ensemble(fx, vox, steps, area, ref)
    return out


def c2_ensemble(fx, vox, steps, area, ref, starts=8, amp=1e-6):
    """Each arithmetic path from `starts` starts: the reset state with every
    population offset perturbed by seeded Gaussian noise of `amp` (seeds
    1..starts), so the chaotic wake follows a different trajectory each time.
    Reports how many starts stay finite over `steps` steps (first non-finite
    steps otherwise) and the spread of C_d over the last half of each finite
    run: at tau 0.505 and Cs 0.1 this flow sits on its stability edge, and
    binary64 arithmetic (the reference's own) loses starts as the float paths
    do, so survival and C_d are properties of the ensemble, not of one run."""
    from windtunnel_amd.lbm import LBM
    p = fx["params"]
    b, bi = fx["voxelize"]["body"], fx["voxelize"]["box_in"]
    res = {"starts": starts, "perturbation": f"N(0, {amp:g}) on every population offset, numpy default_rng(seed), "
                                             f"seeds 1..{starts}", "steps": steps,
           "cd": "mean C_x/(u0^2 A/2) over the last half of each finite run",
           "reference_run_cd": float(np.mean(ref[steps // 2:, 0]) / (0.5 * p["u0"] ** 2 * area))}
    for storage, arith in (("f32", "f64"), ("f32", "f32"), ("f16", "f32")):
        bad, cds = [], []
        for seed in range(1, starts + 1):
            L = LBM(*fx["dims"], storage=storage, arith=arith)
            L.set_params(tau=p["tau"], les=p["les"], csm=p["csm"], trt=p["trt"], u0=p["u0"], **{"lambda": p["lambda"]})
            L.reset()
            L.voxelize_grid(vox, b["M"], b["pos"], b["scale"],
                            dict(x0=bi[0], y0=bi[1], z0=bi[2], x1=bi[3], y1=bi[4], z1=bi[5]))
            f = L.pops(0)
            f += (np.random.default_rng(seed).standard_normal(f.size) * amp).astype(f.dtype)
            L.set_pops(f)
            F = L.step(steps)
            fin = np.isfinite(F).all(axis=1)
            if fin.all():
                cds.append(float(np.mean(F[steps // 2:, 0]) / (0.5 * p["u0"] ** 2 * area)))
            else:
                bad.append(int(np.argmin(fin)) + 1)
            L.close()
        res[f"{storage}_store_{arith}_arith"] = dict(
            finite=len(cds), first_nonfinite_steps=bad,
            cd_mean=float(np.mean(cds)) if cds else None, cd_sd=float(np.std(cds)) if cds else None)
    return res


# ------------------------------------------------------------------- scale
def run_scale(N=3000, K=500):
    """640x320x320 TRT+LES in binary16 storage around the seeded leaky
    icosphere (level 7, R = 56 cells, 0.5 % of the triangles deleted),
    voxelized on the GPU at one lattice cell per voxel; N recorded forward
    steps, then the checkpointed adjoint of the mean C_d over the last K
    steps with respect to a gray skin (weight 0.5 on the first fluid layer
    around the body)."""
    from windtunnel_amd.lbm import LBM, Voxel, device_memory
    from windtunnel_amd.synth import noisy_sphere
    dims = (640, 320, 320)
    res = dict(grid="640x320x320", storage="f16", adjoint_fields="binary16, scaled per step", steps=N, tau=0.505,
               u0=0.08, Lambda=3 / 16, Cs=0.14)
    P = noisy_sphere(level=7, R=56, holes=0.005, seed=5)
    t0 = time.perf_counter()
    vox = Voxel.build(P, 128)
    res["voxel"] = dict(tris=vox["stats"]["tris"], solid=vox["stats"]["solid"], openness=vox["stats"]["openness"],
                        dims=[vox["nx"], vox["ny"], vox["nz"]], seconds=round(time.perf_counter() - t0, 3))
    fx = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "fixtures", "vox_leaky_l7.json")
    if os.path.exists(fx):
        import hashlib
        with open(fx) as fh:
            want = json.load(fh)
        res["voxel"]["matches_voxel_js"] = bool(
            hashlib.sha256(np.ascontiguousarray(vox["grid"]).tobytes()).hexdigest() == want["grid_sha256"]
            and vox["stats"]["openness"] == want["stats"]["openness"])
    L = LBM(*dims, storage="f16")
    L.set_params(tau=0.505, les=True, csm=0.14, trt=True, u0=0.08, **{"lambda": 3 / 16})
    L.reset()
    pos = np.array([0.3 * dims[0], dims[1] / 2, dims[2] / 2])
    b = np.array(vox["bbox"])
    lo, hi = pos + b[:3] - 2, pos + b[3:] + 2
    L.voxelize_grid(vox, np.eye(3).reshape(-1), pos, 1.0,
                    dict(x0=lo[0], y0=lo[1], z0=lo[2], x1=hi[0], y1=hi[1], z1=hi[2]))
    res.update(solid_cells=L.solidCells, frontal_area=L.frontalArea)
    solid, near = L.solid_near()
    w = np.where((near == 1) & (solid == 0), 0.5, 0.0).astype(np.float32)
    del solid, near
    res["gray_cells"] = int(np.count_nonzero(w))
    L.set_gray(w)
    del w
    step_ms = L.time_steps(100) / 100
    res.update(kernel_ms=round(step_ms, 4), mlups=round(L.n / (step_ms * 1e-3) / 1e6, 1))
    L.reset()
    q = 0.5 * 0.08 ** 2 * L.frontalArea
    Fbar = np.zeros((N, 3))
    Fbar[N - K:, 0] = 1.0 / (K * q)
    t0 = time.perf_counter()
    F = L.adjoint_forward(N)
    fwd = time.perf_counter() - t0
    free, total = device_memory()
    g, rev_ms = L.adjoint_backward(Fbar)
    L.adjoint_release()
    gray = np.flatnonzero(g)
    res.update(recorded_forward_ms_per_step=round(fwd * 1e3 / N, 4), reverse_ms_per_step=round(rev_ms / N, 4),
               cost_ratio=round((rev_ms / N) / step_ms, 3), finite=bool(np.all(np.isfinite(F)) and np.all(np.isfinite(g))),
               cd=float(np.mean(F[N - K:, 0]) / q), grad_nonzero_cells=int(gray.size),
               grad_abs_max=float(np.abs(g).max()),
               device_gib_total=round(total / 2 ** 30, 1),
               device_gib_in_use_after_forward=round((total - free) / 2 ** 30, 1))
    L.close()
    return res


def _cpu_mask(dims, scale):
    """Solid mask of the mesh body as the GPU lattice holds it (downloaded)."""
    from windtunnel_amd.lbm import LBM
    L = LBM(*dims, storage="f16")
    place_mesh_body(L, scale)
    s = np.asarray(L.solid).reshape(dims[::-1])
    L.close()
    idx = np.argwhere(s == 1)
    lo, hi = idx.min(0), idx.max(0)
    box = [int(lo[2]), int(lo[1]), int(lo[0]), int(hi[2]), int(hi[1]), int(hi[0])]
    return box, (s[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] == 1).astype(np.uint8).reshape(-1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f16", action="store_true")
    ap.add_argument("--no-exact", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-scale", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-drag", action="store_true")
    ap.add_argument("--no-micro", action="store_true")
    ap.add_argument("--scale-steps", type=int, default=3000,
                    help="recorded forward steps of the 640x320x320 adjoint leg (objective over the last sixth)")
    ap.add_argument("--fd-cells", type=int, default=64, help="central-difference cells of the adjoint leg (0: none)")
    ap.add_argument("--adjoint-steps", type=int, default=3000,
                    help="recorded forward steps of the 192x96x96 adjoint leg (objective over the last sixth)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 512 MiB device copy three times (HBM counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        src = torch.ones(512 * 1024 * 1024 // 4, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("copied 3 x 512 MiB", flush=True)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the z-slab halo exchange that would shard the lattice "
              "over several GPUs is not built", file=sys.stderr)
        sys.exit(2)
    wl = WORKLOADS[a.workload]
    nx, ny, nz = wl["dims"]
    traffic = measured_traffic()
    log(f"{a.workload} {nx}x{ny}x{nz} f32 warmup={a.warmup} steps={a.steps}")
    r32 = run_gpu(wl, "f32", a.steps, a.warmup)
    log(f"f32 {r32['mlups']:.0f} MLUPS kernel {r32['kernel_ms']:.4f} ms")
    out = {
        "metric": METRIC, "value": round(r32["mlups"], 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded mesh body, reset() initial state)",
        "config": {"workload": f"D3Q19 TRT(Lambda=3/16)+LES(Cs=0.1) {nx}x{ny}x{nz}, tau=0.505, u0=0.08, "
                               f"synthetic ellipsoid+box mesh body x{wl['scale']} (GPU voxelized, res 200)",
                   "nx": nx, "ny": ny, "nz": nz, "storage": "f32", "arith": "f32", "layout": "ab",
                   "block_cells": r32["block_cells"], "solid_cells": r32["solid_cells"],
                   "frontal_area": r32["frontal_area"], "body": r32["body"]},
        "roofline": {"bound": "hbm", "achieved": round(r32["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r32["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get("f32"), "kernel_ms": round(r32["kernel_ms"], 5),
                     "kernel_launches": r32["kernel_launches"],
                     "kernel_ms_per_launch_events": round(r32["kernel_ms_per_launch_events"], 5),
                     "bytes_per_cell": 152, "bytes_per_launch": nx * ny * nz * 152},
    }
    if not a.no_exact:
        log("f32 storage, binary64 arithmetic")
        rx = run_gpu(wl, "f32", a.steps, a.warmup, arith="f64")
        log(f"exact {rx['mlups']:.0f} MLUPS kernel {rx['kernel_ms']:.4f} ms")
        out["exact"] = {"what": "float32 storage, binary64 collision arithmetic in the reference's expression order "
                                "(bit for bit the reference's populations)",
                        "value": round(rx["mlups"], 1), "unit": "MLUPS", "ms_per_step": round(rx["ms_per_step"], 4),
                        "kernel_ms": round(rx["kernel_ms"], 5), "kernel_launches": rx["kernel_launches"],
                        "roofline_frac": round(rx["achieved_gbs"] / HBM_PEAK_GBS, 4),
                        "kernel_ms_per_launch_events": round(rx["kernel_ms_per_launch_events"], 5),
                        "traffic": traffic.get("exact")}
    if not a.no_f16:
        log("f16")
        r16 = run_gpu(wl, "f16", a.steps, a.warmup)
        log(f"f16 {r16['mlups']:.0f} MLUPS kernel {r16['kernel_ms']:.4f} ms")
        out["fp16"] = {"value": round(r16["mlups"], 1), "unit": "MLUPS", "ms_per_step": round(r16["ms_per_step"], 4),
                       "speedup_vs_fp32": round(r16["mlups"] / r32["mlups"], 3), "block_cells": r16["block_cells"],
                       "roofline": {"bound": "hbm", "achieved": round(r16["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": round(r16["achieved_gbs"] / HBM_PEAK_GBS, 4),
                                    "traffic": traffic.get("f16"), "kernel_ms": round(r16["kernel_ms"], 5),
                                    "kernel_launches": r16["kernel_launches"],
                                    "kernel_ms_per_launch_events": round(r16["kernel_ms_per_launch_events"], 5),
                                    "bytes_per_cell": 76, "bytes_per_launch": nx * ny * nz * 76}}
    if not a.no_micro:
        log("micro 512x256x256, no body, random populations")
        out["micro"] = run_micro()
        log(f"micro: {json.dumps(out['micro'])}")
    if not a.no_precision:
        log("precision: 128x64x64 laminar and LES (f64/f32/f16), 512x256x256 f16 vs f32, 4000 steps")
        out["precision"] = run_precision()
        log(f"precision: {json.dumps(out['precision'])}")
    if not a.no_adjoint:
        log(f"adjoint 192x96x96, {a.adjoint_steps} steps")
        out["adjoint"] = run_adjoint(a.fd_cells, a.adjoint_steps, max(1, a.adjoint_steps // 6))
    if not a.no_c1:
        log("c1 96x48x48 BGK sphere, reference and channel boundaries, 2000 steps")
        out["c1"] = run_c1()
        log(f"c1: {json.dumps(out['c1'])}")
    if not a.no_c2:
        log("c2 256x128x128 TRT+LES wake of the mesh body, 5000 steps: binary64, float32 and binary16 paths")
        out["c2"] = run_c2()
        log(f"c2: {json.dumps(out['c2'])}")
    if not a.no_drag:
        log("drag: re-voxelize the 512x256x256 body moved by one cell")
        out["drag"] = run_drag()
        log(f"drag: {json.dumps(out['drag'])}")
    if not a.no_scale:
        log("scale 640x320x320 f16, level-7 leaky icosphere, 3000 steps + adjoint")
        out["scale"] = run_scale(a.scale_steps, max(1, a.scale_steps // 6))
        log(f"scale: {json.dumps(out['scale'])}")
    if not a.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline, {threads} threads")
        v, v1 = run_cpu(wl, threads)
        log("cpu detail: 1 and all threads, 96x48x48 and 256x128x128")
        out["cpu_detail"] = run_cpu_detail(threads)
        out["cpu_detail"][f"c3_{nx}x{ny}x{nz}_trt_les_mlups"] = {
            "1_threads": round(v1, 2), f"{threads}_threads": round(v, 2),
            "protocol": "12 warm-up steps, best of 3 x 20 steps; 1 thread: 1 warm-up step, best of 2 x 1 step"}
        out["cpu_baseline"] = {"value": round(v, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
                               "sample": f"oracle (C restatement of js/lbm.js, binary64 arithmetic, float32 "
                                         f"storage) on the full {nx}x{ny}x{nz} lattice with the same body, "
                                         f"{threads} z-slab threads, best of 3x20 steps after 12 warm-up steps; "
                                         f"host {cpu_model()}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
