"""Python host binding of libwindtunnel: the reference's solver object and
voxelizer over the C ABI (include/windtunnel.h).

``LBM`` mirrors ``class LBM`` of js/lbm.js:51-546 -- constructor, settable
solver properties, ``reset``, ``voxelize``, ``step``, ``stepSlab``-style
worker stepping, the macroscopic fields, ``sample`` and ``isSolid`` -- so
scripts written against the reference read the same. ``Voxel`` mirrors
``Voxel.build`` / ``Voxel.bounds`` (js/voxel.js:32-102).

There is no CPU fallback: if the HIP library is missing or no gfx950 device
is present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WT_LIB_PATH") or os.path.join(HERE, "libwindtunnel.so")

STORE = {"f32": 0, "f16": 1, "f64": 2}
ARITH = {"f32": 0, "f64": 1}
LAYOUT = {"ab": 0, "esoteric": 1}

E_ARG, E_DEVICE, E_NOMEM, E_STATE, E_FAULT = -1, -2, -3, -4, -5
E_VOX = (-11, -12, -13, -14)


class WindtunnelError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class Params(C.Structure):
    _fields_ = [("tau", C.c_double), ("csm", C.c_double), ("lam", C.c_double), ("u0", C.c_double),
                ("les", C.c_int), ("trt", C.c_int), ("movingWalls", C.c_int)]


class BodyStats(C.Structure):
    _fields_ = [("frontalArea", C.c_int64), ("planformArea", C.c_int64), ("solidCells", C.c_int64),
                ("rideHeight", C.c_int64), ("movingWalls", C.c_int32), ("ground", C.c_int32)]


class State(C.Structure):
    _fields_ = [("parity", C.c_int64), ("steps", C.c_int64), ("force", C.c_double * 3),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("storage", C.c_int32), ("arith", C.c_int32), ("layout", C.c_int32),
                ("block_cells", C.c_int32), ("reserved", C.c_int32)]


class Part(C.Structure):
    _fields_ = [("type", C.c_int32), ("wheel", C.c_int32), ("has_naca", C.c_int32), ("reserved", C.c_int32),
                ("c", C.c_double * 3), ("s", C.c_double * 3), ("m", C.c_double * 9), ("naca", C.c_double * 3)]


PRIM = {"box": 0, "ell": 1, "cyl": 2, "cone": 3, "foil": 4}


class VoxMeta(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("ntri", C.c_int32),
                ("cell", C.c_double), ("inv", C.c_double), ("min", C.c_double * 3),
                ("ext", C.c_double * 3), ("centre", C.c_double * 3), ("bbox", C.c_double * 6),
                ("solid", C.c_int64), ("disagree", C.c_int64), ("openness", C.c_double)]


# every symbol the header declares, with its ctypes signature
_P, _I, _D, _L = C.c_void_p, C.c_int, C.c_double, C.c_int64
SIGNATURES = {
    "wt_last_error": (C.c_char_p, []),
    "wt_abi_version": (_I, []),
    "wt_create": (_I, [_I, _I, _I, _I, _I, _I, _P]),
    "wt_destroy": (_I, [_P]),
    "wt_set_params": (_I, [_P, _P]),
    "wt_get_params": (_I, [_P, _P]),
    "wt_reset": (_I, [_P]),
    "wt_voxelize": (_I, [_P, _P, _P, _P, _I, _D, _P]),
    "wt_voxelize_grid": (_I, [_P, _P, _P, _P, _P, _D, _P, _I, _D, _P]),
    "wt_step": (_I, [_P, _I, _P]),
    "wt_step_slabs": (_I, [_P, _I, _I, _P]),
    "wt_set_parity": (_I, [_P, _I]),
    "wt_get_state": (_I, [_P, _P]),
    "wt_get_fields": (_I, [_P, _P, _P, _P, _P]),
    "wt_get_pops": (_I, [_P, _I, _P]),
    "wt_get_solid": (_I, [_P, _P, _P]),
    "wt_time_steps": (_I, [_P, _I, _P]),
    "wt_device_memory": (_I, [_P, _P]),
    "wt_set_boundary_mode": (_I, [_P, _I]),
    "wt_slice": (_I, [_P, _I, _I, _I, _P, _P]),
    "wt_set_pops": (_I, [_P, _P]),
    "wt_mesh_parse": (_I, [C.c_char_p, _P, C.c_int64, _P, _P]),
    "wt_sample": (_I, [_P, _P, C.c_int64, _P]),
    "wt_sample_noalloc": (_I, [_P, _P, C.c_int64, _P]),
    "wt_set_timing": (_I, [_P, _I]),
    "wt_get_timing": (_I, [_P, _P, _P]),
    "wt_set_gray": (_I, [_P, _P]),
    "wt_voxelize_parts": (_I, [_P, _P, _I, _P, _P, _D, _P, _I, _D, _D, _P]),
    "wt_get_walls": (_I, [_P, _P, _P, _P]),
    "wt_step_slab": (_I, [_P, _I, _I, _P]),
    "wt_boundaries": (_I, [_P, _I, _I, _I, _I]),
    "wt_adjoint": (_I, [_P, _I, _P, _P, _P, _I, _P]),
    "wt_adjoint_forward": (_I, [_P, _I, _I, _P]),
    "wt_adjoint_backward": (_I, [_P, _I, _P, _P, _P]),
    "wt_adjoint_release": (_I, [_P]),
    "wt_adjoint_record": (_I, [_P, _P]),
    "wt_voxel_bounds": (_I, [_P, _L, _P]),
    "wt_voxel_plan": (_I, [_P, _L, _I, _P]),
    "wt_voxel_build": (_I, [_P, _L, _I, _P, _P]),
}

_lib = None


def lib():
    """Load libwindtunnel.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run: python -m windtunnel_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        msg = lib().wt_last_error().decode()
        if rc in E_VOX or rc == E_ARG:
            raise ValueError(msg)
        raise WindtunnelError(rc, msg)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# velocity set (js/lbm.js:28-40)
Q = 19
EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0], dtype=np.int32)
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1], dtype=np.int32)
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1], dtype=np.int32)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
FLUID, SOLID_OBJ, SOLID_WALL = 0, 1, 2


def mesh_parse(name: str, data: bytes) -> np.ndarray:
    """MeshIO.parse (wt_mesh_parse): a model file's bytes (STL, OBJ, GLB) into
    a float32 triangle soup of 9 values per triangle; ValueError with the
    reference's message for unreadable files."""
    buf = np.frombuffer(data, dtype=np.uint8)
    cnt = C.c_int64(0)
    nm = name.encode()
    _check(lib().wt_mesh_parse(nm, _ptr(buf), int(buf.size), None, C.byref(cnt)))
    out = np.empty(cnt.value, dtype=np.float32)
    _check(lib().wt_mesh_parse(nm, _ptr(buf), int(buf.size), _ptr(out), C.byref(cnt)))
    return out


def device_memory() -> tuple:
    """(free, total) bytes of the current device's memory (wt_device_memory)."""
    f, t = C.c_uint64(), C.c_uint64()
    _check(lib().wt_device_memory(C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


def _toint32(v: float) -> int:
    """ECMAScript ToInt32 (the `|0` operator)."""
    if not math.isfinite(v):
        return 0
    t = int(math.trunc(v)) & 0xFFFFFFFF
    return t - (1 << 32) if t >= (1 << 31) else t


class LBM:
    """D3Q19 solver on one MI355X with the reference object's interface."""

    Q, EX, EY, EZ, W = Q, EX, EY, EZ, W
    FLUID, SOLID_OBJ, SOLID_WALL = FLUID, SOLID_OBJ, SOLID_WALL

    def __init__(self, nx: int, ny: int, nz: int, storage: str = "f32", arith: str = "f32",
                 layout: str = "ab"):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.n = self.nx * self.ny * self.nz
        self.storage, self.arith, self.layout = storage, arith, layout
        h = C.c_void_p()
        _check(lib().wt_create(self.nx, self.ny, self.nz, STORE[storage], ARITH[arith], LAYOUT[layout],
                               C.byref(h)))
        self._h = h
        self._p = Params()
        _check(lib().wt_get_params(self._h, C.byref(self._p)))
        self.frontalArea = 1
        self.planformArea = 1
        self.solidCells = 0
        self.rideHeight = 0
        self.ground = False
        self._fields = None

    def close(self):
        if getattr(self, "_h", None):
            lib().wt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- settable solver properties (js/lbm.js:106-111) ----------------
    def _set(self, name, value):
        setattr(self._p, name, value)
        _check(lib().wt_set_params(self._h, C.byref(self._p)))

    tau = property(lambda s: s._p.tau, lambda s, v: s._set("tau", float(v)))
    csm = property(lambda s: s._p.csm, lambda s, v: s._set("csm", float(v)))
    lam = property(lambda s: s._p.lam, lambda s, v: s._set("lam", float(v)))
    u0 = property(lambda s: s._p.u0, lambda s, v: s._set("u0", float(v)))
    les = property(lambda s: bool(s._p.les), lambda s, v: s._set("les", int(bool(v))))
    trt = property(lambda s: bool(s._p.trt), lambda s, v: s._set("trt", int(bool(v))))
    movingWalls = property(lambda s: bool(s._p.movingWalls), lambda s, v: s._set("movingWalls", int(bool(v))))

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self, "lam" if k == "lambda" else k, v)

    @property
    def nu(self) -> float:
        return (self.tau - 0.5) / 3

    def idx(self, x, y, z) -> int:
        return x + self.nx * (y + self.ny * z)

    # ---- state -----------------------------------------------------------
    def _state(self) -> State:
        s = State()
        _check(lib().wt_get_state(self._h, C.byref(s)))
        return s

    @property
    def parity(self) -> int:
        return int(self._state().parity)

    @parity.setter
    def parity(self, v: int):
        _check(lib().wt_set_parity(self._h, int(v)))

    @property
    def steps(self) -> int:
        return int(self._state().steps)

    @property
    def block_cells(self) -> int:
        """Cells per storage block of the population buffers (0: planar)."""
        return int(self._state().block_cells)

    @property
    def force(self) -> np.ndarray:
        return np.array(self._state().force[:])

    def fields(self) -> dict:
        """rho, ux, uy, uz of the last completed step (downloaded once per step)."""
        if self._fields is None:
            dt = np.float64 if self.storage == "f64" else np.float32
            out = {k: np.empty(self.n, dtype=dt) for k in ("rho", "ux", "uy", "uz")}
            _check(lib().wt_get_fields(self._h, _ptr(out["rho"]), _ptr(out["ux"]), _ptr(out["uy"]),
                                       _ptr(out["uz"])))
            self._fields = out
        return self._fields

    rho = property(lambda s: s.fields()["rho"])
    ux = property(lambda s: s.fields()["ux"])
    uy = property(lambda s: s.fields()["uy"])
    uz = property(lambda s: s.fields()["uz"])

    def pops(self, which: int = 0) -> np.ndarray:
        dt = np.float64 if self.storage == "f64" else np.float32
        out = np.empty(19 * self.n, dtype=dt)
        _check(lib().wt_get_pops(self._h, int(which), _ptr(out)))
        return out

    def set_pops(self, f) -> None:
        """Current populations from deviations in the f[q*n + cell] order
        (wt_set_pops): float, or double for f64 storage."""
        dt = np.float64 if self.storage == "f64" else np.float32
        a = np.ascontiguousarray(f, dtype=dt).reshape(-1)
        if a.size != 19 * self.n:
            raise ValueError(f"populations must hold 19 x {self.n} values")
        _check(lib().wt_set_pops(self._h, _ptr(a)))
        self._fields = None

    f = property(lambda s: s.pops(0))
    g = property(lambda s: s.pops(1))

    def solid_near(self):
        s = np.empty(self.n, dtype=np.uint8)
        nr = np.empty(self.n, dtype=np.uint8)
        _check(lib().wt_get_solid(self._h, _ptr(s), _ptr(nr)))
        return s, nr

    solid = property(lambda s: s.solid_near()[0])
    near = property(lambda s: s.solid_near()[1])

    # ---- operations --------------------------------------------------------
    def reset(self):
        """Equilibrium at u0 in fluid cells, at rest in solid ones (js/lbm.js:164-174)."""
        _check(lib().wt_reset(self._h))
        self._fields = None

    def clamp_box(self, box: dict) -> list:
        """The body box clamped to the interior as voxelize does (js/lbm.js:216-218)."""
        return [max(1, math.floor(box["x0"])), max(1, math.floor(box["y0"])), max(1, math.floor(box["z0"])),
                min(self.nx - 2, math.ceil(box["x1"])), min(self.ny - 2, math.ceil(box["y1"])),
                min(self.nz - 2, math.ceil(box["z1"]))]

    def _apply_stats(self, st: BodyStats):
        self.frontalArea = int(st.frontalArea)
        self.planformArea = int(st.planformArea)
        self.solidCells = int(st.solidCells)
        self.rideHeight = int(st.rideHeight)
        self.ground = bool(st.ground)
        self._p.movingWalls = int(st.movingWalls)
        self._fields = None

    def voxelize_mask(self, box, mask, wv=None, ground=False, roadVel=0.0):
        """voxelize with the inside/wallVel closures already evaluated over
        the clamped box (x fastest): mask 0/1/2, wv 3 floats per mask==2 cell."""
        box = np.ascontiguousarray(box, dtype=np.int32)
        mask = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        cnt = max(0, box[3] - box[0] + 1) * max(0, box[4] - box[1] + 1) * max(0, box[5] - box[2] + 1)
        if mask.size != cnt:
            raise ValueError(f"mask has {mask.size} cells, box holds {cnt}")
        wv = np.ascontiguousarray(wv if wv is not None else np.zeros(0), dtype=np.float32).reshape(-1)
        if wv.size != 3 * int(np.count_nonzero(mask == 2)):
            raise ValueError("wv must hold 3 floats per wall-velocity cell")
        st = BodyStats()
        _check(lib().wt_voxelize(self._h, _ptr(box), _ptr(mask), _ptr(wv), int(bool(ground)),
                                 float(roadVel or 0.0), C.byref(st)))
        self._apply_stats(st)

    def voxelize(self, inside, box, ground=False, wallVel=None, roadVel=0.0):
        """Rasterise a body given as a point test (js/lbm.js:191-266)."""
        b = self.clamp_box(box)
        mask = []
        wv = []
        out = [0.0, 0.0, 0.0]
        for z in range(b[2], b[5] + 1):
            for y in range(b[1], b[4] + 1):
                for x in range(b[0], b[3] + 1):
                    if not inside(x + 0.5, y + 0.5, z + 0.5):
                        mask.append(0)
                        continue
                    if wallVel is not None and wallVel(x + 0.5, y + 0.5, z + 0.5, out):
                        mask.append(2)
                        wv.extend(out)
                    else:
                        mask.append(1)
        self.voxelize_mask(b, np.array(mask, dtype=np.uint8), np.array(wv, dtype=np.float32), ground, roadVel)

    def voxelize_parts(self, parts, M, pos, scale, box, ground=False, roadVel=0.0, wheel_u0=0.0):
        """Analytic body (Body.makeTest / makeWheelVel, js/shapes.js:365-470) on
        the GPU. parts: dicts {t: box|ell|cyl|cone|foil, c, s, m (3x3 row-major
        primitive -> body rotation), naca (optional), wheel (optional)}."""
        arr = (Part * max(1, len(parts)))()
        for i, p in enumerate(parts):
            q = arr[i]
            q.type = PRIM[p["t"]]
            q.wheel = int(bool(p.get("wheel", False)))
            s = list(p["s"]) + [0.0] * (3 - len(p["s"]))
            for k in range(3):
                q.c[k] = float(p["c"][k])
                q.s[k] = float(s[k])
            m = np.asarray(p.get("m", np.eye(3)), dtype=np.float64).reshape(-1)
            for k in range(9):
                q.m[k] = float(m[k])
            if p.get("naca") is not None:
                q.has_naca = 1
                for k in range(3):
                    q.naca[k] = float(p["naca"][k])
        b = np.ascontiguousarray(self.clamp_box(box) if isinstance(box, dict) else box, dtype=np.int32)
        M = np.ascontiguousarray(M, dtype=np.float64)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        st = BodyStats()
        _check(lib().wt_voxelize_parts(self._h, C.cast(arr, C.c_void_p), len(parts), _ptr(M), _ptr(pos), float(scale),
                                       _ptr(b), int(bool(ground)), float(roadVel or 0.0), float(wheel_u0 or 0.0),
                                       C.byref(st)))
        self._apply_stats(st)
        self._fields = None

    def walls(self):
        """Wall velocities (uwx, uwy, uwz) per cell (js/lbm.js:90-93)."""
        w = np.empty((3, self.n), dtype=np.float32)
        _check(lib().wt_get_walls(self._h, _ptr(w[0]), _ptr(w[1]), _ptr(w[2])))
        return w

    def voxelize_grid(self, vox: dict, M, pos, scale, box, ground=False, roadVel=0.0):
        """Grid-backed body (js/shapes.js:369-388) evaluated on the GPU."""
        b = np.ascontiguousarray(self.clamp_box(box) if isinstance(box, dict) else box, dtype=np.int32)
        grid = np.ascontiguousarray(vox["grid"], dtype=np.uint8)
        vm = VoxMeta()
        vm.nx, vm.ny, vm.nz = vox["nx"], vox["ny"], vox["nz"]
        vm.inv = vox["inv"]
        vm.cell = vox["cell"]
        for k in range(3):
            vm.min[k] = vox["min"][k]
        M = np.ascontiguousarray(M, dtype=np.float64)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        st = BodyStats()
        _check(lib().wt_voxelize_grid(self._h, _ptr(grid), C.byref(vm), _ptr(M), _ptr(pos), float(scale),
                                      _ptr(b), int(bool(ground)), float(roadVel or 0.0), C.byref(st)))
        self._apply_stats(st)

    def step(self, n: int = 1) -> np.ndarray:
        """Advance n steps (js/lbm.js:285-291); returns the per-step forces [n,3]."""
        f = np.empty((n, 3))
        _check(lib().wt_step(self._h, int(n), _ptr(f)))
        self._fields = None
        return f

    def step_slabs(self, n: int, nslabs: int) -> np.ndarray:
        """Worker-pool stepping: per-(step, slab) forces [n, nslabs, 3]."""
        f = np.empty((n, nslabs, 3))
        _check(lib().wt_step_slabs(self._h, int(n), int(nslabs), _ptr(f)))
        self._fields = None
        return f

    def step_slab(self, z0: int, z1: int) -> np.ndarray:
        """stepSlab(z0, z1): rows [z0,z1) with their x/y planes, no parity
        flip; returns the slab's force (the reference's _sf)."""
        f = np.zeros(3)
        _check(lib().wt_step_slab(self._h, int(z0), int(z1), _ptr(f)))
        self._fields = None
        return f

    def boundaries(self, z0=None, z1=None, doLo=None, doHi=None) -> None:
        """_boundaries(g, z0, z1, doLo, doHi) after step_slab."""
        whole = z0 is None
        a = 0 if z0 is None else int(z0)
        b = self.nz if z1 is None else int(z1)
        lo = True if whole else bool(doLo)
        hi = True if z1 is None else bool(doHi)
        _check(lib().wt_boundaries(self._h, a, b, int(lo), int(hi)))
        self._fields = None

    SLICE_FIELDS = {"speed": 0, "pressure": 1, "vorticity": 2}

    def slice(self, axis: str, k: int, field: str):
        """The app's slice view on the device (wt_slice, js/main.js:925-965):
        axis 'xy' (plane z = k) or 'xz' (plane y = k); field 'speed' (|u|/u0),
        'pressure' (C_p) or 'vorticity'. Returns (values [h, nx] float32,
        solid [h, nx] uint8)."""
        ax = {"xy": 0, "xz": 1}[axis]
        h = self.ny if ax == 0 else self.nz
        out = np.empty((h, self.nx), dtype=np.float32)
        sol = np.empty((h, self.nx), dtype=np.uint8)
        _check(lib().wt_slice(self._h, ax, int(k), self.SLICE_FIELDS[field], _ptr(out), _ptr(sol)))
        return out, sol

    def sample_many(self, points) -> np.ndarray:
        """sample() (js/lbm.js:515-539) at many points on the device
        (wt_sample): points [..., 3] lattice coordinates -> [..., 3] velocity."""
        p = np.ascontiguousarray(points, dtype=np.float64)
        if p.shape[-1:] != (3,):
            raise ValueError("points must have a last axis of 3")
        out = np.empty_like(p)
        _check(lib().wt_sample(self._h, _ptr(p), int(p.size // 3), _ptr(out)))
        return out

    def set_boundary_mode(self, mode: str) -> None:
        """'reference' (js/lbm.js:473-512) or 'channel': free-slip side walls
        and an equilibrium outflow (wt_set_boundary_mode)."""
        modes = {"reference": 0, "channel": 1}
        if mode not in modes:
            raise ValueError(f"unknown boundary mode {mode!r}")
        _check(lib().wt_set_boundary_mode(self._h, modes[mode]))
        self._fields = None

    def time_steps(self, n: int) -> float:
        """Device milliseconds for n steps, from HIP events on the lattice stream."""
        ms = C.c_double()
        _check(lib().wt_time_steps(self._h, int(n), C.byref(ms)))
        self._fields = None
        return ms.value

    def set_timing(self, on: bool):
        _check(lib().wt_set_timing(self._h, int(bool(on))))

    def timing(self):
        """(total step-kernel milliseconds, launches) since timing was turned on."""
        ms = C.c_double()
        cnt = C.c_int64()
        _check(lib().wt_get_timing(self._h, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def set_gray(self, w):
        """Gray bounce-back weights per cell (0 fluid, (0,1) gray); None clears."""
        if w is None:
            _check(lib().wt_set_gray(self._h, None))
            return
        w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
        if w.size != self.n:
            raise ValueError("one weight per cell")
        _check(lib().wt_set_gray(self._h, _ptr(w)))

    def adjoint(self, nsteps: int, Fbar, interval: int = 0):
        """Run nsteps and the checkpointed reverse sweep for J = sum_k Fbar[k].F_k.

        Returns (forces [nsteps,3], dJ/dw [n], reverse-sweep device ms)."""
        Fbar = np.ascontiguousarray(Fbar, dtype=np.float64).reshape(nsteps, 3)
        g = np.zeros(self.n)
        f = np.zeros((nsteps, 3))
        ms = C.c_double()
        _check(lib().wt_adjoint(self._h, int(nsteps), _ptr(Fbar), _ptr(g), _ptr(f), int(interval), C.byref(ms)))
        self._fields = None
        return f, g, ms.value

    def adjoint_forward(self, nsteps: int, interval: int = 0) -> np.ndarray:
        """Run nsteps like step() and record checkpoints for adjoint_backward.
        Returns the forces [nsteps, 3]."""
        f = np.zeros((int(nsteps), 3))
        _check(lib().wt_adjoint_forward(self._h, int(nsteps), int(interval), _ptr(f)))
        self._fields = None
        self._adj_steps = int(nsteps)
        return f

    def adjoint_backward(self, Fbar):
        """dJ/dw for J = sum_k Fbar[k].F_k over the recorded forward pass.
        Returns (dJ/dw [n], reverse-sweep device ms)."""
        n = getattr(self, "_adj_steps", 0)
        Fbar = np.ascontiguousarray(Fbar, dtype=np.float64)
        if Fbar.size != 3 * n:
            raise ValueError(f"Fbar must hold {n} x 3 values (the recorded forward's forces)")
        g = np.zeros(self.n)
        ms = C.c_double()
        _check(lib().wt_adjoint_backward(self._h, n, _ptr(Fbar), _ptr(g), C.byref(ms)))
        return g, ms.value

    def adjoint_record(self) -> int:
        """Id of the recorded forward pass (increases with every adjoint_forward)."""
        v = C.c_int64(0)
        _check(lib().wt_adjoint_record(self._h, C.byref(v)))
        return int(v.value)

    def adjoint_release(self) -> None:
        _check(lib().wt_adjoint_release(self._h))
        self._adj_steps = 0

    # ---- point queries (js/lbm.js:515-545) -------------------------------
    def sample(self, x, y, z, out=None):
        """sample() at one point: the device's batched sample (wt_sample) with
        a batch of one, on the fields of the last step."""
        v = self.sample_many(np.array([x, y, z], dtype=np.float64))
        out = out if out is not None else [0.0, 0.0, 0.0]
        out[0], out[1], out[2] = float(v[0]), float(v[1]), float(v[2])
        return out

    def isSolid(self, x, y, z) -> bool:
        xi, yi, zi = _toint32(x), _toint32(y), _toint32(z)
        if xi < 0 or yi < 0 or zi < 0 or xi >= self.nx or yi >= self.ny or zi >= self.nz:
            return False
        return bool(self.solid[xi + self.nx * (yi + self.ny * zi)] != 0)


class Voxel:
    """Voxel.bounds / Voxel.build (js/voxel.js:32-102) on the GPU."""

    @staticmethod
    def bounds(P) -> list:
        P = np.ascontiguousarray(P, dtype=np.float32).reshape(-1)
        b = np.zeros(6)
        _check(lib().wt_voxel_bounds(_ptr(P), P.size, _ptr(b)))
        return list(b)

    @staticmethod
    def build(P, res: int = 96) -> dict:
        P = np.ascontiguousarray(P, dtype=np.float32).reshape(-1)
        vm = VoxMeta()
        _check(lib().wt_voxel_plan(_ptr(P), P.size, int(res or 96), C.byref(vm)))
        grid = np.zeros(vm.nx * vm.ny * vm.nz, dtype=np.uint8)
        _check(lib().wt_voxel_build(_ptr(P), P.size, int(res or 96), C.byref(vm), _ptr(grid)))
        return dict(grid=grid, nx=vm.nx, ny=vm.ny, nz=vm.nz, cell=vm.cell, min=list(vm.min), inv=vm.inv,
                    ext=list(vm.ext), centre=list(vm.centre), bbox=list(vm.bbox),
                    stats=dict(tris=vm.ntri, solid=int(vm.solid), openness=vm.openness),
                    disagree=int(vm.disagree))
