"""ctypes binding of the CPU oracle (libwt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke`` and
the CPU-baseline leg of ``bench.py``, never by the product package.

The oracle restates js/lbm.js, js/lbmWorker.js, js/pool.js and js/voxel.js
in C with binary64 arithmetic in the reference's operation order; see
oracle/wt_oracle.c for the per-function citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwt_oracle.so")

_libs = {}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(path: str | None = None):
    """The oracle library (or a variant of it built elsewhere, e.g. with a
    mistake planted by a test)."""
    key = path or LIB_PATH
    if key not in _libs:
        if path is None and not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(key)
        P, I, D, V = C.c_void_p, C.c_int, C.c_double, None
        sig = {
            "wto_create": (P, [I, I, I, I]),
            "wto_destroy": (V, [P]),
            "wto_set_params": (V, [P, D, I, D, I, D, D, I]),
            "wto_reset": (V, [P]),
            "wto_voxelize": (V, [P, P, P, P, I, D, P]),
            "wto_step": (V, [P, I, P]),
            "wto_step_pool": (I, [P, I, I, I, P]),
            "wto_step_slab": (V, [P, I, I, P]),
            "wto_boundaries": (V, [P, I, I, I, I, I]),
            "wto_flip": (V, [P]),
            "wto_array": (P, [P, I]),
            "wto_state": (V, [P, P, P]),
            "wto_sample": (V, [P, D, D, D, P]),
            "wto_is_solid": (I, [P, D, D, D]),
            "wto_voxel_bounds": (V, [P, C.c_long, P]),
            "wto_voxel_plan": (I, [P, C.c_long, I, P]),
            "wto_voxel_build": (I, [P, C.c_long, I, P, P]),
            "wto_grid_mask": (V, [P, P, P, D, P, P, D, P, P]),
            "wto_set_gray": (I, [P, P]),
            "wto_set_channel": (V, [P, I]),
            "wto_adjoint": (I, [P, I, P, P, P]),
            "wto_collide_adj": (V, [P, P, D, D, D, I, I, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[key] = L
    return _libs[key]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


ARRAYS = {"f": 0, "g": 1, "solid": 2, "near": 3, "ux": 4, "uy": 5, "uz": 6, "rho": 7,
          "uwx": 8, "uwy": 9, "uwz": 10, "bufA": 11, "bufB": 12}


class OracleLBM:
    """The reference ``LBM`` object restated on the CPU (js/lbm.js:51-546)."""

    def __init__(self, nx: int, ny: int, nz: int, prec: int = 4, libpath: str | None = None):
        self.nx, self.ny, self.nz, self.prec = nx, ny, nz, prec
        self.n = nx * ny * nz
        self._L = lib(libpath)
        self._h = self._L.wto_create(nx, ny, nz, prec)
        if not self._h:
            raise ValueError("bad lattice dimensions")
        self._params = dict(tau=0.5035, les=True, csm=0.14, trt=True, **{"lambda": 0.25},
                            u0=0.06, movingWalls=False)

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.wto_destroy(self._h)
            self._h = None

    # -- parameters -----------------------------------------------------
    def set_params(self, **kw):
        p = self._params
        for k, v in kw.items():
            if k not in p:
                raise KeyError(k)
            p[k] = v
        self._L.wto_set_params(self._h, float(p["tau"]), int(bool(p["les"])), float(p["csm"]),
                             int(bool(p["trt"])), float(p["lambda"]), float(p["u0"]),
                             int(bool(p["movingWalls"])))

    def params(self) -> dict:
        return dict(self._params)

    def set_boundary_mode(self, mode: str):
        """'reference' (js/lbm.js:473-512) or 'channel' (free-slip side walls,
        equilibrium outflow; applied after the reference boundaries)."""
        if mode not in ("reference", "channel"):
            raise ValueError(mode)
        self._L.wto_set_channel(self._h, int(mode == "channel"))

    # -- state ----------------------------------------------------------
    def array(self, name: str) -> np.ndarray:
        which = ARRAYS[name]
        p = self._L.wto_array(self._h, which)
        if name in ("solid", "near"):
            dt, cnt = np.uint8, self.n
        elif name.startswith("uw"):
            dt, cnt = np.float32, self.n
        elif name in ("f", "g", "bufA", "bufB"):
            dt, cnt = (np.float32 if self.prec == 4 else np.float64), 19 * self.n
        else:
            dt, cnt = (np.float32 if self.prec == 4 else np.float64), self.n
        buf = (C.c_char * (cnt * np.dtype(dt).itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=dt, count=cnt)

    def state(self) -> dict:
        s = np.zeros(8, dtype=np.int64)
        f = np.zeros(3)
        self._L.wto_state(self._h, _ptr(s), _ptr(f))
        keys = ["parity", "steps", "frontalArea", "planformArea", "solidCells", "rideHeight",
                "movingWalls", "ground"]
        d = {k: int(v) for k, v in zip(keys, s)}
        d["force"] = f
        return d

    # -- operations ------------------------------------------------------
    def reset(self):
        self._L.wto_reset(self._h)

    def voxelize_mask(self, box, mask, wv=None, ground=False, roadVel=0.0) -> dict:
        box = np.ascontiguousarray(box, dtype=np.int32)
        mask = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        wv = np.ascontiguousarray(wv if wv is not None else np.zeros(0), dtype=np.float32)
        stats = np.zeros(5, dtype=np.int64)
        self._L.wto_voxelize(self._h, _ptr(box), _ptr(mask), _ptr(wv), int(bool(ground)),
                           float(roadVel), _ptr(stats))
        self._params["movingWalls"] = bool(stats[4])
        return dict(frontalArea=int(stats[0]), planformArea=int(stats[1]),
                    solidCells=int(stats[2]), rideHeight=int(stats[3]), movingWalls=bool(stats[4]))

    def step(self, n: int = 1) -> np.ndarray:
        f = np.zeros((n, 3))
        self._L.wto_step(self._h, n, _ptr(f))
        return f

    def step_pool(self, n: int, nworkers: int, threaded: bool = False) -> np.ndarray:
        f = np.zeros((n, 3))
        rc = self._L.wto_step_pool(self._h, n, nworkers, int(threaded), _ptr(f))
        if rc:
            raise ValueError("bad worker count")
        return f

    def sample(self, x, y, z):
        out = np.zeros(3)
        self._L.wto_sample(self._h, float(x), float(y), float(z), _ptr(out))
        return out

    def set_gray(self, w):
        """Gray bounce-back weights per cell (0 fluid, (0,1) gray); None clears."""
        if w is None:
            rc = self._L.wto_set_gray(self._h, None)
        else:
            self._w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
            rc = self._L.wto_set_gray(self._h, _ptr(self._w))
        if rc:
            raise ValueError("gray weight >= 1 on a fluid cell")

    def adjoint(self, nsteps: int, Fbar: np.ndarray):
        """(forces [n,3], dJ/dw [n]) for J = sum_k Fbar[k].F_k over nsteps steps."""
        Fbar = np.ascontiguousarray(Fbar, dtype=np.float64).reshape(nsteps, 3)
        g = np.zeros(self.n)
        f = np.zeros((nsteps, 3))
        rc = self._L.wto_adjoint(self._h, int(nsteps), _ptr(Fbar), _ptr(g), _ptr(f))
        if rc:
            raise ValueError("adjoint needs FP64 storage and no moving walls")
        return f, g

    def is_solid(self, x, y, z) -> bool:
        return bool(self._L.wto_is_solid(self._h, float(x), float(y), float(z)))


class VoxMeta(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ntri", C.c_int),
                ("cell", C.c_double), ("inv", C.c_double), ("min", C.c_double * 3),
                ("ext", C.c_double * 3), ("centre", C.c_double * 3), ("bbox", C.c_double * 6),
                ("solid", C.c_long), ("disagree", C.c_long), ("openness", C.c_double)]


VOXEL_ERRORS = {-1: "Mesh has no triangles.", -2: "Mesh is degenerate (zero size).",
                -3: "Voxel grid too large; lower the resolution."}


def voxel_bounds(P: np.ndarray) -> list:
    P = np.ascontiguousarray(P, dtype=np.float32).reshape(-1)
    b = np.zeros(6)
    lib().wto_voxel_bounds(_ptr(P), P.size, _ptr(b))
    return list(b)


def voxel_build(P: np.ndarray, res: int = 96) -> dict:
    """``Voxel.build(P,res)`` restated (js/voxel.js:48-102)."""
    P = np.ascontiguousarray(P, dtype=np.float32).reshape(-1)
    m = VoxMeta()
    rc = lib().wto_voxel_plan(_ptr(P), P.size, int(res), C.byref(m))
    if rc:
        raise ValueError(VOXEL_ERRORS[rc])
    grid = np.zeros(m.nx * m.ny * m.nz, dtype=np.uint8)
    lib().wto_voxel_build(_ptr(P), P.size, int(res), C.byref(m), _ptr(grid))
    return dict(grid=grid, nx=m.nx, ny=m.ny, nz=m.nz, cell=m.cell, min=list(m.min), inv=m.inv,
                ext=list(m.ext), centre=list(m.centre), bbox=list(m.bbox),
                stats=dict(tris=m.ntri, solid=m.solid, openness=m.openness), disagree=m.disagree)


def grid_mask(grid, gdim, gmin, ginv, M, pos, scale, box) -> np.ndarray:
    """Grid-backed ``Body.makeTest`` evaluated over a clamped lattice box."""
    grid = np.ascontiguousarray(grid, dtype=np.uint8)
    gdim = np.ascontiguousarray(gdim, dtype=np.int32)
    gmin = np.ascontiguousarray(gmin, dtype=np.float64)
    M = np.ascontiguousarray(M, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.int32)
    n = (box[3] - box[0] + 1) * (box[4] - box[1] + 1) * (box[5] - box[2] + 1)
    mask = np.zeros(max(int(n), 0), dtype=np.uint8)
    lib().wto_grid_mask(_ptr(grid), _ptr(gdim), _ptr(gmin), float(ginv), _ptr(M), _ptr(pos),
                        float(scale), _ptr(box), _ptr(mask))
    return mask


def slice_view(fields: dict, solid: np.ndarray, dims, axis: str, k: int, field: str, u0: float):
    """The app's slice values before colour mapping (js/main.js:925-965) over
    whole fields (arrays of n cells, x fastest): axis 'xy' (plane z = k) or
    'xz' (plane y = k); field 'speed' |u|/u0, 'pressure' C_p of the density
    averaged with its in-plane neighbours, 'vorticity' the in-plane curl
    component by central differences; a missing neighbour at the plane's
    edge is the cell itself. Binary64 from the stored values, returned as
    float32 [h, nx]; solid cells read 0 (the app paints them instead). The
    speed uses sqrt(x^2 + y^2 + z^2) where the app calls Math.hypot (equal
    to within an ulp of binary64)."""
    nx, ny, nz = dims
    sh = (nz, ny, nx)
    pl = (lambda a: np.asarray(a).reshape(sh)[k, :, :]) if axis == "xy" else (lambda a: np.asarray(a).reshape(sh)[:, k, :])
    ux, uy, uz, rho = (pl(fields[n]).astype(np.float64) for n in ("ux", "uy", "uz", "rho"))
    sol = pl(solid) != 0
    h = ux.shape[0]
    a_p = np.minimum(np.arange(nx) + 1, nx - 1)
    a_m = np.maximum(np.arange(nx) - 1, 0)
    b_p = np.minimum(np.arange(h) + 1, h - 1)
    b_m = np.maximum(np.arange(h) - 1, 0)
    if field == "speed":
        v = np.sqrt(ux * ux + uy * uy + uz * uz) / u0
    elif field == "pressure":
        r = (rho * 2 + rho[:, a_p] + rho[:, a_m] + rho[b_p, :] + rho[b_m, :]) / 6
        v = ((r - 1) / 3) / (0.5 * u0 * u0)
    elif axis == "xy":
        v = (uy[:, a_p] - uy[:, a_m]) * 0.5 - (ux[b_p, :] - ux[b_m, :]) * 0.5
    else:
        v = (ux[b_p, :] - ux[b_m, :]) * 0.5 - (uz[:, a_p] - uz[:, a_m]) * 0.5
    out = v.astype(np.float32)
    out[sol] = 0
    return out

