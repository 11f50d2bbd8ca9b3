"""Deterministic synthetic triangle soups for the benchmark and test bodies.

Every operation here is +, -, *, /, sqrt or floor on IEEE doubles, or 32-bit
integer hashing, followed by one rounding to float32. Those are correctly
rounded on every platform, so the same soup comes out byte for byte wherever
it is generated; the JavaScript generator used to write the golden fixtures
(tests/golden/synth.js) performs the identical operation sequence.

Soups are float32 arrays of 9 values per triangle (three xyz vertices), the
input format of ``Voxel.build`` (js/voxel.js:44-48).
"""
from __future__ import annotations

import hashlib

import numpy as np

M32 = 0xFFFFFFFF
_GOLDEN = 0x9E3779B9


def _mix32(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint32)
    z = (z ^ (z >> np.uint32(16))) * np.uint32(0x85EBCA6B)
    z = (z ^ (z >> np.uint32(13))) * np.uint32(0xC2B2AE35)
    return z ^ (z >> np.uint32(16))


class Rng:
    """Counter-based 32-bit generator: state += golden ratio, murmur3 finaliser.

    ``take(k)`` returns the next k uniforms in [0,1) in draw order.
    """

    def __init__(self, seed: int):
        self.count = 0
        self.seed = seed & M32

    def take(self, k: int) -> np.ndarray:
        idx = np.arange(self.count + 1, self.count + 1 + k, dtype=np.uint64)
        s = (np.uint64(self.seed) + idx * np.uint64(_GOLDEN)) & np.uint64(M32)
        self.count += k
        return _mix32(s).astype(np.float64) / 4294967296.0

    def gauss(self, k: int) -> np.ndarray:
        """k Irwin-Hall(12) - 6 deviates (zero mean, unit variance)."""
        u = self.take(12 * k).reshape(k, 12)
        s = np.zeros(k)
        for c in range(12):
            s = s + u[:, c]
        return s - 6.0


def _hash3(seed: int, i: np.ndarray, j: np.ndarray, k: np.ndarray) -> np.ndarray:
    def u32(a):
        return (a.astype(np.int64) & M32).astype(np.uint32)
    h = np.full(i.shape, seed & M32, dtype=np.uint32)
    h = (h ^ u32(i)) * np.uint32(0x27D4EB2D)
    h = h ^ (h >> np.uint32(15))
    h = (h ^ u32(j)) * np.uint32(0x165667B1)
    h = h ^ (h >> np.uint32(15))
    h = (h ^ u32(k)) * np.uint32(0x9E3779B1)
    return h ^ (h >> np.uint32(16))


def _lerp(a, b, t):
    return a + (b - a) * t


def vnoise(seed: int, x, y, z):
    """Value noise in [-1,1], smoothstep-weighted trilinear blend of hashed corners."""
    ix, iy, iz = np.floor(x), np.floor(y), np.floor(z)
    fx, fy, fz = x - ix, y - iy, z - iz
    sx = fx * fx * (3 - 2 * fx)
    sy = fy * fy * (3 - 2 * fy)
    sz = fz * fz * (3 - 2 * fz)
    ix, iy, iz = ix.astype(np.int64), iy.astype(np.int64), iz.astype(np.int64)

    def v(a, b, c):
        return _hash3(seed, ix + a, iy + b, iz + c).astype(np.float64) / 4294967296.0 * 2 - 1
    x00 = _lerp(v(0, 0, 0), v(1, 0, 0), sx)
    x10 = _lerp(v(0, 1, 0), v(1, 1, 0), sx)
    x01 = _lerp(v(0, 0, 1), v(1, 0, 1), sx)
    x11 = _lerp(v(0, 1, 1), v(1, 1, 1), sx)
    return _lerp(_lerp(x00, x10, sy), _lerp(x01, x11, sy), sz)


def fbm(seed: int, x, y, z):
    """Four octaves of value noise, base frequency 2, amplitude halving; in [-1,1]."""
    s = np.zeros(np.shape(x))
    a, f, norm = 1.0, 2.0, 0.0
    for o in range(4):
        s = s + a * vnoise(seed + o, x * f, y * f, z * f)
        norm += a
        a *= 0.5
        f *= 2
    return s / norm


_ICO_FACES = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
    [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
    [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)


def _normalise(x, y, z):
    s = np.sqrt(x * x + y * y + z * z)
    return x / s, y / s, z / s


def _first_occurrence(keys: np.ndarray):
    """Unique keys numbered in order of first appearance; returns (ids, first_idx)."""
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return rank[inv], first[order]


def icosphere(level: int):
    """Unit icosphere: (V[nv,3] float64, F[nf,3] int64), creation-ordered vertices."""
    t = (1 + np.sqrt(5.0)) / 2
    base = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                     [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]])
    V = np.stack(_normalise(base[:, 0], base[:, 1], base[:, 2]), axis=1)
    F = _ICO_FACES.copy()
    for _ in range(level):
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        e0 = np.stack([a, b, c], axis=1).reshape(-1)       # (a,b), (b,c), (c,a) per face
        e1 = np.stack([b, c, a], axis=1).reshape(-1)
        lo, hi = np.minimum(e0, e1), np.maximum(e0, e1)
        ids, first = _first_occurrence(lo * 2097152 + hi)
        nv = V.shape[0]
        pa, pb = e0[first], e1[first]
        mx = (V[pa, 0] + V[pb, 0]) * 0.5
        my = (V[pa, 1] + V[pb, 1]) * 0.5
        mz = (V[pa, 2] + V[pb, 2]) * 0.5
        V = np.concatenate([V, np.stack(_normalise(mx, my, mz), axis=1)], axis=0)
        mid = (ids + nv).reshape(-1, 3)
        ab, bc, ca = mid[:, 0], mid[:, 1], mid[:, 2]
        G = np.stack([a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca], axis=1)
        F = G.reshape(-1, 3)
    return V, F


def box_surface(nxq: int, nyq: int, nzq: int):
    """Box [0,n]^3 surface of unit quads (two triangles each), shared vertices."""
    n = [nxq, nyq, nzq]
    keys, coords = [], []
    faces = [(0, 0, 2, 1), (0, 1, 1, 2), (1, 0, 0, 2), (1, 1, 2, 0), (2, 0, 1, 0), (2, 1, 0, 1)]
    for ax, side, ua, va in faces:
        A, B = np.meshgrid(np.arange(n[ua]), np.arange(n[va]), indexing="ij")
        A, B = A.reshape(-1), B.reshape(-1)
        for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
            p = [None, None, None]
            p[ax] = np.full(A.shape, n[ax] if side else 0)
            p[ua] = A + da
            p[va] = B + db
            coords.append(np.stack(p, axis=1))
    # interleave corners quad by quad: quad q of face f -> its four corners in call order
    per_face = []
    off = 0
    for f in range(6):
        c = np.stack(coords[4 * f:4 * f + 4], axis=1)       # [nq, 4, 3]
        per_face.append(c)
        off += c.shape[0]
    corners = np.concatenate(per_face, axis=0)               # [nquads, 4, 3]
    flat = corners.reshape(-1, 3).astype(np.int64)
    keys = (flat[:, 0] * 4096 + flat[:, 1]) * 4096 + flat[:, 2]
    ids, first = _first_occurrence(keys)
    V = flat[first].astype(np.float64)
    q = ids.reshape(-1, 4)
    F = np.stack([q[:, 0], q[:, 1], q[:, 2], q[:, 0], q[:, 2], q[:, 3]], axis=1).reshape(-1, 3)
    return V, F


def _soup(V, F, keep=None):
    if keep is not None:
        F = F[keep]
    return V[F].reshape(-1).astype(np.float32)


def ellipsoid_box(level=5, ex=40, ey=14, ez=12, bx=30, by=16, bz=10, bc=(22, -6, 0),
                  sigma=0.2, seed=2, scale=1) -> np.ndarray:
    """Ellipsoid (semi-axes ex,ey,ez cells) unioned with a box (bx,by,bz, centre bc).

    Triangles of either part whose centroid lies inside the other part are
    dropped; every vertex is jittered radially by sigma*N(0,1) cells,
    ellipsoid vertices first, then box vertices, in creation order.
    """
    rng = Rng(seed)
    sc = scale
    U, FE = icosphere(level)
    eX, eY, eZ = ex * sc, ey * sc, ez * sc
    d = sigma * sc * rng.gauss(U.shape[0])
    EV = np.stack([U[:, 0] * eX + U[:, 0] * d, U[:, 1] * eY + U[:, 1] * d,
                   U[:, 2] * eZ + U[:, 2] * d], axis=1)
    bX, bY, bZ = bx * sc, by * sc, bz * sc
    BV0, FB = box_surface(bx * sc, by * sc, bz * sc)
    cx, cy, cz = bc[0] * sc, bc[1] * sc, bc[2] * sc
    px, py, pz = BV0[:, 0] - bX * 0.5, BV0[:, 1] - bY * 0.5, BV0[:, 2] - bZ * 0.5
    ux, uy, uz = _normalise(px, py, pz)
    d = sigma * sc * rng.gauss(BV0.shape[0])
    BV = np.stack([cx + px + ux * d, cy + py + uy * d, cz + pz + uz * d], axis=1)

    def centroid(V, F):
        return [(V[F[:, 0], k] + V[F[:, 1], k] + V[F[:, 2], k]) / 3 for k in range(3)]
    x, y, z = centroid(EV, FE)
    in_box = (np.abs(x - cx) < bX * 0.5) & (np.abs(y - cy) < bY * 0.5) & (np.abs(z - cz) < bZ * 0.5)
    x, y, z = centroid(BV, FB)
    a, b, c = x / eX, y / eY, z / eZ
    in_ell = a * a + b * b + c * c < 1
    return np.concatenate([_soup(EV, FE, ~in_box), _soup(BV, FB, ~in_ell)])


def noisy_sphere(level=7, R=56.0, amp=0.15, holes=0.005, seed=5) -> np.ndarray:
    """Icosphere of radius R cells, radius modulated by amp*R*fbm(direction),
    each triangle deleted with probability ``holes`` (a leaky mesh)."""
    U, F = icosphere(level)
    r = R + amp * R * fbm(seed, U[:, 0], U[:, 1], U[:, 2])
    V = U * r[:, None]
    keep = ~(Rng(seed + 1000).take(F.shape[0]) < holes)
    return _soup(V, F, keep)


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
