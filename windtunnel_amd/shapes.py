"""Host-side mesh placement of the reference's Shapes (js/shapes.js).

``orient`` turns an uploaded triangle soup into a body of unit length, the
step before ``Voxel.build`` when the app registers a model
(js/shapes.js:212-247; addModel, js/shapes.js:250). It runs once per upload on a
few hundred thousand vertices, so it stays on the host; the heavy stages
after it (voxelization, the grid lookup, the solver) run on the GPU.
"""
from __future__ import annotations

import numpy as np


def orient(P: np.ndarray, forward: int, up: int, flip: bool, ground: bool) -> float:
    """Rotate P (float32, 9 values per triangle, modified in place) so the
    nose points along -x and up along +y, then centre it and scale its
    length to 1; returns the scale k (js/shapes.js:212-247).

    Arithmetic follows the reference: binary64 products and sums in its
    order, each result stored back into the float32 array; the bounds are
    taken from the binary64 rotated values before they are stored.
    """
    if P.dtype != np.float32:
        raise TypeError("orient works on a float32 triangle soup")
    if forward == up:
        forward = (up + 1) % 3
    f = [0.0, 0.0, 0.0]
    u = [0.0, 0.0, 0.0]
    f[forward] = -1.0 if flip else 1.0
    u[up] = 1.0
    s = [f[1] * u[2] - f[2] * u[1], f[2] * u[0] - f[0] * u[2], f[0] * u[1] - f[1] * u[0]]
    R = np.array([-f[0], -f[1], -f[2], u[0], u[1], u[2], -s[0], -s[1], -s[2]], dtype=np.float64)
    V = P.reshape(-1, 3)
    a, b, c = (V[:, 0].astype(np.float64), V[:, 1].astype(np.float64), V[:, 2].astype(np.float64))
    x = R[0] * a + R[1] * b + R[2] * c
    y = R[3] * a + R[4] * b + R[5] * c
    z = R[6] * a + R[7] * b + R[8] * c
    V[:, 0], V[:, 1], V[:, 2] = x, y, z
    x0, x1 = (x.min(), x.max()) if x.size else (np.inf, -np.inf)
    y0, y1 = (y.min(), y.max()) if y.size else (np.inf, -np.inf)
    z0, z1 = (z.min(), z.max()) if z.size else (np.inf, -np.inf)
    k = 1.0 / max(1e-9, float(x1 - x0))
    cx, cz = (x0 + x1) * 0.5, (z0 + z1) * 0.5
    cy = y0 if ground else (y0 + y1) * 0.5
    V[:, 0] = (V[:, 0].astype(np.float64) - cx) * k
    V[:, 1] = (V[:, 1].astype(np.float64) - cy) * k
    V[:, 2] = (V[:, 2].astype(np.float64) - cz) * k
    return float(k)
