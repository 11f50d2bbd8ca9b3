"""The 640x320x320 binary16 lattice of the full-size adjoint workload, kept
short: the seeded leaky icosphere (level 7, R = 56 cells, 0.5 % of the
triangles deleted) voxelized on the GPU, TRT (Lambda 3/16) + LES (Cs 0.14)
at tau 0.505, u0 0.08, a gray skin of weight 0.5 on the first fluid layer.

* 300 steps stay finite (forces and the final fields);
* the adjoint's recorded forward pass gives the same forces, bit for bit,
  as a plain run of the same steps from the same reset state;
* its checkpoints and replay store fit the device, the gradient is finite,
  nonzero on the gray skin and zero elsewhere, and the reverse sweep leaves
  the lattice's state alone (the next step equals the plain run's).
"""
import hashlib

import numpy as np
import pytest

from windtunnel_amd.lbm import LBM, Voxel, device_memory
from windtunnel_amd.synth import noisy_sphere

pytestmark = pytest.mark.gpu

DIMS = (640, 320, 320)
N = 300


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_full_size_binary16_adjoint_run():
    print(f"[scale] {DIMS} binary16, {N} steps", flush=True)
    vox = Voxel.build(noisy_sphere(level=7, R=56, holes=0.005, seed=5), 128)
    L = LBM(*DIMS, storage="f16")
    L.set_params(tau=0.505, les=True, csm=0.14, trt=True, u0=0.08, **{"lambda": 3 / 16})
    L.reset()
    pos = np.array([0.3 * DIMS[0], DIMS[1] / 2, DIMS[2] / 2])
    b = np.array(vox["bbox"])
    lo, hi = pos + b[:3] - 2, pos + b[3:] + 2
    L.voxelize_grid(vox, np.eye(3).reshape(-1), pos, 1.0,
                    dict(x0=lo[0], y0=lo[1], z0=lo[2], x1=hi[0], y1=hi[1], z1=hi[2]))
    assert L.solidCells > 700000
    solid, near = L.solid_near()
    w = np.where((near == 1) & (solid == 0), 0.5, 0.0).astype(np.float32)
    skin = w > 0
    del solid, near
    L.set_gray(w)
    del w

    F_plain = L.step(N + 1)
    assert np.all(np.isfinite(F_plain))
    for k in ("rho", "ux", "uy", "uz"):
        assert np.all(np.isfinite(getattr(L, k))), k
    plain_rho = sha(L.rho)

    L.reset()
    F_adj = L.adjoint_forward(N)
    assert np.array_equal(F_adj, F_plain[:N])
    free, total = device_memory()
    q = 0.5 * 0.08 ** 2 * L.frontalArea
    Fbar = np.zeros((N, 3))
    Fbar[N // 2:, 0] = 1.0 / ((N - N // 2) * q)
    g, ms = L.adjoint_backward(Fbar)
    print(f"[scale] reverse {ms / N:.3f} ms/step, {(total - free) / 2**30:.1f} GiB in use after the forward "
          f"of {total / 2**30:.1f} GiB", flush=True)
    assert L.steps == N
    assert np.all(np.isfinite(g))
    assert np.count_nonzero(g[~skin]) == 0
    assert np.count_nonzero(g[skin]) > 0.9 * np.count_nonzero(skin)
    # the reverse sweep left the lattice's state alone: the next step is
    # the plain run's step N + 1
    assert np.array_equal(L.step(1), F_plain[N:])
    assert sha(L.rho) == plain_rho
    L.adjoint_release()
    L.close()
