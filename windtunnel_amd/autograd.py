"""Gray-weight force response as a torch.autograd.Function.

    forces = gray_forces(w, lattice, nsteps)        # [nsteps, 3] float64
    J = (forces[-500:, 0] / q).mean()                # e.g. mean C_d
    J.backward()                                     # w.grad = dJ/dw per cell

forward sets the lattice's gray bounce-back weights to w, runs nsteps from
the lattice's current state and records sqrt(N) checkpoints on the device
(wt_adjoint_forward); backward hands the cotangent of the forces to the
discrete adjoint (wt_adjoint_backward), which returns dJ/dw for every cell
(0 outside the gray band). The lattice's state advances by nsteps, as after
step(nsteps); backward leaves it alone. A lattice holds one recorded
forward pass: a second gray_forces call on it replaces the first, and the
first graph's backward then raises instead of returning the other graph's
gradient. torch is plumbing here: the tensors
are copied to and from host memory, the work runs in libwindtunnel.
"""
from __future__ import annotations

import numpy as np
import torch

from .lbm import LBM


class GrayForces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w: torch.Tensor, lattice: LBM, nsteps: int, interval: int = 0):
        if w.numel() != lattice.n:
            raise ValueError(f"w must hold one weight per cell ({lattice.n}), got {w.numel()}")
        lattice.set_gray(w.detach().to("cpu", torch.float32).numpy().reshape(-1))
        f = lattice.adjoint_forward(int(nsteps), int(interval))
        ctx.lattice = lattice
        ctx.record = lattice.adjoint_record()
        ctx.w_meta = (w.shape, w.dtype, w.device)
        return torch.from_numpy(f)

    @staticmethod
    def backward(ctx, gF: torch.Tensor):
        shape, dtype, device = ctx.w_meta
        if ctx.lattice.adjoint_record() != ctx.record:
            raise RuntimeError("the lattice's recorded forward pass was replaced by a later gray_forces call on "
                               "the same lattice; backward of this graph would use the other graph's states")
        g, _ = ctx.lattice.adjoint_backward(np.ascontiguousarray(gF.detach().to("cpu", torch.float64).numpy()))
        return torch.from_numpy(g).reshape(shape).to(device=device, dtype=dtype), None, None, None


def gray_forces(w: torch.Tensor, lattice: LBM, nsteps: int, interval: int = 0) -> torch.Tensor:
    """Forces [nsteps, 3] of nsteps steps with gray weights w; differentiable in w."""
    return GrayForces.apply(w, lattice, nsteps, interval)
