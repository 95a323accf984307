"""Test-only torch.autograd.Function over the engine's adjoint C entry points
(sph_adjoint_forward / sph_adjoint_backward): the loss of an FP64 engine run
as a differentiable function of the initial positions, velocities and
internal energies.  CPU tensors in and out; the engine does the work on the
GPU."""
from __future__ import annotations

import numpy as np
import torch


class SphRunLoss(torch.autograd.Function):
    """loss = SphRunLoss.apply(x0, v0, u0, engine, steps, ckpt_every)

    x0, v0: [n, 3] float64, u0: [n] float64.  The engine (FP64 form) keeps its
    masses, mean molecular weights and parameters; its positions, velocities
    and u are replaced by the arguments before the run."""

    @staticmethod
    def forward(ctx, x0, v0, u0, engine, steps, ckpt_every):
        n = engine.n
        pos, vel, hyd = engine.get("pos"), engine.get("vel"), engine.get("hydro")
        pos[:n, :3] = x0.detach().numpy()
        vel[:n, :3] = v0.detach().numpy()
        hyd[:n, 2] = u0.detach().numpy()
        engine.set("pos", pos)
        engine.set("vel", vel)
        engine.set("hydro", hyd)
        loss = engine.adjoint_forward(int(steps), int(ckpt_every))
        ctx.engine = engine
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gx, gv, gu = ctx.engine.adjoint_backward()
        s = float(g)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a) * s)
        return t(gx), t(gv), t(gu), None, None, None
