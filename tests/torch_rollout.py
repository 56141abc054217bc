"""torch.autograd.Function over the solver's adjoint (test harness only: torch here is
plumbing for the loss; the forward rollout and its reverse run in the HIP library)."""
from __future__ import annotations

import numpy as np
import torch


class Rollout(torch.autograd.Function):
    """final positions x(T) of one instance as a function of log per-edge stretch
    compliance and the initial velocities; backward = xrt_adjoint_backward."""

    @staticmethod
    def forward(ctx, log_c, v0, solver, inst, x0, n_frames, checkpoint_every):
        solver.SetStretchCompliance(inst, np.exp(log_c.detach().cpu().numpy()).astype(np.float32))
        solver.SetState(inst, x0, v0.detach().cpu().numpy().astype(np.float32))
        solver.AdjointBegin(n_frames, checkpoint_every)
        x, _ = solver.GetState(inst)
        ctx.solver, ctx.inst = solver, inst
        ctx.save_for_backward(log_c)
        return torch.from_numpy(x.astype(np.float64))

    @staticmethod
    def backward(ctx, gx):
        s, i = ctx.solver, ctx.inst
        (log_c,) = ctx.saved_tensors
        s.AdjointSeed(i, gx.detach().cpu().numpy().astype(np.float32), None)
        s.AdjointBackward()
        gs, _ = s.GradCompliance(i)
        _, gv = s.GradState(i)
        g_log_c = torch.from_numpy(gs.astype(np.float64)) * torch.exp(log_c.detach())
        return g_log_c, torch.from_numpy(gv.astype(np.float64)), None, None, None, None, None
