"""torch.autograd bridge of the discrete adjoint: C_L(alpha, bumps) as a
differentiable torch function.

Forward records the SSPRK3 run of the NACA0012 case for the given angle
(degrees) and Hicks-Henne bump amplitudes (``dg_sens_set_alpha`` +
``dg_sens_record``); backward runs the checkpointed adjoint sweep
(``dg_sens_gradient``) and scales dC_L/dalpha and dC_L/dbumps by the
incoming cotangent.  The dt sequence is the one recorded by the forward, so
forward and backward differentiate the same discrete map.  Thin by design:
a test and experimentation wrapper over the C ABI, one case per call.
"""
import numpy as np

from .dg import Sensitivity


def lift(sens, alpha, bumps, nsteps=400, cfl=1.0, ckpt_every=20):
    """C_L of the recorded run as a 0-d float64 tensor; differentiable in
    alpha (0-d tensor, degrees) and bumps (1-d tensor, chords)."""
    import torch

    class _Lift(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a, b):
            if not isinstance(sens, Sensitivity):
                raise TypeError("sens must be a Sensitivity")
            sens.set_alpha(float(a.detach().cpu()))
            bb = b.detach().cpu().numpy().astype(np.float64)
            cl = sens.record(bb, nsteps=nsteps, cfl=cfl)
            ctx.dev = b.device
            return torch.tensor(cl, dtype=torch.float64, device=b.device)

        @staticmethod
        def backward(ctx, g):
            _, da, db = sens.gradient(ckpt_every=ckpt_every)
            gs = g.detach().to(torch.float64).cpu()
            ga = gs * torch.tensor(da, dtype=torch.float64)
            gb = gs * torch.from_numpy(np.asarray(db, dtype=np.float64))
            return ga.to(ctx.dev), gb.to(ctx.dev)

    return _Lift.apply(alpha, bumps)
