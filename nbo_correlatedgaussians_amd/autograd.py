"""The ECG energy as a differentiable torch function.

forward(x) -> E.  When x needs a gradient, the forward runs the gradient pair
kernel of outprod (the matel loop and the gradel pair terms of make_grad,
python/outprod.py:49-68, python/make_grad.py:34-55) and the on-device quotient
E = C^T H C / C^T S C, and keeps the per-function pair partials on the engine
as its tape; otherwise it runs the element-only pair kernel.
backward(gE) is the hand-written adjoint contraction of make_grad
(python/make_grad.py:57-90): GRADA from the tape's partials and GRADC from
H C and S C, scaled by gE.  If another evaluation ran on the engine between
the two (the tape is gone), backward recomputes E, H C, S C for its own x
and runs the adjoint pair kernel, which accumulates
sum_kl C_k C_l (dH_kl - E dS_kl) per function: the vector-Jacobian product
with H-bar = C C^T / C^T S C and S-bar = -E C C^T / C^T S C.
"""

from __future__ import annotations

import torch

from .engine import Engine


class ECGEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, engine: Engine):
        if x.dtype != torch.float64:
            raise TypeError("the ECG energy is FP64 throughout: x must be float64")
        ctx.tape = 0
        if ctx.needs_input_grad[0]:
            f, ctx.tape = engine.energy_fwd_t(x)
        else:
            f = engine.energy_t(x)
        ctx.engine = engine
        ctx.save_for_backward(x)
        return f.reshape(())

    @staticmethod
    def backward(ctx, gE):
        (x,) = ctx.saved_tensors
        eng = ctx.engine
        gb = gE.reshape(1).to(torch.float64)
        gx = eng.energy_bwd_t(x, ctx.tape, gb) if ctx.tape else None
        if gx is None:
            eng.energy_t(x)
            gx = eng.vjp_t(x, gb)
        return gx.to(x.device), None


def ecg_energy(x: torch.Tensor, engine: Engine) -> torch.Tensor:
    """Differentiable E(x) for x = [iv | C] (float64, on the engine's GPU)."""
    return ECGEnergy.apply(x, engine)
