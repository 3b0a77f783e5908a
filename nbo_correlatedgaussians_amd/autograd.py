"""The ECG energy as a differentiable torch function.

forward(x) -> E runs the element pair kernel (the matel loop,
python/outprod.py:49-60) and the on-device quotient E = C^T H C / C^T S C
(python/outprod.py:65-68).  backward(gE) runs the hand-written adjoint of
gradel/make_grad: with the energy known, one pair kernel accumulates
sum_kl C_k C_l (dH_kl - E dS_kl) per function (python/make_grad.py:37-90),
i.e. the vector-Jacobian product with H-bar = C C^T / C^T S C and
S-bar = -E C C^T / C^T S C, scaled by gE.
"""

from __future__ import annotations

import torch

from .engine import Engine


class ECGEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, engine: Engine):
        if x.dtype != torch.float64:
            raise TypeError("the ECG energy is FP64 throughout: x must be float64")
        f = engine.energy_t(x)
        ctx.engine = engine
        ctx.generation = engine.energy_generation
        ctx.save_for_backward(x)
        return f.reshape(())

    @staticmethod
    def backward(ctx, gE):
        (x,) = ctx.saved_tensors
        eng = ctx.engine
        if eng.energy_generation != ctx.generation:
            # another forward ran on this engine since: restore its state (E,
            # H C, S C) for this x before the adjoint pass
            eng.energy_t(x)
        gx = eng.vjp_t(x, gE.reshape(1).to(torch.float64))
        return gx.to(x.device), None


def ecg_energy(x: torch.Tensor, engine: Engine) -> torch.Tensor:
    """Differentiable E(x) for x = [iv | C] (float64, on the engine's GPU)."""
    return ECGEnergy.apply(x, engine)
