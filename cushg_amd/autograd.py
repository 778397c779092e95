"""torch.autograd.Function over the C ABI's shg_adjoint.

``eta(waist, focus, temperature, poling, params)`` returns the SH
conversion efficiency of a run as a 0-dim float64 tensor whose backward
pass applies the gradient computed by the library's hand-written adjoint of
the discrete SSFM/RK4 march (checkpointed recomputation; with THERMAL, at
the converged thermo-optic steady state; no reference counterpart -- the
reference computes no gradients).  `temperature` moves T_C and T_peltier_C
together (the reference's argv[3]).  The forward
and the adjoint run together in the forward call, so ``backward`` only
scales the stored gradient.  PyTorch is plumbing here: the inputs are read
into FP64 host scalars / a host poling array and passed through the C ABI.
"""
from __future__ import annotations

import numpy as np
import torch

from . import shg
from .abi import ShgParams


class EtaFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, waist, focus, temperature, poling, params: ShgParams, handle: shg.Handle,
                checkpoint_every: int):
        p = params.copy()
        p.waist_um = float(waist.detach().cpu())
        p.focal_um = float(focus.detach().cpu())
        p.T_C = p.T_peltier_C = float(temperature.detach().cpu())
        pol = poling.detach().to("cpu", torch.float64).contiguous().numpy()
        if pol.shape != (p.nz,):
            raise ValueError(f"poling must have shape ({p.nz},), got {tuple(pol.shape)}")
        p.set_poling(np.ascontiguousarray(pol))
        r, g = handle.adjoint(p, checkpoint_every)
        grad = torch.from_numpy(g)
        ctx.save_for_backward(grad)
        ctx.meta = [(t.dtype, t.device) for t in (waist, focus, temperature, poling)]
        return torch.tensor(r.eta, dtype=torch.float64, device=waist.device)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        go = float(gout.detach().cpu())
        parts = [grad[0], grad[1], grad[2], grad[3:]]
        outs = [(go * v).to(dtype=dt, device=dev) for v, (dt, dev) in zip(parts, ctx.meta)]
        return outs[0], outs[1], outs[2], outs[3], None, None, None


def eta(waist, focus, temperature, poling, params: ShgParams, handle: shg.Handle | None = None,
        checkpoint_every: int = 0):
    """Differentiable eta(waist [um], focus [um], T [C], poling [rad, nz]) of
    the run described by `params` (its other fields are kept)."""
    h = handle if handle is not None else shg.Handle(0)
    return EtaFunction.apply(waist, focus, temperature, poling, params, h, checkpoint_every)
