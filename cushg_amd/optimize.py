"""Gradient-based eta maximisation over (waist, focus, oven temperature).

No reference counterpart: the reference sweeps its parameters with a bash
loop of one process per point (src/cuSHG.sh:58-94) and computes no
gradients.  Here a batch of parameter triples is evaluated with eta and its
gradient (shg_adjoint, THERMAL runs at their converged thermo-optic steady
state) and every triple takes one step uphill:

    x_j <- x_j + r_j g_j s_j / ||g s||_2,   s = (r_waist, r_focus, r_T),

a step of length 1 in coordinates scaled by the trust radii r (so no
parameter moves by more than its radius), accepted when the forward run at
the new point raises eta and halved (at most `backtrack` times) otherwise.
"""
from __future__ import annotations

import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import abi, shg
from .abi import ShgParams

#: default trust radii: waist [um], focus [um], oven temperature [K]
RADII = (2.0, 500.0, 0.05)


@dataclass
class StepResult:
    """One ascent step of one parameter triple."""
    eta0: float                  # eta at the starting point
    grad: np.ndarray             # d eta / d(waist, focus, T) there
    params: ShgParams            # the accepted point (the start if no step was accepted)
    eta1: float                  # eta at the accepted point
    halvings: int                # step halvings before acceptance
    accepted: bool


@dataclass
class BatchStep:
    steps: list = field(default_factory=list)
    wall_s: float = 0.0          # the whole step: gradients, trial forwards
    grad_wall_s: float = 0.0     # the forward + adjoint evaluations


_FP32_FIELD_FORMS = (abi.PREC_FP32_MIXED, abi.PREC_FP32_STORAGE, abi.PREC_FP32_SPLIT)


def _moved(p: ShgParams, d) -> ShgParams:
    """p moved by d, in the arithmetic of its gradient (shg_adjoint
    differentiates every FP32-field form as FP32_STORAGE), so that the trial
    eta and the starting eta differ by the step, not by the form."""
    q = p.copy()
    if q.precision in _FP32_FIELD_FORMS:
        q.precision = abi.PREC_FP32_STORAGE
    q.waist_um = p.waist_um + d[0]
    q.focal_um = p.focal_um + d[1]
    q.T_C = q.T_peltier_C = p.T_C + d[2]
    return q


def gradients(hs, ps, checkpoint_every: int = 0):
    """(eta, d eta / d(waist, focus, T)) of every triple in `ps` by
    shg_adjoint.  With several handles (on one device, each used by one host
    thread) the instances are taken in turn by one thread per handle, so that
    the single-instance adjoints, which do not fill the GPU alone, run side
    by side on the handles' streams."""
    out = [None] * len(ps)
    nxt = iter(range(len(ps)))
    lock = threading.Lock()

    def work(h):
        while True:
            with lock:
                i = next(nxt, None)
            if i is None:
                return
            res, g = h.adjoint(ps[i], checkpoint_every)
            out[i] = (res.eta, np.array(g[:3], dtype=np.float64))

    if len(hs) == 1:
        work(hs[0])
    else:
        with ThreadPoolExecutor(len(hs)) as ex:
            for f in [ex.submit(work, h) for h in hs]:
                f.result()
    return out


def ascent_step(h: shg.Handle, ps, radii=RADII, backtrack: int = 3, checkpoint_every: int = 0,
                extra_handles=()) -> BatchStep:
    """One eta-maximisation step for every triple in `ps` (ShgParams; T_C
    and T_peltier_C move together).  Gradients by shg_adjoint, one instance
    at a time per handle (`h` and `extra_handles`, see gradients()); the
    trial points of all instances run as one batch on `h`."""
    t0 = time.perf_counter()
    r = np.asarray(radii, dtype=np.float64)
    out = BatchStep()
    evals = gradients([h, *extra_handles], ps, checkpoint_every)
    out.grad_wall_s = time.perf_counter() - t0
    scale = [1.0] * len(ps)
    pending = list(range(len(ps)))
    best = [None] * len(ps)
    halvings = [0] * len(ps)
    for _ in range(backtrack + 1):
        if not pending:
            break
        trial = []
        for i in pending:
            eta0, g = evals[i]
            gs = g * r
            n = float(np.linalg.norm(gs))
            d = (r * gs / n) * scale[i] if n > 0 and math.isfinite(n) else np.zeros(3)
            trial.append(_moved(ps[i], d))
        rc, rs, _ = h.run_batch(trial)
        if rc != abi.SHG_OK:
            raise shg.ShgError(rc, shg.last_error())
        still = []
        for i, q, rr in zip(pending, trial, rs):
            if rr.eta > evals[i][0]:
                q.precision = ps[i].precision
                best[i] = (q, rr.eta)
            else:
                scale[i] *= 0.5
                halvings[i] += 1
                still.append(i)
        pending = still
    for i, p in enumerate(ps):
        eta0, g = evals[i]
        if best[i] is None:
            out.steps.append(StepResult(eta0, g, p.copy(), eta0, halvings[i], False))
        else:
            out.steps.append(StepResult(eta0, g, best[i][0], best[i][1], halvings[i], True))
    out.wall_s = time.perf_counter() - t0
    return out
