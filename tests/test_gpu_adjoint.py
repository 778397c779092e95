"""Adjoint gradients (shg_adjoint) against central finite differences of the
FP64 oracle and of the GPU's own FP64 forward, the invariance of the
result under the checkpoint interval, and the torch.autograd wrapper.

d eta / d(waist, focus, T, poling[s]) of a non-THERMAL run.  Finite
differences: central, steps chosen so the truncation error (h^2 f'''/6) is
below 1e-8 of the gradient scale for these cases: waist 1e-3 um, focus
0.5 um, T 2e-4 K, poling 1e-4 rad; the FP64 oracle's round-off (~1e-16
eta / h) is far below that.  Tolerances are relative to the largest
component of each group (the poling vector as one group).
"""
import numpy as np
import pytest

from cushg_amd import abi, configs
from oracle import oracle as O

STEPS = {"waist_um": 1e-3, "focal_um": 0.5, "T_C": 2e-4}


def small_case(precision):
    p = configs.adjoint256(nz=30, n=64)
    p.power_W = 20.0
    p.precision = precision
    return p


def set_param(p, name, v):
    setattr(p, name, v)
    if name == "T_C":
        p.T_peltier_C = v


def fd_scalar(run_eta, p, name, h):
    qp, qm = p.copy(), p.copy()
    v = getattr(p, name)
    set_param(qp, name, v + h)
    set_param(qm, name, v - h)
    return (run_eta(qp) - run_eta(qm)) / (2 * h)


def fd_poling(run_eta, p, s, h=1e-4):
    ph = np.array([p.poling_phase[i] for i in range(p.nz)])
    qp, qm = p.copy(), p.copy()
    a = ph.copy()
    a[s] += h
    qp.set_poling(a)
    b = ph.copy()
    b[s] -= h
    qm.set_poling(b)
    return (run_eta(qp) - run_eta(qm)) / (2 * h)


def oracle_eta(p):
    return O.run(p, fields=False)[0].eta


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(abi.PREC_FP64_FUSED, 1e-6), (abi.PREC_FP32_STORAGE, 2e-4)])
def test_adjoint_vs_oracle_finite_differences(handle, precision, tol):
    p = small_case(precision)
    print(f"case adjoint 64x64x30 precision={precision}", flush=True)
    r, g = handle.adjoint(p)
    q = p.copy()
    q.precision = abi.PREC_FP64_PLAIN
    ro, _ = O.run(q, fields=False)
    assert abs(r.eta - ro.eta) / ro.eta < (1e-11 if precision == abi.PREC_FP64_FUSED else 1e-5)
    fd = [fd_scalar(oracle_eta, q, n, h) for n, h in STEPS.items()]
    for i, name in enumerate(STEPS):
        err = abs(g[i] - fd[i]) / abs(fd[i])
        print(f"  d eta/d {name}: adjoint {g[i]:.10e} fd {fd[i]:.10e} rel {err:.2e}")
        assert err < tol
    fdp = np.array([fd_poling(oracle_eta, q, s) for s in range(p.nz)])
    scale = np.abs(fdp).max()
    perr = np.abs(g[3:] - fdp).max() / scale
    print(f"  d eta/d poling: max |adjoint - fd| / max |fd| = {perr:.2e} (scale {scale:.3e})")
    assert perr < tol
    # the last slice does not reach the saved output
    assert g[3 + p.nz - 1] == 0.0 and abs(fdp[-1]) < 1e-9 * scale


@pytest.mark.gpu
def test_adjoint_walkoff_loss_and_strang(handle):
    """Walk-off, 2PA and loss on, Strang scheme: every derivative path of
    dA/dz and the propagator."""
    p = small_case(abi.PREC_FP64_FUSED)
    p.nz = 24
    p.set_poling(np.linspace(-0.1, 0.1, p.nz))
    p.rho_s = 0.01
    p.alpha_p, p.alpha_s, p.beta_s = 1e-6, 2e-5, 5e-4
    p.scheme = abi.SCHEME_STRANG
    p.focal_um = 0.3 * p.lz_um
    r, g = handle.adjoint(p)
    q = p.copy()
    q.precision = abi.PREC_FP64_PLAIN
    fd = [fd_scalar(oracle_eta, q, n, h) for n, h in STEPS.items()]
    for i, name in enumerate(STEPS):
        err = abs(g[i] - fd[i]) / abs(fd[i])
        print(f"  d eta/d {name}: adjoint {g[i]:.10e} fd {fd[i]:.10e} rel {err:.2e}")
        assert err < 1e-6
    for s in (0, 7, p.nz - 2):
        f = fd_poling(oracle_eta, q, s)
        assert abs(g[3 + s] - f) <= 1e-6 * np.abs(g[3:]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("nz", [2, 3, 5])
def test_adjoint_few_slices(handle, nz):
    """Marches of one to four slices: one segment recorded by the forward
    itself, the output boundary and the first slice in the same segment."""
    p = small_case(abi.PREC_FP64_FUSED)
    p.nz = nz
    p.set_poling(np.linspace(0.05, -0.05, nz))
    print(f"case adjoint 64x64x{nz}", flush=True)
    q = p.copy()
    q.precision = abi.PREC_FP64_PLAIN
    for ck in (0, 1, 7):
        r, g = handle.adjoint(p, ck)
        for i, (name, h) in enumerate(STEPS.items()):
            f = fd_scalar(oracle_eta, q, name, h)
            assert abs(g[i] - f) <= 1e-6 * abs(f), (ck, name, g[i], f)
        fdp = np.array([fd_poling(oracle_eta, q, s) for s in range(nz)])
        # absolute floor: the central difference cannot resolve less than
        # eta's FP64 round-off over the step, ~ eta * 1e-16 / 1e-4
        assert np.abs(g[3:] - fdp).max() <= 1e-6 * np.abs(fdp).max() + 1e-10 * r.eta, ck
        assert g[3 + nz - 1] == 0.0


@pytest.mark.gpu
def test_checkpoint_interval_does_not_change_the_gradient(handle):
    p = small_case(abi.PREC_FP64_FUSED)
    p.nz = 40
    p.set_poling(np.zeros(p.nz))
    _, g0 = handle.adjoint(p, 1)
    for ck in (0, 4, 7, 39, 100):
        _, g = handle.adjoint(p, ck)
        assert np.array_equal(g, g0), ck
    p.precision = abi.PREC_FP32_STORAGE
    _, h0 = handle.adjoint(p, 1)
    _, h1 = handle.adjoint(p, 6)
    assert np.array_equal(h0, h1)
    # every FP32-field form differentiates in FP64 arithmetic (FP32 storage)
    for prec in (abi.PREC_FP32_MIXED, abi.PREC_FP32_SPLIT):
        p.precision = prec
        _, h2 = handle.adjoint(p, 6)
        assert np.array_equal(h2, h0), prec


@pytest.mark.gpu
def test_adjoint256_vs_gpu_fp64_finite_differences(handle):
    """The 256x256, 2000-slice adjoint case (checkpoints every 45 slices):
    the FP32-storage adjoint against central differences of the GPU's FP64
    fused forward, and the cost of the gradient relative to one forward."""
    import time
    p = configs.adjoint256()
    p.precision = abi.PREC_FP32_STORAGE
    print("case adjoint 256x256x2000 ck=45", flush=True)
    handle.run(p, fields=False)
    t = time.perf_counter()
    r0, _ = handle.run(p, fields=False)
    tf = time.perf_counter() - t
    handle.adjoint(p, 45)
    t = time.perf_counter()
    r, g = handle.adjoint(p, 45)
    ta = time.perf_counter() - t
    print(f"  forward {tf * 1e3:.1f} ms, forward + adjoint {ta * 1e3:.1f} ms, ratio {ta / tf:.2f}")
    assert abs(r.eta - r0.eta) / r0.eta < 1e-12
    # the gradient (forward, recomputation and reverse sweep) costs at most
    # three forwards (measured about 2.2)
    assert ta / tf <= 3.0
    q = p.copy()
    q.precision = abi.PREC_FP64_FUSED

    def gpu_eta(x):
        return handle.run(x, fields=False)[0].eta

    for i, (name, h) in enumerate(STEPS.items()):
        f = fd_scalar(gpu_eta, q, name, h)
        err = abs(g[i] - f) / abs(f)
        print(f"  d eta/d {name}: adjoint {g[i]:.10e} fd {f:.10e} rel {err:.2e}")
        assert err < 2e-4
    gmax = np.abs(g[3:]).max()
    for s in (0, 500, 1000, 1998):
        f = fd_poling(gpu_eta, q, s)
        print(f"  d eta/d poling[{s}]: adjoint {g[3 + s]:.6e} fd {f:.6e}")
        assert abs(g[3 + s] - f) <= 2e-4 * gmax


@pytest.mark.gpu
def test_torch_autograd_wrapper(handle):
    import torch

    from cushg_amd import autograd
    p = small_case(abi.PREC_FP64_FUSED)
    w = torch.tensor(p.waist_um, dtype=torch.float64, requires_grad=True)
    f = torch.tensor(p.focal_um, dtype=torch.float64, requires_grad=True)
    T = torch.tensor(p.T_C, dtype=torch.float64, requires_grad=True)
    pol = torch.tensor([p.poling_phase[i] for i in range(p.nz)], dtype=torch.float64, requires_grad=True)
    e = autograd.eta(w, f, T, pol, p, handle)
    (3.0 * e).backward()
    r, g = handle.adjoint(p)
    assert float(e.detach()) == r.eta
    assert float(w.grad) == 3.0 * g[0] and float(f.grad) == 3.0 * g[1] and float(T.grad) == 3.0 * g[2]
    assert np.array_equal(pol.grad.numpy(), 3.0 * g[3:])
