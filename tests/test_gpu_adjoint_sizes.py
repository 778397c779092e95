"""Adjoint gradients at every transform length the adjoint passes tile
differently: 192 (two-column tiles, no XCD regrouping), 384 (two-column
tiles, regrouped), 512 (four-column tiles, regrouped) -- without THERMAL,
and 512 with THERMAL -- against central finite differences of the GPU's
FP64 fused forward (the oracle would take minutes at these sizes).  The
512 THERMAL gradient is the same with one and with two points per thread
in the adjoint row pass and with two- and four-column adjoint tiles (equal
to all ten printed digits when the kernels changed).
Steps and tolerances as in test_gpu_adjoint.py: waist 1e-3 um (1e-2 um
with THERMAL), focus 0.5 um, T 2e-4 K, poling 1e-4 rad; FP64 adjoint within 1e-6 of the largest
component of each group, the FP32-storage adjoint within 2e-4 (optics) or
2e-5 (THERMAL, per component as in test_gpu_adjoint_thermal.py).
"""
import numpy as np
import pytest

from cushg_amd import abi, configs

pytestmark = pytest.mark.gpu

STEPS = {"waist_um": 1e-3, "focal_um": 0.5, "T_C": 2e-4}


def set_param(p, name, v):
    setattr(p, name, v)
    if name == "T_C":
        p.T_peltier_C = v


def fd(handle, p, name, h):
    qp, qm = p.copy(), p.copy()
    set_param(qp, name, getattr(p, name) + h)
    set_param(qm, name, getattr(p, name) - h)
    return (handle.run(qp, fields=False)[0].eta - handle.run(qm, fields=False)[0].eta) / (2 * h)


def fd_poling(handle, p, s, h=1e-4):
    ph = np.array([p.poling_phase[i] for i in range(p.nz)])
    qp, qm = p.copy(), p.copy()
    a = ph.copy()
    a[s] += h
    qp.set_poling(a)
    b = ph.copy()
    b[s] -= h
    qm.set_poling(b)
    return (handle.run(qp, fields=False)[0].eta - handle.run(qm, fields=False)[0].eta) / (2 * h)


@pytest.mark.parametrize("n", [192, 384, 512])
def test_adjoint_sizes_vs_gpu_fp64_finite_differences(handle, n):
    p = configs.adjoint256(nz=24, n=n)
    p.lx_um = p.ly_um = 2000.0 * n / 256
    p.power_W = 20.0
    q = p.copy()
    q.precision = abi.PREC_FP64_FUSED
    f = [fd(handle, q, name, h) for name, h in STEPS.items()]
    fp = [fd_poling(handle, q, s) for s in (0, 11, 22)]
    for prec, tol in ((abi.PREC_FP64_FUSED, 1e-6), (abi.PREC_FP32_STORAGE, 2e-4)):
        p.precision = prec
        print(f"case: adjoint {n}x{n}x{p.nz}, precision {prec}", flush=True)
        r, g = handle.adjoint(p)
        assert r.status == abi.SHG_OK
        scale = max(abs(x) for x in f)
        for i, name in enumerate(STEPS):
            print(f"  d eta/d {name}: adjoint {g[i]:.10e} fd {f[i]:.10e}", flush=True)
            assert abs(g[i] - f[i]) <= tol * scale
        gmax = np.abs(g[3:]).max()
        for s, x in zip((0, 11, 22), fp):
            assert abs(g[3 + s] - x) <= tol * gmax


def test_thermal_adjoint_512_vs_gpu_fp64_finite_differences(handle):
    """512 x 512 THERMAL (MgO:PPLN, 96 slices, 6 heat slabs).  The finite
    differences here carry the forward's outer-loop truncation: each of the
    two runs stops on the reference's one-sweep test after 26-27 slowly
    contracting thermo-optic passes, which leaves ~1e-6 of the gradient in
    the difference (the adjoint is the derivative at the exact fixed point;
    the 64 x 64 test against the tightly converged FP64 oracle holds 1e-6).
    So: FP64 adjoint vs differences within 2e-5 per component, and the
    FP32-storage adjoint vs the FP64 one within 5e-5 per component (the
    focus derivative here is a near-cancelling sum, 1.6e-5 of the
    temperature one, so the FP32 fields' rounding weighs more in it)."""
    p = configs.ppln512_thermal()
    p.nz = 96
    p.nz_heat = 6
    p.heat_solver = abi.HEAT_DIRECT
    p.jacobi_tol = 1e-10
    p.heat_tol_K = 3e-11  # the FP64 defect correction's floor at this size is ~1.5e-11 K
    p.max_outer = 40
    p.set_poling(np.random.default_rng(5).normal(0.0, 0.05, p.nz))
    q = p.copy()
    q.precision = abi.PREC_FP64_FUSED
    steps = {"waist_um": 1e-2, "focal_um": 0.5, "T_C": 2e-4}
    f = [fd(handle, q, name, h) for name, h in steps.items()]
    grads = {}
    for prec in (abi.PREC_FP64_FUSED, abi.PREC_FP32_STORAGE):
        p.precision = prec
        print(f"case: THERMAL adjoint 512x512x{p.nz} heat {p.nz_heat}, precision {prec}", flush=True)
        r, g = handle.adjoint(p)
        assert r.status == abi.SHG_OK and r.peak_dT_K > 0.5
        grads[prec] = g
        for i, name in enumerate(steps):
            print(f"  d eta/d {name}: adjoint {g[i]:.10e} fd {f[i]:.10e} rel {abs(g[i] - f[i]) / abs(f[i]):.2e}",
                  flush=True)
    g64, g32 = grads[abi.PREC_FP64_FUSED], grads[abi.PREC_FP32_STORAGE]
    for i in range(3):
        assert abs(g64[i] - f[i]) <= 2e-5 * abs(f[i])
        assert abs(g32[i] - g64[i]) <= 5e-5 * abs(g64[i])
