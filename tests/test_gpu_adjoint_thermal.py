"""THERMAL adjoint: d eta / d(waist, focus, oven temperature, poling[s]) at
the converged thermo-optic steady state, against central finite
differences of the FP64 oracle.

The forward being differentiated is the reference's outer loop
(src/headers/Solver.h:321-373): a heat solve from the last march's heat
source, the march on that temperature, until a heat solve needs one sweep.
The oven temperature moves T_C and T_peltier together (the reference's
argv[3], src/cuSHG.cu:46-53); in a THERMAL run only T_peltier reaches the
steady state.  The case heats the crystal by about a kelvin (strong
absorption), so the gradient's heat-feedback part (the heat source's
dependence on every parameter, through T back into the phase matching) is
a sizeable part of each component.  The oracle's outer loop is converged
tightly (one-sweep change below 1e-10 K) so that its eta is a smooth
function of the parameters.

Finite-difference steps: waist 1e-3 um, focus 0.5 um, T 2e-4 K, poling
1e-4 rad.  Tolerances relative to the largest component of each group.
"""
import numpy as np
import pytest

from cushg_amd import abi, configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STEPS = {"waist_um": 1e-3, "focal_um": 0.5, "T_C": 2e-4}


def thermal_case(precision, bc=abi.HEAT_BC_OPEN_TOP):
    p = configs.spplt256_thermal()
    p.nx = p.ny = 64
    p.nz = 24
    p.nz_heat = 8
    p.lx_um = p.ly_um = 400.0
    p.lz_um = 10000.0
    p.focal_um = 5000.0
    p.power_W = 20.0
    p.waist_um = 25.0
    p.alpha_p, p.alpha_s, p.beta_s = 2e-6, 2e-5, 5e-5
    p.heat_solver = abi.HEAT_DIRECT
    p.heat_bc = bc
    p.T_C = p.T_peltier_C = 40.0
    p.Lambda_um = abi.qpm_period(p, 40.6)  # phase matched slightly above the oven: eta moves with T
    p.jacobi_tol = 1e-10
    p.max_outer = 40
    p.set_poling(np.random.default_rng(4).normal(0.0, 0.05, p.nz))
    p.precision = precision
    return p


def set_param(p, name, v):
    setattr(p, name, v)
    if name == "T_C":
        p.T_peltier_C = v


def oracle_eta(p):
    r, _ = O.run(p, fields=False)
    return r.eta


def fd_scalar(p, name, h):
    qp, qm = p.copy(), p.copy()
    v = getattr(p, name)
    set_param(qp, name, v + h)
    set_param(qm, name, v - h)
    return (oracle_eta(qp) - oracle_eta(qm)) / (2 * h)


def fd_poling(p, s, h=1e-4):
    ph = np.array([p.poling_phase[i] for i in range(p.nz)])
    qp, qm = p.copy(), p.copy()
    a = ph.copy()
    a[s] += h
    qp.set_poling(a)
    b = ph.copy()
    b[s] -= h
    qm.set_poling(b)
    return (oracle_eta(qp) - oracle_eta(qm)) / (2 * h)


@pytest.mark.parametrize("precision,bc,tol", [
    pytest.param(abi.PREC_FP64_FUSED, abi.HEAT_BC_OPEN_TOP, 1e-6, id="fp64-open"),
    pytest.param(abi.PREC_FP32_STORAGE, abi.HEAT_BC_OPEN_TOP, 2e-5, id="fp32storage-open"),
    pytest.param(abi.PREC_FP64_FUSED, abi.HEAT_BC_OVEN_SURROUNDED, 1e-6, id="fp64-oven"),
])
def test_thermal_adjoint_vs_oracle_finite_differences(handle, precision, bc, tol):
    p = thermal_case(precision, bc)
    print(f"case THERMAL adjoint 64x64x24 heat 8, precision {precision}, bc {bc}", flush=True)
    r, g = handle.adjoint(p)
    q = p.copy()
    q.precision = abi.PREC_FP64_PLAIN
    ro, _ = O.run(q, fields=False)
    print(f"  eta gpu={r.eta:.10e} oracle={ro.eta:.10e} dT={ro.peak_dT_K:.4f} K outer {r.outer_iters}/"
          f"{ro.outer_iters}", flush=True)
    assert ro.peak_dT_K > 0.3
    assert abs(r.eta - ro.eta) / ro.eta < 1e-5
    fd = [fd_scalar(q, n, h) for n, h in STEPS.items()]
    for i, name in enumerate(STEPS):
        err = abs(g[i] - fd[i]) / abs(fd[i])
        print(f"  d eta/d {name}: adjoint {g[i]:.10e} fd {fd[i]:.10e} rel {err:.2e}", flush=True)
        assert err < tol
    sl = [0, 5, 11, p.nz - 2, p.nz - 1]
    fdp = np.array([fd_poling(q, s) for s in sl])
    scale = np.abs(g[3:]).max()
    perr = np.abs(g[3 + np.array(sl)] - fdp).max() / scale
    print(f"  d eta/d poling at {sl}: max |adjoint - fd| / max |adjoint| = {perr:.2e} (scale {scale:.3e})",
          flush=True)
    assert perr < tol


def test_thermal_adjoint_checkpoint_interval_invariance(handle):
    """The gradient does not depend on the checkpoint interval (one
    segment, a few, one slice per segment): every segment is recomputed
    from the same stored states."""
    p = thermal_case(abi.PREC_FP64_FUSED)
    p.nz = 12
    p.set_poling(np.linspace(-0.05, 0.05, p.nz))
    gs = [handle.adjoint(p, ck)[1] for ck in (0, 1, 5, 7, 12)]
    for g in gs[1:]:
        assert np.abs(g - gs[0]).max() <= 1e-12 * np.abs(gs[0]).max()


def test_thermal_adjoint_refusals(handle):
    from cushg_amd import shg
    p = thermal_case(abi.PREC_FP64_FUSED)
    p.heat_solver = abi.HEAT_JACOBI_REFERENCE
    with pytest.raises(shg.ShgError) as e:
        handle.adjoint(p)
    assert e.value.status == abi.SHG_EUNSUPPORTED
