"""The gradient-based eta maximisation step (cushg_amd.optimize.ascent_step)
over (waist, focus, oven temperature) triples of THERMAL runs.

Checks: every accepted step raises eta; with small trust radii the rise
equals the directional derivative g . d to first order (to 0.5% of it); a
step of the full radius along the gradient is never a descent direction.
"""
import numpy as np
import pytest

from cushg_amd import abi, configs, optimize, shg

pytestmark = pytest.mark.gpu


def triples(n=4):
    base = configs.ppln512_thermal()
    base.nx = base.ny = 64
    base.lx_um = base.ly_um = 500.0
    base.nz = 24
    base.nz_heat = 6
    base.precision = abi.PREC_FP32_SPLIT
    r = configs.splitmix64(9)
    out = []
    for _ in range(n):
        p = base.copy()
        p.waist_um = 40.0 + 20.0 * next(r)
        p.focal_um = p.lz_um * (0.4 + 0.2 * next(r))
        p.T_C = p.T_peltier_C = 38.0 + 4.0 * next(r)
        out.append(p)
    return out


def test_ascent_step_raises_eta(handle):
    ps = triples()
    print("case: ascent step, 4 THERMAL PPLN 64x64x24 triples", flush=True)
    st = optimize.ascent_step(handle, ps)
    for s in st.steps:
        print(f"  eta {s.eta0:.8e} -> {s.eta1:.8e} halvings {s.halvings} accepted {s.accepted}", flush=True)
        assert s.accepted and s.eta1 > s.eta0


def test_small_step_matches_directional_derivative(handle):
    """The outer loop is converged tightly here (one-sweep change below
    1e-10 K) so that the trial eta carries no fixed-point truncation."""
    ps = triples(2)
    for p in ps:
        p.jacobi_tol = 1e-10
        p.max_outer = 40
    radii = (0.01, 2.0, 2e-4)
    st = optimize.ascent_step(handle, ps, radii=radii, backtrack=0)
    for p, s in zip(ps, st.steps):
        d = np.array([s.params.waist_um - p.waist_um, s.params.focal_um - p.focal_um, s.params.T_C - p.T_C])
        lin = float(s.grad @ d)
        print(f"  d eta {s.eta1 - s.eta0:.6e} linear {lin:.6e}", flush=True)
        assert lin > 0
        assert abs((s.eta1 - s.eta0) - lin) <= 0.005 * lin
        assert np.all(np.abs(d) <= np.array(radii) * (1 + 1e-12))


def test_gradients_on_several_handles_equal_one_handle(handle):
    """Gradients taken side by side on three handles (one host thread each)
    equal those taken one after another on one handle."""
    ps = triples(5)
    print("case: gradients of 5 THERMAL PPLN 64x64x24 triples, 1 handle vs 3", flush=True)
    one = optimize.gradients([handle], ps)
    extra = [shg.Handle(0), shg.Handle(0)]
    try:
        three = optimize.gradients([handle, *extra], ps)
    finally:
        for h in extra:
            h.close()
    for (e1, g1), (e3, g3) in zip(one, three):
        assert e1 == e3
        assert np.allclose(g1, g3, rtol=1e-10, atol=0.0)
