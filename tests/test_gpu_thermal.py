"""GPU parity of the THERMAL path against the FP64 oracle.

Tolerances: eta and peak dT relative error <= 1e-5 (the accuracy bar of the
FP32 batched form); the temperature field to 1e-5 of the peak rise (both
solve to the same discrete
steady state -- the FP32-mixed form refines its FP32 fast-diagonalisation
solve with FP64 residuals until the update is < 1e-8 K -- but the heat
source deposited by the FP32 fields differs from the FP64 one at ~1e-7).
"""
import numpy as np
import pytest

from cushg_amd import abi, configs
from oracle import oracle as O


def rel(a, b):
    return abs(a - b) / abs(b)


def small_thermal(n=64, nz=40, nzh=12, solver=abi.HEAT_DIRECT):
    p = configs.spplt256_thermal()
    p.nx = p.ny = n
    p.nz = nz
    p.nz_heat = nzh
    p.lx_um = p.ly_um = 400.0
    p.lz_um = 10000.0
    p.focal_um = 5000.0
    p.power_W = 20.0
    p.waist_um = 25.0
    p.alpha_p, p.alpha_s = 2e-6, 2e-5  # strong absorption: kelvin-scale heating
    p.heat_solver = solver
    p.Lambda_um = abi.qpm_period(p, p.T_C)
    return p


@pytest.mark.gpu
def test_thermal_uniform_equals_plain(handle):
    """No heating and T_inf = T_oven = T: the THERMAL march (per-voxel n(T)
    and running dk integral) reproduces the plain march."""
    p = configs.spplt64()
    p.nz = 60
    p.T_inf_C = p.T_peltier_C = p.T_C
    r0, o0 = handle.run(p)
    q = p.copy()
    q.thermal = 1
    q.nz_heat = 15
    q.heat_solver = abi.HEAT_DIRECT
    r1, o1 = handle.run(q)
    print(f"eta plain={r0.eta:.12e} thermal={r1.eta:.12e} outer={r1.outer_iters}")
    assert r1.outer_iters == 2
    assert rel(r1.eta, r0.eta) < 2e-6
    assert abs(r1.peak_dT_K) < 1e-9


FORMS = [pytest.param(abi.PREC_FP32_MIXED, id="mixed"), pytest.param(abi.PREC_FP32_STORAGE, id="storage"),
         pytest.param(abi.PREC_FP32_SPLIT, id="split")]


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("nzh", [40, 12])
def test_thermal_direct_vs_oracle(handle, nzh, form):
    p = small_thermal(nzh=nzh)
    p.precision = form
    print(f"case thermal direct nx={p.nx} nz={p.nz} nzh={nzh} precision={form}", flush=True)
    ro, o = O.run(p, temperature=True)
    rg, g = handle.run(p, temperature=True)
    print(f"  eta gpu={rg.eta:.10e} cpu={ro.eta:.10e} rel={rel(rg.eta, ro.eta):.2e} "
          f"dT gpu={rg.peak_dT_K:.8f} cpu={ro.peak_dT_K:.8f} outer {rg.outer_iters}/{ro.outer_iters} "
          f"maxTerr={np.abs(g['T'] - o['T']).max():.2e}")
    assert ro.peak_dT_K > 0.1
    assert rg.outer_iters == ro.outer_iters
    assert rel(rg.eta, ro.eta) <= 1e-5
    assert rel(rg.peak_dT_K, ro.peak_dT_K) <= 1e-5
    # T agrees to 1e-5 of the peak rise (the heat source itself is FP32)
    assert np.abs(g["T"] - o["T"]).max() <= 1e-5 * ro.peak_dT_K
    # the last pass is the one that converged in one sweep: its one-sweep
    # change is below the reference tolerance
    assert rg.heat_residual_K < p.jacobi_tol


@pytest.mark.gpu
@pytest.mark.parametrize("n,bc", [(192, abi.HEAT_BC_OPEN_TOP), (384, abi.HEAT_BC_OPEN_TOP),
                                  (192, abi.HEAT_BC_OVEN_SURROUNDED)])
def test_thermal_direct_other_sizes_vs_oracle(handle, n, bc):
    """The non-power-of-two grids: the heat solver's x fold at 192 (96 + 96) and 384, the z fold at 8 slices,
    and the 3 x 64 / 6 x 64 transforms, against the oracle."""
    p = small_thermal(n=n, nz=12, nzh=8)
    p.heat_bc = bc
    p.precision = abi.PREC_FP32_SPLIT
    print(f"case thermal direct n={n} bc={bc}", flush=True)
    ro, o = O.run(p, temperature=True)
    rg, g = handle.run(p, temperature=True)
    print(f"  eta gpu={rg.eta:.10e} cpu={ro.eta:.10e} rel={rel(rg.eta, ro.eta):.2e} "
          f"dT {rg.peak_dT_K:.6e}/{ro.peak_dT_K:.6e} outer {rg.outer_iters}/{ro.outer_iters}")
    assert rg.outer_iters == ro.outer_iters
    assert rel(rg.eta, ro.eta) <= 1e-5
    assert rel(rg.peak_dT_K, ro.peak_dT_K) <= 1e-5
    assert np.abs(g["T"] - o["T"]).max() <= 1e-5 * ro.peak_dT_K


@pytest.mark.gpu
def test_thermal_jacobi_reference_loop_vs_oracle(handle):
    """The reference loop (sweeps + norm test every check_every) on both sides."""
    p = small_thermal(n=64, nz=16, nzh=6, solver=abi.HEAT_JACOBI_REFERENCE)
    p.lz_um = 2000.0
    p.focal_um = 1000.0
    p.check_every = 200
    p.jacobi_tol = 1e-3
    p.max_outer = 4
    ro, o = O.run(p, temperature=True)
    rg, g = handle.run(p, temperature=True)
    print(f"  sweeps gpu={rg.jacobi_sweeps} cpu={ro.jacobi_sweeps} eta rel={rel(rg.eta, ro.eta):.2e}")
    assert rg.jacobi_sweeps == ro.jacobi_sweeps
    assert rg.outer_iters == ro.outer_iters
    assert np.abs(g["T"] - o["T"]).max() <= 1e-5 * ro.peak_dT_K
    assert rel(rg.eta, ro.eta) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS)
def test_spplt256_thermal_scaled_vs_oracle(handle, form):
    """The 256x256 THERMAL sPPLT physics (50 mm, 10 W, 40 um, oven 40 C) at
    128x128, 200 z-steps, heat 128x128x25; FP64 arithmetic holds eta to
    2e-6."""
    p = configs.scaled(configs.spplt256_thermal(), n=128, nz=200, nz_heat=25)
    p.precision = form
    tol = 2e-6 if form == abi.PREC_FP32_STORAGE else 1e-5
    ro, o = O.run(p, temperature=True)
    rg, g = handle.run(p, temperature=True)
    print(f"  eta gpu={rg.eta:.10e} cpu={ro.eta:.10e} rel={rel(rg.eta, ro.eta):.2e} "
          f"dT {rg.peak_dT_K:.6e}/{ro.peak_dT_K:.6e} outer {rg.outer_iters}/{ro.outer_iters}")
    assert rg.outer_iters == ro.outer_iters
    assert rel(rg.eta, ro.eta) <= tol
    assert rel(rg.peak_dT_K, ro.peak_dT_K) <= tol


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS)
def test_oven_surrounded_direct_vs_oracle(handle, form):
    """The oven-surrounded boundary set (four lateral planes at the oven
    temperature, convective end faces; src/headers/Tfield.h:205-240)."""
    p = small_thermal(nzh=12)
    p.heat_bc = abi.HEAT_BC_OVEN_SURROUNDED
    p.precision = form
    print(f"case oven-surrounded direct precision={form}", flush=True)
    ro, o = O.run(p, temperature=True)
    rg, g = handle.run(p, temperature=True)
    print(f"  eta gpu={rg.eta:.10e} cpu={ro.eta:.10e} rel={rel(rg.eta, ro.eta):.2e} "
          f"dT gpu={rg.peak_dT_K:.8f} cpu={ro.peak_dT_K:.8f} outer {rg.outer_iters}/{ro.outer_iters}")
    assert ro.peak_dT_K > 0.05
    assert rg.outer_iters == ro.outer_iters
    assert rel(rg.eta, ro.eta) <= 1e-5
    assert rel(rg.peak_dT_K, ro.peak_dT_K) <= 1e-5
    assert np.abs(g["T"] - o["T"]).max() <= 1e-5 * ro.peak_dT_K
    # the lateral planes stay at the oven temperature
    for sl in (g["T"][:, :, 0], g["T"][:, :, -1], g["T"][:, 0, :], g["T"][:, -1, :]):
        assert np.all(sl == p.T_peltier_C)


@pytest.mark.gpu
def test_oven_surrounded_jacobi_reference_loop_vs_oracle(handle):
    p = small_thermal(n=64, nz=16, nzh=6, solver=abi.HEAT_JACOBI_REFERENCE)
    p.heat_bc = abi.HEAT_BC_OVEN_SURROUNDED
    p.lz_um = 2000.0
    p.focal_um = 1000.0
    p.check_every = 200
    p.jacobi_tol = 1e-3
    p.max_outer = 4
    ro, o = O.run(p, temperature=True)
    rg, g = handle.run(p, temperature=True)
    print(f"  sweeps gpu={rg.jacobi_sweeps} cpu={ro.jacobi_sweeps} eta rel={rel(rg.eta, ro.eta):.2e}")
    assert rg.jacobi_sweeps == ro.jacobi_sweeps
    assert rg.outer_iters == ro.outer_iters
    assert np.abs(g["T"] - o["T"]).max() <= 1e-5 * max(ro.peak_dT_K, 1e-3)
    assert rel(rg.eta, ro.eta) <= 1e-5
