"""GPU parity of the plain FP64 form against the FP64 oracle, and agreement
of the fused FP32-field forms with it (the FP32 batched form's accuracy bar
is eta relative error <= 1e-5 against FP64)."""
import numpy as np
import pytest

from cushg_amd import abi, configs
from oracle import oracle as O


def rel(a, b):
    return abs(a - b) / abs(b)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def plain(p):
    q = p.copy()
    q.precision = abi.PREC_FP64_PLAIN
    return q


@pytest.mark.gpu
@pytest.mark.parametrize("n,nz,scheme", [(64, 100, abi.SCHEME_REFERENCE), (192, 20, abi.SCHEME_REFERENCE),
                                          (128, 30, abi.SCHEME_STRANG), (384, 8, abi.SCHEME_REFERENCE)])
def test_plain_fp64_vs_oracle(handle, n, nz, scheme):
    p = configs.spplt64()
    p.nx = p.ny = n
    p.nz = nz
    p.scheme = scheme
    p.alpha_p, p.alpha_s, p.beta_s = 0.17e-6, 1.57e-6, 5e-5
    print(f"case plain n={n} nz={nz} scheme={scheme}", flush=True)
    ro, o = O.run(p)
    rg, g = handle.run(plain(p))
    print(f"  eta gpu={rg.eta:.15e} cpu={ro.eta:.15e} rel={rel(rg.eta, ro.eta):.2e}")
    assert rel(rg.eta, ro.eta) < 1e-11
    assert rel_l2(g["pump"], o["pump"]) < 1e-11
    assert rel_l2(g["signal"], o["signal"]) < 1e-11


@pytest.mark.gpu
def test_plain_thermal_vs_oracle(handle):
    p = configs.scaled(configs.spplt256_thermal(), n=64, nz=60, nz_heat=15)
    p.lx_um = p.ly_um = 400.0
    p.alpha_p, p.alpha_s = 2e-6, 2e-5
    ro, o = O.run(p, temperature=True)
    rg, g = handle.run(plain(p), temperature=True)
    print(f"  eta gpu={rg.eta:.12e} cpu={ro.eta:.12e} rel={rel(rg.eta, ro.eta):.2e} dT {rg.peak_dT_K:.9f}/{ro.peak_dT_K:.9f}")
    assert rg.outer_iters == ro.outer_iters
    # the heat source is stored in FP32 on the GPU (as in the reference)
    assert rel(rg.eta, ro.eta) < 1e-7
    assert rel(rg.peak_dT_K, ro.peak_dT_K) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("form,tol", [(abi.PREC_FP32_MIXED, 1e-5), (abi.PREC_FP32_STORAGE, 2e-6)])
def test_fused_forms_agree_with_plain(handle, form, tol):
    """Both FP32-field forms stay within 1e-5 of the FP64 form; with FP64
    arithmetic (only the stores round) within 2e-6."""
    p = configs.scaled(configs.spplt256_thermal(), n=128, nz=300, nz_heat=30)
    p.precision = form
    rm, _ = handle.run(p, fields=False)
    rp, _ = handle.run(plain(p), fields=False)
    print(f"  eta fused={rm.eta:.12e} plain={rp.eta:.12e} rel={rel(rm.eta, rp.eta):.2e}")
    assert rel(rm.eta, rp.eta) <= tol
    assert rel(rm.peak_dT_K, rp.peak_dT_K) <= tol


@pytest.mark.gpu
@pytest.mark.parametrize("form,bound", [(abi.PREC_FP32_MIXED, 3e-6), (abi.PREC_FP32_STORAGE, 5e-7)])
@pytest.mark.parametrize("n", [64, 128, 192, 256, 384, 512])
def test_fp32_diffraction_has_no_power_drift(handle, n, form, bound):
    """Pure diffraction (no coupling, no loss) is unitary: over 2000 FP32
    slices the pump power must not drift.  A fixed rounding error in a
    transform constant (3e-8 per transform round trip) accumulates
    coherently to ~6e-5 here, and an FP32-rounded twiddle table to ~5e-6 at
    N = 192; data-dependent rounding, partly correlated between consecutive
    slices of a slowly varying beam, stays near 1e-6 with FP32 arithmetic;
    with FP64 arithmetic only the two FP32 stores per slice round."""
    p = configs.ppln512_thermal()
    p.thermal = 0
    p.nx = p.ny = n
    p.nz = 2000
    p.deff_um_per_V = 0.0
    p.alpha_p = p.alpha_s = p.beta_s = 0.0
    p.precision = form
    r, _ = handle.run(p, fields=False)
    drift = r.P_pump_out_W / r.P_pump_in_W - 1.0
    print(f"  n={n} precision={form}: power drift over 2000 slices {drift:.2e}")
    assert abs(drift) < bound
