"""Pins the FP64 oracle (CPU, no GPU needed).

Each check is chosen so that a plausible restatement error (dropped term,
wrong sign or index, transposed operand) fails it:
  - FFT against numpy golden vectors and a long-double naive DFT;
  - n(lambda,T) and dk(T) against a 40-digit decimal evaluation and the
    phase-matching temperature;
  - pump + propagator + FFT normalisation against the analytic Gaussian
    beam after free-space propagation;
  - the RK4 schemes against the closed-form SHG solution (sech/tanh);
  - power conservation (Manley-Rowe) of the full Strang march;
  - the Jacobi sweep and the direct heat solver against an independently
    assembled finite-volume system solved by scipy;
  - the THERMAL march with a uniform temperature against the non-THERMAL
    march (pins the running dk integral and the heat-slice mapping);
  - the reference Jacobi loop against the direct solver;
  - the heat source against the exponential pump decay and the power the
    fields lose slice by slice (energy balance with two-photon absorption).
"""
import json
import math
import os

import numpy as np
import pytest

from cushg_amd import abi, configs
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C_LIGHT = 299792458.0 * 1e6 / 1e12
EPS0 = 8.8541878128e-12 * 1e12 / 1e6


# ---------------------------------------------------------------- FFT

@pytest.fixture(scope="module")
def fftgold():
    return np.load(os.path.join(GOLD, "fft_vectors.npz"))


@pytest.mark.parametrize("n", [64, 128, 192, 256, 384, 512])
def test_fft1d_golden(fftgold, n):
    x = fftgold[f"x{n}"]
    X = O.fft1d(x)
    Xi = O.fft1d(x, inverse=True)
    assert np.abs(X - fftgold[f"X{n}"]).max() / np.abs(fftgold[f"X{n}"]).max() < 1e-14
    assert np.abs(Xi - fftgold[f"I{n}"]).max() / np.abs(fftgold[f"I{n}"]).max() < 1e-14


@pytest.mark.parametrize("shape", [(64, 64), (128, 192)])
def test_fft2d_golden(fftgold, shape):
    ny, nx = shape
    x = fftgold[f"x2d_{ny}_{nx}"]
    X = O.fft2d(x)
    ref = fftgold[f"X2d_{ny}_{nx}"]
    assert np.abs(X - ref).max() / np.abs(ref).max() < 1e-14


@pytest.mark.parametrize("n", [7, 12, 30, 48])
def test_fft_vs_naive_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for inv in (False, True):
        a = O.fft1d(x, inverse=inv)
        b = O.dft_naive(x, inverse=inv)
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()


def test_fft_roundtrip_parseval():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((64, 128)) + 1j * rng.standard_normal((64, 128))
    X = O.fft2d(x)
    assert abs(np.sum(np.abs(X) ** 2) / x.size - np.sum(np.abs(x) ** 2)) < 1e-12 * np.sum(np.abs(x) ** 2)
    y = O.fft2d(X, inverse=True) / x.size
    assert np.abs(y - x).max() < 1e-14


# ---------------------------------------------------------------- material

def test_sellmeier_and_dk_decimal():
    g = json.load(open(os.path.join(GOLD, "sellmeier.json")))
    p = abi.default_params()
    p.Lambda_um = g["Lambda_um"]
    for row in g["rows"]:
        T = row["T"]
        assert O.refractive_index(0, 1.064, T) == pytest.approx(row["n_p"], rel=1e-15)
        assert O.refractive_index(0, 0.532, T) == pytest.approx(row["n_s"], rel=1e-15)
        assert O.delta_k(p, T) == pytest.approx(row["dk"], rel=1e-11, abs=1e-16)


def _ppln_params():
    p = abi.default_params()
    abi.set_crystal(p, abi.CRYSTAL_MGO_PPLN)
    return p


def test_ppln_sellmeier_and_dk_decimal():
    """MgO:PPLN against a 40-digit decimal evaluation of Gayer et al.'s
    published n_e(lambda, T), coefficients typed afresh from the paper
    (tests/golden/make_golden.py), and dk(T) with the database's period and
    Jundt's expansion of the grating."""
    g = json.load(open(os.path.join(GOLD, "sellmeier_ppln.json")))
    p = _ppln_params()
    assert (p.Lambda_um, p.therm_exp_a, p.therm_exp_b) == (g["Lambda_um"], g["therm_exp_a"], g["therm_exp_b"])
    for row in g["rows"]:
        T = row["T"]
        assert O.refractive_index(abi.CRYSTAL_MGO_PPLN, 1.064, T) == pytest.approx(row["n_p"], rel=1e-15)
        assert O.refractive_index(abi.CRYSTAL_MGO_PPLN, 0.532, T) == pytest.approx(row["n_s"], rel=1e-15)
        # dk cancels terms of ~25 1/um: FP64 leaves ~1e-15 of that
        assert O.delta_k(p, T) == pytest.approx(row["dk"], rel=1e-11, abs=3e-14)
    for row in g["index"]:
        assert O.refractive_index(abi.CRYSTAL_MGO_PPLN, row["lambda_um"], row["T"]) == pytest.approx(
            row["n"], rel=1e-15)


def test_ppln_against_published_magnitudes():
    """Independent of the coefficient table: MgO:LN's extraordinary index at
    1064 nm is about 2.148 near room temperature (congruent LN without MgO:
    2.156), the 532 nm index about 2.22, and the first-order QPM period for
    1064 -> 532 nm SHG in MgO:PPLN is about 6.9 um, shrinking as T rises.
    A swapped pole or a dropped term moves these far outside the bounds."""
    n1 = O.refractive_index(abi.CRYSTAL_MGO_PPLN, 1.064, 25.0)
    n2 = O.refractive_index(abi.CRYSTAL_MGO_PPLN, 0.532, 25.0)
    assert 2.140 < n1 < 2.156 and 2.215 < n2 < 2.235
    periods = [1.064 / (2 * (O.refractive_index(abi.CRYSTAL_MGO_PPLN, 0.532, T)
                             - O.refractive_index(abi.CRYSTAL_MGO_PPLN, 1.064, T))) for T in (25.0, 60.0, 100.0)]
    assert all(6.8 < L < 7.0 for L in periods) and periods[0] > periods[1] > periods[2]
    # normal dispersion: n falls with wavelength across the fit range
    lams = np.linspace(0.5, 3.0, 26)
    ns = [O.refractive_index(abi.CRYSTAL_MGO_PPLN, float(l), 40.0) for l in lams]
    assert np.all(np.diff(ns) < 0)


def test_dk_zero_near_56p92C():
    """dk vanishes at 56.9158 C for Lambda = 7.97 um (plane-wave phase matching)."""
    p = abi.default_params()
    lo, hi = 50.0, 60.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if O.delta_k(p, mid) > 0:
            lo = mid
        else:
            hi = mid
    assert abs(lo - 56.9158) < 1e-3


def test_qpm_period_helper_phase_matches():
    p = configs.spplt256_thermal()
    assert abs(O.delta_k(p, 40.0)) < 1e-14


def test_tridiag_eigen_vs_numpy():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 40, 101):
        d = rng.standard_normal(n)
        o = rng.standard_normal(max(n - 1, 0))
        lam, V = O.tridiag_eigen(d, o if n > 1 else np.zeros(1))
        A = np.diag(d) + (np.diag(o, 1) + np.diag(o, -1) if n > 1 else 0)
        assert np.allclose(np.sort(lam), np.linalg.eigvalsh(A), atol=1e-12)
        assert np.abs(A @ V - V * lam).max() < 1e-12
        assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12


# ---------------------------------------------------------------- optics

def _gauss(p, Tc, zrel):
    """Analytic paraxial Gaussian beam, focus at focal_um, observed zrel after
    the input plane; same amplitude normalisation as the pump init."""
    n = O.refractive_index(p.crystal, p.lambda_p_um, Tc)
    w2 = p.waist_um ** 2
    zR = math.pi * n * w2 / p.lambda_p_um
    Ap0 = math.sqrt(4 * p.power_W / (EPS0 * C_LIGHT * math.pi * n * w2))
    M = 1 + 1j * (zrel - p.focal_um) / zR
    dx, dy = p.lx_um / (p.nx - 1), p.ly_um / (p.ny - 1)
    x = (np.arange(p.nx) - p.nx / 2) * dx
    y = (np.arange(p.ny) - p.ny / 2) * dy
    r2 = x[None, :] ** 2 + y[:, None] ** 2
    return (Ap0 / M) * np.exp(-r2 / (w2 * M))


def test_pump_init_closed_form_and_power():
    p = configs.spplt64()
    p.nx = p.ny = 128
    A = O.pump_init(p, p.T_C)
    assert np.abs(A - _gauss(p, p.T_C, 0.0)).max() < 1e-12 * np.abs(A).max()
    n = O.refractive_index(0, 1.064, p.T_C)
    assert O.power(p, A, n) == pytest.approx(p.power_W, rel=1e-6)


def test_free_space_gaussian_matches_analytic():
    """No coupling, no loss: the march is pure diffraction; the saved output
    (after nz-1 slices = the crystal length) is the analytic beam there."""
    p = configs.spplt64()
    p.nx = p.ny = 128
    p.deff_um_per_V = 0.0
    p.nz = 30
    r, o = O.run(p)
    ref = _gauss(p, p.T_C, p.lz_um)
    assert np.abs(o["pump"] - ref).max() < 1e-8 * np.abs(ref).max()
    assert np.abs(o["signal"]).max() == 0.0
    # diffraction is unitary: power kept to round-off
    assert r.P_pump_out_W == pytest.approx(r.P_pump_in_W, rel=1e-12)


def test_walkoff_free_space_is_the_shifted_analytic_beam():
    """Walk-off in free space: the propagator's linear phase
    exp(-2 pi i fx dz tan(rho)) per slice shifts the beam along +x by
    tan(rho) per unit length (shift theorem), so after nz-1 slices the saved
    pump is the analytic Gaussian centred at x = L tan(rho).  A sign slip,
    sin/rho for tan, a missing 2 pi or the wrong axis fails this."""
    p = configs.spplt64()
    p.nx = p.ny = 256
    p.lx_um = p.ly_um = 1600.0
    p.deff_um_per_V = 0.0
    p.nz = 30
    p.rho_p = 0.0042
    r, o = O.run(p)
    n = O.refractive_index(p.crystal, p.lambda_p_um, p.T_C)
    w2 = p.waist_um ** 2
    zR = math.pi * n * w2 / p.lambda_p_um
    Ap0 = math.sqrt(4 * p.power_W / (EPS0 * C_LIGHT * math.pi * n * w2))
    M = 1 + 1j * (p.lz_um - p.focal_um) / zR
    dx, dy = p.lx_um / (p.nx - 1), p.ly_um / (p.ny - 1)
    shift = p.lz_um * math.tan(p.rho_p)
    assert shift > 10 * dx
    x = (np.arange(p.nx) - p.nx / 2) * dx - shift
    y = (np.arange(p.ny) - p.ny / 2) * dy
    ref = (Ap0 / M) * np.exp(-(x[None, :] ** 2 + y[:, None] ** 2) / (w2 * M))
    assert np.abs(o["pump"] - ref).max() < 1e-8 * np.abs(ref).max()
    assert r.P_pump_out_W == pytest.approx(r.P_pump_in_W, rel=1e-12)


def test_walkoff_of_both_fields_is_a_lattice_shift_of_the_coupled_march():
    """With the same walk-off for pump and SH and tan(rho) dz equal to one
    grid step, each slice's diffraction moves both fields by exactly one
    sample along +x.  The nonlinear step is pointwise and the same at every x
    (uniform dk), so the coupled march with walk-off equals the march without
    it, rolled by nz-1 samples: this pins the SH walk-off term and its sign
    with real conversion on."""
    p = configs.spplt64()
    p.nx = p.ny = 128
    p.lx_um = p.ly_um = 1600.0
    p.power_W = 20.0
    p.nz = 24
    dx = p.lx_um / (p.nx - 1)
    dz = p.lz_um / (p.nz - 1)
    r0, o0 = O.run(p)
    q = p.copy()
    q.rho_p = q.rho_s = math.atan(dx / dz)
    r1, o1 = O.run(q)
    assert r0.eta > 1e-4
    for f in ("pump", "signal"):
        ref = np.roll(o0[f], p.nz - 1, axis=1)
        assert np.abs(o1[f] - ref).max() < 1e-10 * np.abs(ref).max(), f
    assert r1.eta == pytest.approx(r0.eta, rel=1e-10)
    # a walk-off of the SH alone is not a lattice shift: the fields separate
    s = p.copy()
    s.rho_s = q.rho_s
    r2, _ = O.run(s)
    assert abs(r2.eta - r0.eta) > 1e-3 * r0.eta


def test_constant_poling_phase_rotates_the_sh_only():
    """A phase phi on every slice enters the SH drive as exp(+i phi) and the
    pump's back-conversion as exp(-i phi).  The map (P, S) -> (P, S e^{i phi})
    then commutes with every step, so the march returns the pump unchanged
    and the SH rotated by phi (with depletion and absorption on).  A wrong
    sign on either field breaks the invariance."""
    p = configs.spplt64()
    p.power_W = 30.0
    p.alpha_p, p.alpha_s, p.beta_s = 0.17e-6, 1.57e-6, 5e-5
    p.nz = 40
    r0, o0 = O.run(p)
    assert r0.eta > 1e-3
    phi = 0.7
    q = p.copy()
    q.set_poling(np.full(p.nz, phi))
    r1, o1 = O.run(q)
    sc = np.abs(o0["pump"]).max()
    assert np.abs(o1["pump"] - o0["pump"]).max() < 1e-12 * sc
    assert np.abs(o1["signal"] - np.exp(1j * phi) * o0["signal"]).max() < 1e-12 * sc
    assert r1.eta == pytest.approx(r0.eta, rel=1e-12)


def test_poling_phase_first_order_closed_form():
    """Plane wave, small signal (undepleted pump, no loss): the SH after the
    march is a sum of each slice's drive, to first order in the coupling.
    Per RK4 call the reference scheme adds dz/6 (k1 + 2k2 + 2k3 + k4) with
    k_j = i ks P0^2 exp(i(z_j dk + phi_s)) at z_j = s dz + {0, dz/4, dz/4,
    dz/2}, both calls of slice s at the same z (src/headers/Solver.h:231-273),
    and nz-1 slices precede the saved state.  Random per-slice phases make a
    shifted slice index, a sign slip or a phase missing from one stage fail."""
    p = _pw_params(31)
    P0 = 1e-3
    rng = np.random.default_rng(11)
    phi = rng.uniform(-math.pi, math.pi, p.nz)
    p.Lambda_um = 7.96  # off phase matching: dk dz is about 1.2 rad
    p.set_poling(phi)
    dz = p.lz_um / (p.nz - 1)
    dk = O.delta_k(p, p.T_C)
    assert abs(dk * dz) > 0.5
    P = np.full((p.ny, p.nx), P0, np.complex128)
    S = np.zeros_like(P)
    for s in range(p.nz - 1):
        P, S = O.ssfm_step(p, s, P, S)
    dQ = 2 * p.deff_um_per_V / math.pi
    n_s = O.refractive_index(0, 0.5 * p.lambda_p_um, p.T_C)
    ks = 0.5 * math.pi * dQ / (n_s * 0.5 * p.lambda_p_um)
    s_idx = np.arange(p.nz - 1)
    e = np.exp(1j * (s_idx * dz * dk + phi[:-1]))
    stages = 1 + 4 * np.exp(1j * dk * dz / 4) + np.exp(1j * dk * dz / 2)
    S_ref = 2 * (dz / 6) * 1j * ks * P0 ** 2 * np.sum(e) * stages
    assert abs(S[0, 0] - S_ref) < 1e-6 * abs(S_ref)
    assert np.abs(S - S[0, 0]).max() < 1e-12 * abs(S_ref)
    # the same march with the phases moved by one slice is far off
    q = p.copy()
    q.set_poling(np.roll(phi, 1))
    P1 = np.full((p.ny, p.nx), P0, np.complex128)
    S1 = np.zeros_like(P1)
    for s in range(p.nz - 1):
        P1, S1 = O.ssfm_step(q, s, P1, S1)
    assert abs(S1[0, 0] - S_ref) > 0.1 * abs(S_ref)


def test_poling_ramp_acts_as_a_dk_shift():
    """A linear ramp phi_s = delta s dz adds delta to the phase mismatch at
    the start of every slice; it differs from dk + delta only by delta times
    the sub-step offsets (at most dz/2), so the ramp run sits far closer to the
    dk + delta run than to the unramped or the dk - delta run."""
    p = configs.spplt64()
    p.nz = 60
    dz = p.lz_um / (p.nz - 1)
    delta = 2e-4
    ramp = p.copy()
    ramp.set_poling(delta * dz * np.arange(p.nz))
    r_ramp, o_ramp = O.run(ramp)
    r0, o0 = O.run(p)
    # dk + delta through the grating period: 1/(Lambda' e) = 1/(Lambda e) + delta/(2 pi),
    # e the grating's thermal expansion factor at T_C
    d = p.T_C - 25.0
    e = 1 + p.therm_exp_a * d + p.therm_exp_b * d * d
    up = p.copy()
    up.Lambda_um = 1.0 / (1.0 / p.Lambda_um + e * delta / (2 * math.pi))
    r_up, o_up = O.run(up)
    dn = p.copy()
    dn.Lambda_um = 1.0 / (1.0 / p.Lambda_um - e * delta / (2 * math.pi))
    r_dn, o_dn = O.run(dn)
    assert O.delta_k(up, p.T_C) - O.delta_k(p, p.T_C) == pytest.approx(delta, rel=1e-9)
    d_up = np.abs(o_ramp["signal"] - o_up["signal"]).max()
    assert d_up < 0.05 * np.abs(o_ramp["signal"] - o0["signal"]).max()
    assert d_up < 0.05 * np.abs(o_ramp["signal"] - o_dn["signal"]).max()
    assert r_ramp.eta == pytest.approx(r_up.eta, rel=2e-2)


def _astigmatic_gauss(p, Tc, zrel):
    """Analytic elliptical (astigmatic) Gaussian beam: each transverse axis is
    an independent 1D Gaussian beam, amplitude 1/sqrt(M) per axis, with
    M = 1 + i (z - f) / zR of its own waist; normalised to power_W."""
    n = O.refractive_index(p.crystal, p.lambda_p_um, Tc)
    wx, wy = p.waist_um, p.waist_y_um
    Ap0 = math.sqrt(4 * p.power_W / (EPS0 * C_LIGHT * math.pi * n * wx * wy))
    Mx = 1 + 1j * (zrel - p.focal_um) / (math.pi * n * wx ** 2 / p.lambda_p_um)
    My = 1 + 1j * (zrel - p.focal_um) / (math.pi * n * wy ** 2 / p.lambda_p_um)
    dx, dy = p.lx_um / (p.nx - 1), p.ly_um / (p.ny - 1)
    x = (np.arange(p.nx) - p.nx / 2) * dx
    y = (np.arange(p.ny) - p.ny / 2) * dy
    return (Ap0 / (np.sqrt(Mx) * np.sqrt(My))) * np.exp(-x[None, :] ** 2 / (wx ** 2 * Mx)
                                                         - y[:, None] ** 2 / (wy ** 2 * My))


def test_elliptical_pump_is_the_astigmatic_gaussian():
    """waist_y_um > 0: the pump is the separable astigmatic Gaussian (rank 1,
    each axis with its own Rayleigh range), carries power_W, reduces to the
    circular beam when the waists agree, and free-space diffraction carries
    it to the analytic beam at the crystal end (a wrong axis, a missing
    sqrt(M) or a circular M would each fail one of these)."""
    p = configs.spplt64()
    p.nx = p.ny = 128
    p.lx_um = p.ly_um = 1000.0
    p.waist_y_um = 1.7 * p.waist_um
    p.focal_um = 0.3 * p.lz_um
    A = O.pump_init(p, p.T_C)
    assert np.abs(A - _astigmatic_gauss(p, p.T_C, 0.0)).max() < 1e-12 * np.abs(A).max()
    sv = np.linalg.svd(A, compute_uv=False)
    assert sv[1] < 1e-12 * sv[0]
    n = O.refractive_index(0, 1.064, p.T_C)
    assert O.power(p, A, n) == pytest.approx(p.power_W, rel=1e-6)
    # |A|^2 falls to 1/e^2 at x = wx on the focus plane (f = 0)
    q = p.copy()
    q.focal_um = 0.0
    B = O.pump_init(q, q.T_C)
    assert B[q.ny // 2, q.nx // 2] == pytest.approx(math.sqrt(4 * q.power_W / (
        EPS0 * C_LIGHT * math.pi * n * q.waist_um * q.waist_y_um)), rel=1e-14)
    c = q.copy()
    c.waist_y_um = c.waist_um
    q.waist_y_um = 0.0
    assert np.abs(O.pump_init(c, c.T_C) - O.pump_init(q, q.T_C)).max() < 1e-14 * np.abs(B).max()
    # propagation through the crystal without coupling
    p.deff_um_per_V = 0.0
    p.nz = 30
    r, o = O.run(p)
    ref = _astigmatic_gauss(p, p.T_C, p.lz_um)
    assert np.abs(o["pump"] - ref).max() < 1e-8 * np.abs(ref).max()


def _plane_wave_run(p, steps, P0):
    P = np.full((p.ny, p.nx), P0, np.complex128)
    S = np.zeros_like(P)
    for s in range(steps):
        P, S = O.ssfm_step(p, s, P, S)
    return P[0, 0], S[0, 0]


def _tanh_closed_form(p, P0, z, kmul):
    dQ = 2 * p.deff_um_per_V / math.pi
    n_p = O.refractive_index(0, p.lambda_p_um, p.T_C)
    n_s = O.refractive_index(0, 0.5 * p.lambda_p_um, p.T_C)
    kp = kmul * math.pi * dQ / (n_p * p.lambda_p_um)
    ks = kmul * 0.5 * math.pi * dQ / (n_s * 0.5 * p.lambda_p_um)
    G = math.sqrt(kp * ks) * P0
    return P0 / math.cosh(G * z), 1j * math.sqrt(ks / kp) * P0 * math.tanh(G * z)


def _pw_params(nz):
    p = configs.spplt64()
    p.alpha_p = p.alpha_s = p.beta_s = 0.0
    p.nx = p.ny = 8
    p.nz = nz
    p.Lambda_um = abi.qpm_period(p, p.T_C)
    return p


def test_strang_rk4_matches_tanh():
    """Plane wave at dk = 0: textbook SHG, P = P0 sech(Gz), S = i sqrt(ks/kp) P0 tanh(Gz)."""
    P0 = 1.5
    p = _pw_params(41)
    p.scheme = abi.SCHEME_STRANG
    steps = p.nz - 1
    dz = p.lz_um / (p.nz - 1)
    P, S = _plane_wave_run(p, steps, P0)
    Pe, Se = _tanh_closed_form(p, P0, steps * dz, 2.0)
    assert abs(S - Se) / abs(Se) < 1e-9
    assert abs(P - Pe) / abs(Pe) < 1e-9


def test_reference_scheme_first_order_to_tanh():
    """The reference RK4 (stage inputs without dz, full-dz half slices) is a
    first-order integrator of the same ODE with the textbook coupling:
    halving dz halves the error."""
    P0 = 1.5
    errs = []
    for nz in (21, 41, 81):
        p = _pw_params(nz)
        steps = p.nz - 1
        P, S = _plane_wave_run(p, steps, P0)
        _, Se = _tanh_closed_form(p, P0, p.lz_um, 2.0)
        errs.append(abs(S - Se) / abs(Se))
    assert errs[2] < 2e-2
    assert 1.7 < errs[0] / errs[1] < 2.3
    assert 1.7 < errs[1] / errs[2] < 2.3


def test_strang_march_conserves_power():
    """Manley-Rowe: with alpha = beta = 0, n_p|P|^2 + n_s|S|^2 is invariant."""
    p = configs.spplt64()
    p.scheme = abi.SCHEME_STRANG
    p.power_W = 50.0
    p.nz = 60
    r, _ = O.run(p)
    assert r.P_sh_W > 1e-3 * p.power_W
    assert (r.P_pump_out_W + r.P_sh_W) == pytest.approx(r.P_pump_in_W, rel=1e-7)


# ---------------------------------------------------------------- heat

@pytest.fixture(scope="module")
def heatgold():
    return np.load(os.path.join(GOLD, "heat_fv.npz"))


def _heat_params(spec, bc=abi.HEAT_BC_OPEN_TOP):
    nx, ny, nz, lx, ly, lz, k, h, Ti, Tp = spec
    p = abi.default_params()
    p.nx, p.ny = int(nx), int(ny)
    p.nz = int(nz)
    p.lx_um, p.ly_um, p.lz_um = lx, ly, lz
    p.k_th, p.h_conv, p.T_inf_C, p.T_peltier_C = k, h, Ti, Tp
    p.heat_bc = bc
    return p, int(nz)


# the open-top set (the reference's) and the oven-surrounded set, both
# against their own independently assembled scipy solutions
BCS = [pytest.param("", abi.HEAT_BC_OPEN_TOP, id="open-top"),
       pytest.param("o", abi.HEAT_BC_OVEN_SURROUNDED, id="oven-surrounded")]


@pytest.mark.parametrize("pre,bc", BCS)
@pytest.mark.parametrize("case", [0, 1, 2])
def test_direct_heat_solve_vs_scipy_fv(heatgold, case, pre, bc):
    p, nzh = _heat_params(heatgold[f"{pre}spec{case}"], bc)
    T = O.heat_solve_direct(p, nzh, heatgold[f"{pre}Q{case}"])
    ref = heatgold[f"{pre}T{case}"]
    assert np.abs(T - ref).max() < 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("pre,bc", BCS)
@pytest.mark.parametrize("case", [0, 1, 2])
def test_jacobi_sweep_fixed_point_is_fv_solution(heatgold, case, pre, bc):
    """Every boundary formula of the sweep (faces, edges, vertices) leaves the
    finite-volume steady state unchanged."""
    p, nzh = _heat_params(heatgold[f"{pre}spec{case}"], bc)
    ref = heatgold[f"{pre}T{case}"]
    Tf = O.jacobi_sweep(p, nzh, ref, heatgold[f"{pre}Q{case}"])
    assert np.abs(Tf - ref).max() < 1e-9 * np.abs(ref).max()


def test_oven_boundary_differs_from_open_top(heatgold):
    """The two boundary sets give different steady states for the same
    source (a restatement that ignored heat_bc would pass the tests above
    only by accident of the fixtures)."""
    spec, Q = heatgold["ospec0"], heatgold["oQ0"]
    To = O.heat_solve_direct(_heat_params(spec, abi.HEAT_BC_OVEN_SURROUNDED)[0], int(spec[2]), Q)
    Tt = O.heat_solve_direct(_heat_params(spec, abi.HEAT_BC_OPEN_TOP)[0], int(spec[2]), Q)
    assert np.abs(To - Tt).max() > 1e-3 * np.abs(To - spec[9]).max()
    # the four lateral planes sit at the oven temperature
    assert np.all(To[:, :, 0] == spec[9]) and np.all(To[:, :, -1] == spec[9])
    assert np.all(To[:, 0, :] == spec[9]) and np.all(To[:, -1, :] == spec[9])


@pytest.mark.parametrize("pre,bc", BCS)
def test_jacobi_iteration_converges_to_direct(heatgold, pre, bc):
    p, nzh = _heat_params(heatgold[f"{pre}spec1"], bc)
    Q = heatgold[f"{pre}Q1"]
    T = np.full_like(Q, p.T_peltier_C)
    for _ in range(20000):
        Tn = O.jacobi_sweep(p, nzh, T, Q)
        if np.abs(Tn - T).max() < 1e-13:
            break
        T = Tn
    assert np.abs(T - heatgold[f"{pre}T1"]).max() < 1e-9


def test_thermal_uniform_equals_plain():
    """T_inf = T_oven = T and no absorption: T stays uniform, the running dk
    integral reduces to z*dk and the THERMAL march equals the plain one."""
    p = configs.spplt64()
    p.nx = p.ny = 32
    p.nz = 24
    p.T_inf_C = p.T_peltier_C = p.T_C
    p.alpha_p = p.alpha_s = p.beta_s = 0.0
    r0, o0 = O.run(p)
    for nzh, solver in ((24, abi.HEAT_DIRECT), (7, abi.HEAT_DIRECT), (24, abi.HEAT_JACOBI_REFERENCE)):
        q = p.copy()
        q.thermal = 1
        q.nz_heat = nzh
        q.heat_solver = solver
        r1, o1 = O.run(q)
        assert r1.eta == pytest.approx(r0.eta, rel=1e-12)
        assert np.abs(o1["signal"] - o0["signal"]).max() < 1e-12 * np.abs(o0["signal"]).max()
        assert r1.outer_iters == 2  # pass 1 solves, pass 2 converges in one sweep


def test_thermal_jacobi_reference_loop_vs_direct():
    """With a tight tolerance the reference Jacobi loop reaches the direct
    solver's steady state, so the thermo-optic results agree."""
    p = configs.spplt256_thermal()
    p.nx = p.ny = 16
    p.nz = 12
    p.nz_heat = 6
    p.lx_um = p.ly_um = 200.0
    p.lz_um = 2000.0
    p.power_W = 50.0
    p.waist_um = 20.0
    p.alpha_p, p.alpha_s = 2e-4, 2e-3
    p.focal_um = 1000.0
    p.jacobi_tol = 1e-11
    p.check_every = 50
    rd, od = O.run(p, temperature=True)
    q = p.copy()
    q.heat_solver = abi.HEAT_JACOBI_REFERENCE
    rj, oj = O.run(q, temperature=True)
    assert rd.peak_dT_K > 1e-3
    assert np.abs(oj["T"] - od["T"]).max() < 1e-8
    assert rj.eta == pytest.approx(rd.eta, rel=1e-9)
    assert rj.jacobi_sweeps > rd.jacobi_sweeps


def _march_heat_setup(nz, nzh):
    p = configs.spplt64()
    p.scheme = abi.SCHEME_STRANG
    p.thermal = 1
    p.power_W = 50.0
    p.nz, p.nz_heat = nz, nzh
    p.alpha_p, p.alpha_s, p.beta_s = 2e-5, 2e-4, 1.0
    dA = (p.lx_um / (p.nx - 1)) * (p.ly_um / (p.ny - 1))
    dz = p.lz_um / (p.nz - 1)
    n_p = O.refractive_index(p.crystal, p.lambda_p_um, p.T_C)
    P_in = O.power(p, O.pump_init(p, p.T_C), n_p)
    return p, dA, dz, P_in, np.full((nzh, p.ny, p.nx), p.T_C)


def test_heat_source_without_conversion_is_the_absorbed_pump():
    """deff = 0: the pump decays as exp(-alpha_p z), so the heat of optical
    slice s integrated over the plane is alpha_p P_in exp(-alpha_p (s+1) dz);
    a heat slice holds the mean over the optical slices s with
    floor(s nzh / nz) = h.  Pins Q's alpha_p I_p term (the 1/2 c eps0 n_p
    intensity) against the beam power and the slice-to-heat-slab averaging."""
    p, dA, dz, P_in, T = _march_heat_setup(40, 7)
    p.deff_um_per_V = 0.0
    Q, pw = O.march_heat(p, T)
    assert np.all(pw[:, 1] == 0.0)
    s = np.arange(p.nz)
    per_slice = p.alpha_p * P_in * np.exp(-p.alpha_p * (s + 1) * dz)
    slab = (s * p.nz_heat) // p.nz
    want = np.array([per_slice[slab == h].mean() for h in range(p.nz_heat)])
    got = Q.sum(axis=(1, 2)) * dA
    assert np.abs(got / want - 1).max() < 1e-9
    # uniform in z within a slab only through the average: the first and last
    # slabs differ by the decay over the crystal
    assert got[0] / got[-1] == pytest.approx(want[0] / want[-1], rel=1e-12)


def test_heat_source_closes_the_energy_balance():
    """With conversion, linear SH absorption and two-photon absorption: the
    power the two fields lose over a slice equals the heat deposited in it
    (trapezoid of the plane-integrated Q at the slice ends; the z = 0 end
    holds only the pump).  alpha_p I_p, alpha_s I_s and beta_s I_s^2 each
    carry more than 5% of the heat, so a wrong intensity factor (n_s for
    n_p, a dropped 1/2, I_s for I_s^2) fails it."""
    p, dA, dz, P_in, T = _march_heat_setup(120, 120)
    Q, pw = O.march_heat(p, T)
    H = Q.sum(axis=(1, 2)) * dA
    ends = np.concatenate([[p.alpha_p * P_in], H])
    total = np.concatenate([[P_in], pw.sum(axis=1)])
    loss = total[:-1] - total[1:]
    assert np.abs(loss / (0.5 * dz * (ends[:-1] + ends[1:])) - 1).max() < 4e-4
    share_p = p.alpha_p * pw[:, 0].sum() / H.sum()
    share_s = p.alpha_s * pw[:, 1].sum() / H.sum()
    assert share_p > 0.05 and share_s > 0.05 and 1 - share_p - share_s > 0.05
    assert pw[-1, 1] > 1e-3 * p.power_W
