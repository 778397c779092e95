"""ctypes mirror of include/shg.h (the C ABI of the HIP library).

The structs here must match include/shg.h byte for byte; tests check the
sizes against the library's own view.  Field meanings and defaults follow
the reference's compile-time constants (src/headers/Crystal.h:19-52,
src/cuSHG.cu:11-15,46-53) -- see ``default_params``.
"""
from __future__ import annotations

import ctypes as C
import math

SHG_OK = 0
SHG_EINVAL = -1
SHG_ENOMEM = -2
SHG_EDEVICE = -3
SHG_ENOCONV = -4
SHG_ENAN = -5
SHG_EUNSUPPORTED = -6

STATUS_NAMES = {
    SHG_OK: "SHG_OK",
    SHG_EINVAL: "SHG_EINVAL",
    SHG_ENOMEM: "SHG_ENOMEM",
    SHG_EDEVICE: "SHG_EDEVICE",
    SHG_ENOCONV: "SHG_ENOCONV",
    SHG_ENAN: "SHG_ENAN",
    SHG_EUNSUPPORTED: "SHG_EUNSUPPORTED",
}

CRYSTAL_MGO_SPPLT = 0
CRYSTAL_MGO_PPLN = 1
SCHEME_REFERENCE = 0
SCHEME_STRANG = 1
PREC_FP32_MIXED = 0
PREC_FP64_PLAIN = 1
PREC_FP32_STORAGE = 2
PREC_FP64_FUSED = 3
PREC_FP32_SPLIT = 4
HEAT_JACOBI_REFERENCE = 0
HEAT_DIRECT = 1
# row-pass plan value of the wide form (include/shg.h SHG_PLAN_ROWS_WIDE)
PLAN_ROWS_WIDE = 16
HEAT_BC_OPEN_TOP = 0
HEAT_BC_OVEN_SURROUNDED = 1

FFT_SIZES = (64, 128, 192, 256, 384, 512)


class ShgParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nz_heat", C.c_int32),
        ("lx_um", C.c_double), ("ly_um", C.c_double), ("lz_um", C.c_double),
        ("crystal", C.c_int32), ("scheme", C.c_int32),
        ("lambda_p_um", C.c_double), ("deff_um_per_V", C.c_double), ("Lambda_um", C.c_double),
        ("alpha_p", C.c_double), ("alpha_s", C.c_double), ("beta_s", C.c_double),
        ("rho_p", C.c_double), ("rho_s", C.c_double),
        ("k_th", C.c_double), ("h_conv", C.c_double),
        ("therm_exp_a", C.c_double), ("therm_exp_b", C.c_double),
        ("power_W", C.c_double), ("waist_um", C.c_double), ("focal_um", C.c_double),
        ("T_C", C.c_double), ("T_inf_C", C.c_double), ("T_peltier_C", C.c_double),
        ("thermal", C.c_int32), ("precision", C.c_int32), ("heat_solver", C.c_int32),
        ("max_outer", C.c_int32), ("check_every", C.c_int32), ("heat_bc", C.c_int32),
        ("max_sweeps", C.c_int64),
        ("jacobi_tol", C.c_double), ("heat_tol_K", C.c_double),
        ("waist_y_um", C.c_double),
        ("poling_phase", C.POINTER(C.c_double)),
    ]

    def copy(self) -> "ShgParams":
        q = ShgParams()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(ShgParams))
        # keep the poling buffer alive with the copy
        q._keep = getattr(self, "_keep", None)
        return q

    def set_poling(self, phases) -> None:
        """Attach a per-slice poling phase vector (len nz) or None."""
        if phases is None:
            self.poling_phase = C.POINTER(C.c_double)()
            self._keep = None
            return
        buf = (C.c_double * len(phases))(*[float(v) for v in phases])
        self._keep = buf
        self.poling_phase = C.cast(buf, C.POINTER(C.c_double))

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "poling_phase"}


class ShgResult(C.Structure):
    _fields_ = [
        ("eta", C.c_double), ("P_sh_W", C.c_double), ("P_pump_out_W", C.c_double),
        ("P_pump_in_W", C.c_double), ("peak_dT_K", C.c_double), ("T_max_C", C.c_double),
        ("heat_residual_K", C.c_double), ("wall_s", C.c_double),
        ("steps_done", C.c_int64), ("jacobi_sweeps", C.c_int64),
        ("outer_iters", C.c_int32), ("status", C.c_int32),
        ("pump_out", C.POINTER(C.c_double)), ("signal_out", C.POINTER(C.c_double)),
        ("temperature_out", C.POINTER(C.c_double)),
        ("adjoint_sweeps", C.c_int32), ("reserved0", C.c_int32),
    ]


# -- material helpers in plain Python (used to pick QPM periods) -----------

_PI = 3.14159265358979323846


def _n_spplt(L: float, T: float) -> float:
    # src/headers/Crystal.h:54-70 (Bruner et al., MgO:sPPLT)
    A, B, Cc, D = 4.502483, 0.007294, 0.185087, -0.02357
    E, F, G, H = 0.073423, 0.199595, 0.001, 7.99724
    tk = T + 273.15
    b = 3.483933e-8 * tk * tk
    c = 1.607839e-8 * tk * tk
    L2 = L * L
    return math.sqrt(A + (B + b) / (L2 - (Cc + c) ** 2) + E / (L2 - F * F)
                     + G / (L2 - H * H) + D * L2)


def _n_ppln(L: float, T: float) -> float:
    # Gayer et al., Appl. Phys. B 91, 343 (2008), 5% MgO:CLN, e-ray
    a1, a2, a3, a4, a5, a6 = 5.756, 0.0983, 0.2020, 189.32, 12.52, 1.32e-2
    b1, b2, b3, b4 = 2.860e-6, 4.700e-8, 6.113e-8, 1.516e-4
    f = (T - 24.5) * (T + 570.82)
    L2 = L * L
    g = a3 + b3 * f
    return math.sqrt(a1 + b1 * f + (a2 + b2 * f) / (L2 - g * g)
                     + (a4 + b4 * f) / (L2 - a5 * a5) - a6 * L2)


def refractive_index(crystal: int, lam_um: float, T_C: float) -> float:
    return _n_ppln(lam_um, T_C) if crystal == CRYSTAL_MGO_PPLN else _n_spplt(lam_um, T_C)


def qpm_period(p: ShgParams, T_C: float) -> float:
    """First-order grating period [um] (at 25 C) that phase-matches at T_C,
    i.e. dk(T_C) = 0 in src/headers/PhaseMatching.h:80."""
    lp = p.lambda_p_um
    ls = 0.5 * lp
    np_ = refractive_index(p.crystal, lp, T_C)
    ns_ = refractive_index(p.crystal, ls, T_C)
    d = T_C - 25.0
    return 1.0 / ((ns_ / ls - 2.0 * np_ / lp) * (1.0 + p.therm_exp_a * d + p.therm_exp_b * d * d))


def set_crystal(p: ShgParams, crystal: int) -> None:
    """Crystal database; mirrors shg_params_set_crystal()."""
    p.crystal = crystal
    if crystal == CRYSTAL_MGO_PPLN:
        # literature values (no reference counterpart): d33 of MgO:LN
        # (Shoji et al. 1997), conductivity and grating expansion of LN
        p.deff_um_per_V = 25.0e-6
        p.Lambda_um = 6.97
        p.k_th = 4.6e-6
        p.therm_exp_a = 1.54e-5
        p.therm_exp_b = 5.3e-9
    else:
        # src/headers/Crystal.h:40-52, src/headers/PhaseMatching.h:53,74
        p.deff_um_per_V = 11.00e-6
        p.Lambda_um = 7.97
        p.k_th = 8.4e-6
        p.therm_exp_a = 2.2e-6
        p.therm_exp_b = -5.9e-9
    p.alpha_p = 0.17e-6
    p.alpha_s = 1.57e-6
    p.beta_s = 5e-5
    p.rho_p = 0.0
    p.rho_s = 0.0
    p.h_conv = 10e-12


def default_params() -> ShgParams:
    """The reference's defaults: src/headers/Crystal.h:19-52,
    src/cuSHG.cu:11-15 and :46-53, src/cuSHG.sh:43-45."""
    p = ShgParams()
    p.nx = p.ny = 128
    p.nz = 100
    p.nz_heat = 0
    p.lx_um, p.ly_um, p.lz_um = 2000.0, 1000.0, 30000.0
    set_crystal(p, CRYSTAL_MGO_SPPLT)
    p.scheme = SCHEME_REFERENCE
    p.lambda_p_um = 1.064
    p.power_W = 30.0
    p.waist_um = 29.0
    p.focal_um = 0.5 * p.lz_um
    p.T_C = 55.66
    p.T_inf_C = 25.0
    p.T_peltier_C = 55.66
    p.thermal = 0
    p.precision = PREC_FP32_STORAGE
    p.heat_solver = HEAT_JACOBI_REFERENCE
    p.max_outer = 20
    p.check_every = 5000
    p.max_sweeps = 50_000_000
    p.jacobi_tol = 5e-4
    p.heat_tol_K = 1e-8
    p.set_poling(None)
    return p
