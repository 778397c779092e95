"""The five workloads the path is measured and checked on, as ShgParams.

Shared conventions: MgO:sPPLT per src/headers/Crystal.h unless stated,
lambda_p = 1.064 um, focus at the crystal centre (src/cuSHG.cu:48),
T_inf = 25 C, the SH field starts at zero.  Window 800 um x 800 um.
"""
from __future__ import annotations

import math

from . import abi
from .abi import ShgParams


def _base(n: int, nz: int, lz_um: float) -> ShgParams:
    p = abi.default_params()
    p.nx = p.ny = n
    p.nz = nz
    p.lx_um = p.ly_um = 800.0
    p.lz_um = lz_um
    p.focal_um = 0.5 * lz_um
    return p


def spplt64() -> ShgParams:
    """Plain SSFM, CPU-runnable: 64x64 over 0.8 mm, 100 slices, 30 mm
    MgO:sPPLT, Lambda = 7.97 um, 1 W, 30 um waist at the centre, T = 55.66 C
    (src/cuSHG.sh:45), no absorption, THERMAL off."""
    p = _base(64, 100, 30000.0)
    p.power_W = 1.0
    p.waist_um = 30.0
    p.T_C = p.T_peltier_C = 55.66
    p.alpha_p = p.alpha_s = p.beta_s = 0.0
    p.thermal = 0
    return p


def spplt256_thermal() -> ShgParams:
    """THERMAL: 256x256, 1000 slices, 50 mm, 10 W, 40 um waist,
    alpha_p = 0.0002/cm, alpha_SH = 0.02/cm (linear absorption only),
    heat grid 256x256x100, oven 40 C, heat tolerance 1e-8 K.  The grating is
    the first-order period that phase-matches at the oven temperature."""
    p = _base(256, 1000, 50000.0)
    p.power_W = 10.0
    p.waist_um = 40.0
    p.alpha_p = 0.0002 * 1e-4
    p.alpha_s = 0.02 * 1e-4
    p.beta_s = 0.0
    p.T_C = p.T_peltier_C = 40.0
    p.Lambda_um = abi.qpm_period(p, 40.0)
    p.thermal = 1
    p.nz_heat = 100
    p.heat_solver = abi.HEAT_DIRECT
    p.heat_tol_K = 1e-8
    return p


def splitmix64(seed: int):
    """SplitMix64 stream (Steele et al. 2014) -> floats in [0, 1)."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        yield (z >> 11) / float(1 << 53)


def sweep_batch(seed: int = 0, n: int = 64, scale_steps: float = 1.0):
    """Batched sweep: n instances, seeded uniform P 0.1-20 W, waist 15-80 um,
    focus -0.5L..+0.5L about the centre, oven 30-60 C, grid from
    {128,192,256,384,512}, 500-2000 slices (times scale_steps)."""
    r = splitmix64(seed)
    out = []
    for _ in range(n):
        grid = (128, 192, 256, 384, 512)[min(4, int(next(r) * 5))]
        nz = 500 + int(next(r) * 1501)
        nz = max(2, int(round(nz * scale_steps)))
        p = _base(grid, nz, 30000.0)
        p.power_W = 0.1 + 19.9 * next(r)
        p.waist_um = 15.0 + 65.0 * next(r)
        p.focal_um = p.lz_um * (0.5 + (next(r) - 0.5))
        p.T_C = p.T_peltier_C = 30.0 + 30.0 * next(r)
        p.thermal = 0
        out.append(p)
    return out


def adjoint256(nz: int = 2000, n: int = 256, seed: int = 0):
    """Adjoint: 256x256, 2000 slices, 30 mm MgO:sPPLT, 10 W, 40 um, THERMAL off,
    poling phases ~ N(0, 0.05 rad) from SplitMix64 + Box-Muller (seed 0)."""
    p = _base(n, nz, 30000.0)
    p.power_W = 10.0
    p.waist_um = 40.0
    p.T_C = p.T_peltier_C = 55.66
    p.thermal = 0
    r = splitmix64(seed)
    ph = []
    while len(ph) < nz:
        u1, u2 = next(r), next(r)
        rad = math.sqrt(-2.0 * math.log(max(u1, 1e-300)))
        ph.append(0.05 * rad * math.cos(2 * math.pi * u2))
        ph.append(0.05 * rad * math.sin(2 * math.pi * u2))
    p.set_poling(ph[:nz])
    return p


def ppln512_thermal() -> ShgParams:
    """Full coupled case: 512x512, 4000 slices, 50 mm MgO:PPLN (literature
    constants), 30 W, 50 um, THERMAL with a 512x512x400 heat grid."""
    p = _base(512, 4000, 50000.0)
    abi.set_crystal(p, abi.CRYSTAL_MGO_PPLN)
    p.power_W = 30.0
    p.waist_um = 50.0
    p.T_C = p.T_peltier_C = 40.0
    p.Lambda_um = abi.qpm_period(p, 40.0)
    p.thermal = 1
    p.nz_heat = 400
    p.heat_solver = abi.HEAT_DIRECT
    return p


def ppln512_thermal_batch(n: int = 16, seed: int = 5):
    """n THERMAL MgO:PPLN 512x512x4000 instances (heat 512x512x400)
    over seeded (waist, focus, oven temperature) triples around the nominal
    point: waist U(40, 60) um, focus U(0.4 L, 0.6 L), oven U(38, 42) C, the
    grating fixed at the 40 C first-order period."""
    base = ppln512_thermal()
    r = splitmix64(seed)
    out = []
    for _ in range(n):
        p = base.copy()
        p.waist_um = 40.0 + 20.0 * next(r)
        p.focal_um = p.lz_um * (0.4 + 0.2 * next(r))
        p.T_C = p.T_peltier_C = 38.0 + 4.0 * next(r)
        out.append(p)
    return out


def scaled(p: ShgParams, n: int | None = None, nz: int | None = None, nz_heat: int | None = None) -> ShgParams:
    """Same physics on a smaller grid (for parity tests the oracle finishes fast)."""
    q = p.copy()
    if n is not None:
        q.nx = q.ny = n
    if nz is not None:
        q.nz = nz
    if nz_heat is not None:
        q.nz_heat = nz_heat
    return q
