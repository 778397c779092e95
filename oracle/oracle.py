"""ctypes binding of the FP64 CPU oracle (oracle/liboracle_shg.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / CPU baseline.  The product
package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from cushg_amd.abi import ShgParams, ShgResult

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_shg.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_refractive_index.restype = C.c_double
        L.orc_refractive_index.argtypes = [C.c_int32, C.c_double, C.c_double]
        L.orc_delta_k.restype = C.c_double
        L.orc_delta_k.argtypes = [C.POINTER(ShgParams), C.c_double]
        L.orc_fft1d.argtypes = [dp, C.c_int32, C.c_int32, C.c_int32]
        L.orc_fft2d.argtypes = [dp, C.c_int32, C.c_int32, C.c_int32]
        L.orc_dft1d_naive.argtypes = [dp, dp, C.c_int32, C.c_int32]
        L.orc_pump_init.argtypes = [C.POINTER(ShgParams), C.c_double, dp]
        L.orc_propagators.argtypes = [C.POINTER(ShgParams), C.c_double, dp, dp]
        L.orc_ssfm_step.argtypes = [C.POINTER(ShgParams), C.c_int32, dp, dp]
        L.orc_power.restype = C.c_double
        L.orc_power.argtypes = [C.POINTER(ShgParams), dp, C.c_double]
        L.orc_jacobi_sweep.argtypes = [C.POINTER(ShgParams), C.c_int32, dp, dp, dp]
        L.orc_heat_solve_direct.argtypes = [C.POINTER(ShgParams), C.c_int32, dp, dp]
        L.orc_tridiag_eigen.argtypes = [C.c_int32, dp, dp, dp, dp]
        L.orc_run.argtypes = [C.POINTER(ShgParams), C.POINTER(ShgResult)]
        L.orc_march_heat.argtypes = [C.POINTER(ShgParams), dp, dp, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def refractive_index(crystal: int, lam: float, T: float) -> float:
    return lib().orc_refractive_index(crystal, lam, T)


def delta_k(p: ShgParams, T: float) -> float:
    return lib().orc_delta_k(C.byref(p), T)


def fft1d(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Batched 1D FFT over the last axis (complex128)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = a.shape[-1]
    rc = lib().orc_fft1d(_dp(a.view(np.float64)), n, a.size // n, int(inverse))
    assert rc == 0
    return a


def fft2d(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    ny, nx = a.shape
    rc = lib().orc_fft2d(_dp(a.view(np.float64)), nx, ny, int(inverse))
    assert rc == 0
    return a


def dft_naive(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(a)
    lib().orc_dft1d_naive(_dp(a.view(np.float64)), _dp(out.view(np.float64)), a.size, int(inverse))
    return out


def pump_init(p: ShgParams, Tc: float) -> np.ndarray:
    a = np.zeros((p.ny, p.nx), np.complex128)
    lib().orc_pump_init(C.byref(p), Tc, _dp(a.view(np.float64)))
    return a


def propagators(p: ShgParams, Tc: float):
    pp = np.zeros((p.ny, p.nx), np.complex128)
    ps = np.zeros((p.ny, p.nx), np.complex128)
    lib().orc_propagators(C.byref(p), Tc, _dp(pp.view(np.float64)), _dp(ps.view(np.float64)))
    return pp, ps


def ssfm_step(p: ShgParams, s: int, P: np.ndarray, S: np.ndarray):
    P = np.ascontiguousarray(P, np.complex128).copy()
    S = np.ascontiguousarray(S, np.complex128).copy()
    lib().orc_ssfm_step(C.byref(p), s, _dp(P.view(np.float64)), _dp(S.view(np.float64)))
    return P, S


def power(p: ShgParams, A: np.ndarray, n: float) -> float:
    A = np.ascontiguousarray(A, np.complex128)
    return lib().orc_power(C.byref(p), _dp(A.view(np.float64)), n)


def jacobi_sweep(p: ShgParams, nzh: int, Ti: np.ndarray, Q: np.ndarray) -> np.ndarray:
    Ti = np.ascontiguousarray(Ti, np.float64)
    Q = np.ascontiguousarray(Q, np.float64)
    Tf = np.empty_like(Ti)
    lib().orc_jacobi_sweep(C.byref(p), nzh, _dp(Ti), _dp(Tf), _dp(Q))
    return Tf


def heat_solve_direct(p: ShgParams, nzh: int, Q: np.ndarray) -> np.ndarray:
    Q = np.ascontiguousarray(Q, np.float64)
    T = np.empty_like(Q)
    rc = lib().orc_heat_solve_direct(C.byref(p), nzh, _dp(Q), _dp(T))
    assert rc == 0
    return T


def march_heat(p: ShgParams, T: np.ndarray):
    """orc_march_heat: one THERMAL march on T[nzh, ny, nx]; returns the heat
    source Q[nzh, ny, nx] and the (pump, SH) powers after every slice."""
    T = np.ascontiguousarray(T, np.float64)
    Q = np.empty_like(T)
    pw = np.empty((p.nz, 2))
    rc = lib().orc_march_heat(C.byref(p), _dp(T), _dp(Q), _dp(pw))
    assert rc == 0
    return Q, pw


def tridiag_eigen(diag: np.ndarray, off: np.ndarray):
    n = len(diag)
    d = np.ascontiguousarray(diag, np.float64)
    o = np.ascontiguousarray(off if n > 1 else np.zeros(1), np.float64)
    lam = np.empty(n)
    V = np.empty((n, n))
    rc = lib().orc_tridiag_eigen(n, _dp(d), _dp(o), _dp(lam), _dp(V))
    assert rc == 0
    return lam, V


def run(p: ShgParams, fields: bool = True, temperature: bool = False):
    """orc_run: returns (ShgResult, dict of arrays)."""
    r = ShgResult()
    out = {}
    if fields:
        out["pump"] = np.zeros((p.ny, p.nx), np.complex128)
        out["signal"] = np.zeros((p.ny, p.nx), np.complex128)
        r.pump_out = _dp(out["pump"].view(np.float64))
        r.signal_out = _dp(out["signal"].view(np.float64))
    if temperature and p.thermal:
        nzh = p.nz_heat if p.nz_heat > 0 else p.nz
        out["T"] = np.zeros((nzh, p.ny, p.nx), np.float64)
        r.temperature_out = _dp(out["T"])
    rc = lib().orc_run(C.byref(p), C.byref(r))
    if rc != 0:
        raise RuntimeError(f"orc_run failed with status {rc}")
    return r, out
