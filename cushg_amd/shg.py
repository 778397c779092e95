"""ctypes binding of libshg_hip.so (the C ABI of include/shg.h).

``Solver.run`` mirrors the reference's
``Solver::run(Power, waist, focalpoint, Temp, T_inf, Tpeltier,
save_only_last, save_temperature)`` (src/headers/Solver.h:149,311): same
argument names and meaning, results returned instead of only written to
files.  Errors raise ``ShgError`` carrying the C status code; the library
never exits the process (the reference's CHECK does,
src/headers/Common.h:10-20).

There is no CPU fallback: if the HIP library is missing or no GPU is
present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from .abi import ShgParams, ShgResult

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libshg_hip.so")

# every symbol include/shg.h declares
EXPORTS = (
    "shg_fft_size_supported", "shg_params_default", "shg_params_set_crystal",
    "shg_refractive_index", "shg_delta_k", "shg_validate", "shg_last_error",
    "shg_handle_create", "shg_handle_destroy", "shg_run", "shg_run_traced", "shg_run_batch", "shg_adjoint",
    "shg_cli_main", "shg_last_kernel_times", "shg_handle_set_profiling", "shg_set_plan",
    "shg_get_plan", "shg_autotune", "shg_set_graph_steps", "shg_run_noalloc", "shg_fft2d",
)

_lib = None

# typedef int (*shg_slice_fn)(int32_t slice, int32_t pass, const double* pump,
#                             const double* signal, void* user)
SLICE_FN = C.CFUNCTYPE(C.c_int, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


class ShgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ShgError(abi.SHG_EUNSUPPORTED,
                       f"{LIB_PATH} not built; run __graft_entry__.build() or cushg_amd/build.py")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(ShgParams)
    R = C.POINTER(ShgResult)
    H = C.c_void_p
    L.shg_fft_size_supported.argtypes = [C.c_int32]
    L.shg_params_default.argtypes = [P]
    L.shg_params_default.restype = None
    L.shg_params_set_crystal.argtypes = [P, C.c_int32]
    L.shg_refractive_index.argtypes = [C.c_int32, C.c_double, C.c_double]
    L.shg_refractive_index.restype = C.c_double
    L.shg_delta_k.argtypes = [P, C.c_double]
    L.shg_delta_k.restype = C.c_double
    L.shg_validate.argtypes = [P]
    L.shg_last_error.restype = C.c_char_p
    L.shg_handle_create.argtypes = [C.c_int32, C.POINTER(H)]
    L.shg_handle_destroy.argtypes = [H]
    L.shg_run.argtypes = [H, P, R]
    L.shg_run_traced.argtypes = [H, P, C.c_int32, SLICE_FN, C.c_void_p, R]
    L.shg_run_batch.argtypes = [H, P, C.c_int32, R]
    L.shg_adjoint.argtypes = [H, P, C.c_int32, C.POINTER(C.c_double), R]
    L.shg_cli_main.argtypes = [C.c_int, C.POINTER(C.c_char_p), P, C.c_char_p]
    L.shg_last_kernel_times.argtypes = [H, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                        C.POINTER(C.c_int32), C.c_int32]
    L.shg_handle_set_profiling.argtypes = [H, C.c_int32]
    L.shg_set_plan.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.shg_get_plan.argtypes = [H, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.shg_autotune.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.shg_set_graph_steps.argtypes = [H, C.c_int32]
    L.shg_run_noalloc.argtypes = [H, P, R]
    L.shg_fft2d.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                            C.POINTER(C.c_double)]
    _lib = L
    return L


def last_error() -> str:
    return (lib().shg_last_error() or b"").decode()


def _check(rc: int):
    if rc != abi.SHG_OK:
        raise ShgError(rc, last_error())


def default_params() -> ShgParams:
    """Reference defaults as filled by the library (shg_params_default)."""
    p = ShgParams()
    lib().shg_params_default(C.byref(p))
    return p


def validate(p: ShgParams) -> int:
    return lib().shg_validate(C.byref(p))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Handle:
    """Owns device buffers, plans and the HIP stream (shg_handle)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().shg_handle_create(device, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().shg_handle_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _prep_result(self, p: ShgParams, fields: bool, temperature: bool):
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
        return r, out

    def run(self, p: ShgParams, fields: bool = True, temperature: bool = False):
        """shg_run: returns (ShgResult, dict of output arrays)."""
        r, out = self._prep_result(p, fields, temperature)
        _check(lib().shg_run(self._h, C.byref(p), C.byref(r)))
        return r, out

    def run_noalloc(self, p: ShgParams, fields: bool = True):
        """shg_run_noalloc: one non-THERMAL fused-form run with every device
        buffer in the library's fixed per-GPU scratch (no allocation).
        Returns (ShgResult, dict of output arrays)."""
        r, out = self._prep_result(p, fields, False)
        _check(lib().shg_run_noalloc(self._h, C.byref(p), C.byref(r)))
        return r, out

    def run_traced(self, p: ShgParams, every: int, fn, fields: bool = False, temperature: bool = False):
        """shg_run_traced: fn(slice, pass, pump, signal) receives the state
        before the SSFM step of slices 0, every, ... of every thermo-optic
        pass as (ny, nx) complex128 arrays (copies); a truthy return value
        cancels the run.  Returns (ShgResult, dict of output arrays)."""
        r, out = self._prep_result(p, fields, temperature)
        n2 = p.nx * p.ny
        err = []

        def tramp(slice_, pass_, pp, ps_, _user):
            try:
                a = np.ctypeslib.as_array(pp, (2 * n2,)).view(np.complex128).reshape(p.ny, p.nx).copy()
                b = np.ctypeslib.as_array(ps_, (2 * n2,)).view(np.complex128).reshape(p.ny, p.nx).copy()
                return 1 if fn(slice_, pass_, a, b) else 0
            except Exception as e:  # an exception in the callback cancels the run
                err.append(e)
                return 1

        cb = SLICE_FN(tramp)
        rc = lib().shg_run_traced(self._h, C.byref(p), every, cb, None, C.byref(r))
        if err:
            raise err[0]
        _check(rc)
        return r, out

    def run_batch(self, ps, fields: bool = False, temperature: bool = False):
        """shg_run_batch over a list of ShgParams; per-instance status in r.status."""
        n = len(ps)
        arr = (ShgParams * n)()
        for i, p in enumerate(ps):
            C.memmove(C.byref(arr[i]), C.byref(p), C.sizeof(ShgParams))
        res = (ShgResult * n)()
        outs = []
        for i, p in enumerate(ps):
            r, out = self._prep_result(p, fields, temperature)
            res[i] = r
            outs.append(out)
        rc = lib().shg_run_batch(self._h, arr, n, res)
        return rc, list(res), outs

    def adjoint(self, p: ShgParams, checkpoint_every: int = 0):
        """shg_adjoint: returns (ShgResult, d_eta array [waist, focus, T, poling...])."""
        g = np.zeros(3 + p.nz, np.float64)
        r = ShgResult()
        _check(lib().shg_adjoint(self._h, C.byref(p), checkpoint_every, _dp(g), C.byref(r)))
        return r, g

    def fft2d(self, x: np.ndarray, precision: int = abi.PREC_FP64_FUSED, inverse: bool = False) -> np.ndarray:
        """shg_fft2d: the library's unnormalised 2D DFT of x[ny, nx] (complex)."""
        a = np.ascontiguousarray(x, dtype=np.complex128)
        if a.ndim != 2:
            raise ValueError("x must be 2-D [ny, nx]")
        ny, nx = a.shape
        out = np.empty_like(a)
        _check(lib().shg_fft2d(self._h, nx, ny, precision, int(inverse), _dp(a.view(np.float64)),
                               _dp(out.view(np.float64))))
        return out

    def set_profiling(self, on: bool):
        _check(lib().shg_handle_set_profiling(self._h, int(on)))

    def kernel_times(self) -> dict:
        cap = 64
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        nl = (C.c_int32 * cap)()
        n = lib().shg_last_kernel_times(self._h, names, ms, nl, cap)
        return {names[i].decode(): (ms[i], nl[i]) for i in range(n)}

    def set_plan(self, n: int, rows: int = 0, cols: int = 0, precision: int = abi.PREC_FP32_MIXED):
        _check(lib().shg_set_plan(self._h, n, precision, rows, cols))

    def get_plan(self, n: int, precision: int = abi.PREC_FP32_MIXED) -> tuple[int, int]:
        """(rows per row-pass workgroup, columns per column-pass workgroup)."""
        r, c = C.c_int32(), C.c_int32()
        _check(lib().shg_get_plan(self._h, n, precision, C.byref(r), C.byref(c)))
        return r.value, c.value

    def autotune(self, n: int, precision: int = abi.PREC_FP32_MIXED, batch: int = 1, nz: int = 64):
        """Time every supported plan for n x n on a synthetic march of `batch`
        instances and keep the fastest; returns the chosen plan."""
        _check(lib().shg_autotune(self._h, n, precision, batch, nz))
        return self.get_plan(n, precision)

    def set_graph_steps(self, steps: int):
        """Replay the fused march as hipGraphs of `steps` z-steps (0: direct
        launches; -1, the default: graphs for batches that march on several
        streams)."""
        _check(lib().shg_set_graph_steps(self._h, steps))


def save_matrix_complex(A: np.ndarray, name: str, out_dir: str = ".") -> None:
    """Text format of saveMatrixComplex (src/headers/Files.h:69-90):
    <name>_r.dat / <name>_i.dat, ny rows of nx tab-terminated values."""
    for suf, part in (("_r.dat", A.real), ("_i.dat", A.imag)):
        with open(os.path.join(out_dir, name + suf), "w") as f:
            for row in part.astype(np.float32):
                f.write("".join(f"{float(v):.20g}\t" for v in row) + "\n")


def save_tensor_real(T: np.ndarray, name: str, out_dir: str = ".") -> None:
    """Text format of saveTensorReal (src/headers/Files.h:169-188):
    <name>_<z>.dat per z-slice, ny rows of nx tab-terminated values at the
    default stream precision (6 significant digits)."""
    for z in range(T.shape[0]):
        with open(os.path.join(out_dir, f"{name}_{z}.dat"), "w") as f:
            for row in T[z].astype(np.float32):
                f.write("".join(f"{float(v):.6g}\t" for v in row) + "\n")


class Solver:
    """Mirror of the reference's class Solver (src/headers/Solver.h:116-150)
    over the C ABI; `params` holds what the reference fixes at compile time."""

    def __init__(self, params: ShgParams | None = None, device: int = 0):
        self.params = params.copy() if params is not None else default_params()
        self.handle = Handle(device)

    def run(self, Power, waist, focalpoint, Temp, T_inf, Tpeltier, save_only_last=True,
            save_temperature=False, out_dir=None):
        p = self.params.copy()
        p.power_W, p.waist_um, p.focal_um = float(Power), float(waist), float(focalpoint)
        p.T_C, p.T_inf_C, p.T_peltier_C = float(Temp), float(T_inf), float(Tpeltier)
        if save_only_last or out_dir is None:
            r, out = self.handle.run(p, fields=True, temperature=bool(save_temperature))
        else:
            # every slice's state before its SSFM step (src/headers/Solver.h:289-294)
            def dump(s, _pass, pump, signal):
                save_matrix_complex(pump, f"Pump_{s}", out_dir)
                save_matrix_complex(signal, f"Signal_{s}", out_dir)
            r, out = self.handle.run_traced(p, 1, dump, fields=True, temperature=bool(save_temperature))
        if out_dir is not None:
            if save_only_last:
                save_matrix_complex(out["pump"], "Pump_out", out_dir)
                save_matrix_complex(out["signal"], "Signal_out", out_dir)
            if save_temperature and "T" in out:
                save_tensor_real(out["T"], "Tfinal", out_dir)
        return r, out
