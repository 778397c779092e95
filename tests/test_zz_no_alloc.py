"""shg_run_noalloc (the zero-allocation run) against shg_run, through the C ABI.

shg_run_noalloc has no counterpart in the reference: it serves callers that
cannot allocate device memory.  Checks:
  - two calls in a row, the second on a smaller input (512x512 FP32-split
    with walk-off and poling phases, then 128x128 FP32-storage), each equal
    to shg_run's result on the same parameters bit for bit (eta and both
    output fields);
  - a call on a fresh handle leaves the device's free memory unchanged (a
    run through the handle's arena would take >= 16 MiB at this size);
  - the limits: nz beyond SHG_NOALLOC_MAX_NZ is SHG_EINVAL with a message,
    THERMAL and the FP64 plain form are SHG_EUNSUPPORTED.
"""
import ctypes as C

import numpy as np
import pytest

from cushg_amd import abi, configs, shg

pytestmark = pytest.mark.gpu

MAX_NZ = 16384  # SHG_NOALLOC_MAX_NZ in include/shg.h


def _hip_runtime():
    """The HIP runtime the library itself runs on (the copy mapped from the
    ROCm install; torch, when imported, maps a second copy of its own)."""
    shg.lib()
    with open("/proc/self/maps") as f:
        for line in f:
            path = line.split()[-1]
            if "/libamdhip64.so" in path and "/torch/" not in path:
                return C.CDLL(path)
    raise RuntimeError("HIP runtime not mapped")


def _device_free_bytes() -> int:
    hip = _hip_runtime()
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _case_a():
    p = configs.spplt64()
    p.nx = p.ny = 512
    p.lx_um = p.ly_um = 1200.0
    p.nz = 80
    p.power_W = 12.0
    p.rho_p, p.rho_s = 0.002, 0.004
    p.set_poling(np.random.default_rng(3).normal(0.0, 0.05, p.nz))
    p.precision = abi.PREC_FP32_SPLIT
    return p


def _case_b():
    p = configs.spplt64()
    p.nx = p.ny = 128
    p.nz = 50
    p.power_W = 3.0
    p.precision = abi.PREC_FP32_STORAGE
    return p


def test_noalloc_equals_shg_run_on_two_calls_in_a_row():
    a, b = _case_a(), _case_b()
    print("case: noalloc 512x512x80 split (walk-off, poling), then 128x128x50 storage", flush=True)
    with shg.Handle(0) as h:
        ra, oa = h.run_noalloc(a)
        rb, ob = h.run_noalloc(b)
        ra0, oa0 = h.run(a)
        rb0, ob0 = h.run(b)
    for r, o, r0, o0 in ((ra, oa, ra0, oa0), (rb, ob, rb0, ob0)):
        print(f"  eta noalloc={r.eta:.12e} shg_run={r0.eta:.12e}", flush=True)
        assert r.status == abi.SHG_OK
        assert r.eta == r0.eta and r.P_pump_out_W == r0.P_pump_out_W and r.P_pump_in_W == r0.P_pump_in_W
        assert np.array_equal(o["pump"], o0["pump"])
        assert np.array_equal(o["signal"], o0["signal"])
        assert r.steps_done == r0.steps_done
    assert ra.eta > 1e-4


def test_noalloc_allocates_no_device_memory():
    p = _case_a()
    p.precision = abi.PREC_FP64_FUSED
    print("case: noalloc 512x512x80 FP64 on a fresh handle, device free memory before/after", flush=True)
    with shg.Handle(0) as warm:  # loads every kernel's code object first
        warm.run_noalloc(p)
        warm.run(p)
    with shg.Handle(0) as h:
        before = _device_free_bytes()
        r, _ = h.run_noalloc(p)
        after = _device_free_bytes()
        print(f"  free before {before} after {after} (diff {before - after})", flush=True)
        assert r.status == abi.SHG_OK
        assert before - after < (4 << 20)


def test_noalloc_limits():
    with shg.Handle(0) as h:
        p = configs.spplt64()
        p.nz = MAX_NZ + 1
        p.lz_um = 3e5
        p.precision = abi.PREC_FP32_SPLIT
        with pytest.raises(shg.ShgError) as e:
            h.run_noalloc(p, fields=False)
        assert e.value.status == abi.SHG_EINVAL and "maximum" in str(e.value)
        q = configs.scaled(configs.spplt256_thermal(), n=64, nz=20, nz_heat=4)
        with pytest.raises(shg.ShgError) as e:
            h.run_noalloc(q, fields=False)
        assert e.value.status == abi.SHG_EUNSUPPORTED
        q = configs.spplt64()
        q.precision = abi.PREC_FP64_PLAIN
        with pytest.raises(shg.ShgError) as e:
            h.run_noalloc(q, fields=False)
        assert e.value.status == abi.SHG_EUNSUPPORTED
        # the handle still serves shg_run_noalloc after the refusals
        r, _ = h.run_noalloc(_case_b(), fields=False)
        assert r.status == abi.SHG_OK
