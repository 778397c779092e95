"""Failure paths of the C ABI on the GPU.

The reference's only guards are exit(1) (src/headers/Common.h:10-20) and the
Jacobi loop's stagnation test (src/headers/Solver.h:349).  Here every
failure is a status code, per call and per batch instance.  Each condition
is forced from the host, so every kernel still runs to its normal end:
  - SHG_ENOCONV: a sweep cap (max_sweeps) on the reference Jacobi loop;
  - SHG_ENAN: a pump power whose FP32 amplitude overflows, so the fields
    hold inf/NaN (kernels index nothing by field values);
  - SHG_ENOMEM: a batch whose device arena exceeds the GPU's memory; the
    allocation fails before any launch;
  - SHG_EINVAL: an invalid instance inside a batch.
In a batch, the failing instances report their own status and the others
still give the bits of their single runs.
"""
import numpy as np
import pytest

from cushg_amd import abi, configs, shg

pytestmark = pytest.mark.gpu


def _thermal_small():
    return configs.scaled(configs.spplt256_thermal(), n=64, nz=20, nz_heat=5)


def test_jacobi_sweep_cap_returns_enoconv(handle):
    p = _thermal_small()
    p.heat_solver = abi.HEAT_JACOBI_REFERENCE
    p.max_sweeps = 7
    p.precision = abi.PREC_FP32_SPLIT
    print("case: 64x64x20 THERMAL, reference Jacobi loop capped at 7 sweeps", flush=True)
    rc, res, _ = handle.run_batch([p])
    assert rc == abi.SHG_ENOCONV and res[0].status == abi.SHG_ENOCONV
    assert "max_sweeps" in shg.last_error()
    # the FP64 plain form takes the same path
    q = p.copy()
    q.precision = abi.PREC_FP64_PLAIN
    rc, res, _ = handle.run_batch([q])
    assert rc == abi.SHG_ENOCONV and res[0].status == abi.SHG_ENOCONV


@pytest.mark.parametrize("form", [abi.PREC_FP32_SPLIT, abi.PREC_FP32_STORAGE])
def test_overflowing_fields_return_enan(handle, form):
    p = configs.spplt64()
    p.nz = 12
    p.power_W = 1e200
    p.precision = form
    print(f"case: 64x64x12 power 1e200 W, form {form}", flush=True)
    rc, res, _ = handle.run_batch([p])
    assert rc == abi.SHG_ENAN and res[0].status == abi.SHG_ENAN


def test_failing_instances_inside_a_batch(handle):
    """An invalid instance (EINVAL), a capped Jacobi loop (ENOCONV) and an
    overflowing one (ENAN) beside three good ones: each reports its own
    status, the batch returns the first failure, and the good ones equal
    their single runs bit for bit."""
    good = []
    for i, n in enumerate((64, 128, 64)):
        g = configs.spplt64()
        g.nx = g.ny = n
        g.nz = 16 + 4 * i
        g.power_W = 2.0 + i
        g.precision = abi.PREC_FP32_SPLIT
        good.append(g)
    good[2].precision = abi.PREC_FP32_STORAGE
    bad_inval = configs.spplt64()
    bad_inval.nx = 100
    bad_conv = _thermal_small()
    bad_conv.heat_solver = abi.HEAT_JACOBI_REFERENCE
    bad_conv.max_sweeps = 5
    bad_conv.precision = abi.PREC_FP32_SPLIT
    bad_nan = configs.spplt64()
    bad_nan.nz = 12
    bad_nan.power_W = 1e200
    bad_nan.precision = abi.PREC_FP32_SPLIT
    batch = [good[0], bad_inval, good[1], bad_conv, bad_nan, good[2]]
    print("case: batch of 6 with EINVAL, ENOCONV and ENAN instances", flush=True)
    rc, res, outs = handle.run_batch(batch, fields=True)
    assert rc == abi.SHG_EINVAL  # the first failure in instance order
    assert [r.status for r in res] == [abi.SHG_OK, abi.SHG_EINVAL, abi.SHG_OK, abi.SHG_ENOCONV,
                                       abi.SHG_ENAN, abi.SHG_OK]
    for k, g in zip((0, 2, 5), good):
        r1, o1 = handle.run(g, fields=True)
        assert res[k].eta == r1.eta
        assert np.array_equal(outs[k]["pump"], o1["pump"])
        assert np.array_equal(outs[k]["signal"], o1["signal"])


def test_oversized_batch_returns_enomem_and_handle_recovers(handle):
    """24000 FP64 512x512 instances need ~400 GB of device arena, more than
    the GPU holds: the call returns SHG_ENOMEM before launching anything,
    every instance reports it, and the handle then runs as before."""
    p = configs.spplt64()
    p.nx = p.ny = 512
    p.nz = 2
    p.precision = abi.PREC_FP64_FUSED
    n = 24000
    print(f"case: batch of {n} x 512x512 FP64 (arena over the device memory)", flush=True)
    with shg.Handle(0) as h2:
        rc, res, _ = h2.run_batch([p] * n)
        assert rc == abi.SHG_ENOMEM
        assert all(r.status == abi.SHG_ENOMEM for r in res)
        assert "hipMalloc" in shg.last_error()
        q = configs.spplt64()
        q.nz = 20
        q.precision = abi.PREC_FP32_SPLIT
        r2, o2 = h2.run(q, fields=True)
    r1, o1 = handle.run(q, fields=True)
    assert r2.status == abi.SHG_OK and r2.eta == r1.eta
    assert np.array_equal(o1["signal"], o2["signal"])
