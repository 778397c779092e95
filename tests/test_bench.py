"""bench.py without a GPU: the single-GPU guard, and the algorithmic byte
count behind its roofline, checked against a launch-by-launch count of what
the THERMAL row pass reads and writes (the heat-slab map floor(s nzh / nz)
of the march, the voxel arrays of every slab a launch touches, the heat
source read-modify-write except on a slab's first optical slice)."""
import math
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


class _P:
    def __init__(self, n, nz, nzh):
        self.nx = self.ny = n
        self.nz = nz
        self.nz_heat = nzh


def _count(nx, ny, nz, nzh):
    """Bytes of the nz + 1 row-pass launches, launch by launch."""
    def slab(s):
        return (s * nzh) // nz
    total = 0
    for g in range(nz + 1):
        b = 32  # P and S, read and write, FP32 complex
        first = slab(max(g - 1, 0))
        slabs = {first}
        if 1 <= g < nz and slab(g) != slab(g - 1):
            slabs.add(slab(g))
        b += 24 * len(slabs)  # (n_p, n_s) 8 + dk-base 8 + dk 8 per slab touched
        if g >= 1:  # heat source of slice g-1
            new_slab = g - 1 == 0 or slab(g - 2) != slab(g - 1)
            b += 4 if new_slab else 8
        total += b
    return total * nx * ny


@pytest.mark.parametrize("nz,nzh", [(4000, 400), (1000, 100), (200, 20), (37, 5), (10, 10), (100, 7)])
def test_rowpass_thermal_bytes_match_a_launch_count(nz, nzh):
    p = _P(64, nz, nzh)
    assert bench.rowpass_thermal_march_bytes(p) == _count(64, 64, nz, nzh)


def test_rowpass_thermal_bytes_per_point_of_the_bench_workload():
    # 4000 z-steps over 400 heat slices: 65.99 bytes per grid point and
    # launch; 65.83 at the 200-step, 20-slice cut the counter runs use
    per = bench.rowpass_thermal_march_bytes(_P(512, 4000, 400)) / (512 * 512 * 4001)
    assert math.isclose(per, 65.99, abs_tol=0.01)
    per = bench.rowpass_thermal_march_bytes(_P(512, 200, 20)) / (512 * 512 * 201)
    assert math.isclose(per, 65.83, abs_tol=0.01)


class _Q:
    def __init__(self, n, nz, precision, thermal=0, nzh=0):
        self.nx = self.ny = n
        self.nz = nz
        self.nz_heat = nzh
        self.precision = precision
        self.thermal = thermal


class _R:
    def __init__(self, steps, outer):
        self.steps_done = steps
        self.outer_iters = outer


def test_column_pass_roofline_counts_every_launch():
    """pass_roofline for the FP64-arithmetic column pass of the split form:
    32 B per grid point and z-step over the instances of its size and form,
    divided by the kernel's time; other sizes and forms are not counted."""
    from cushg_amd import abi
    ps = [_Q(512, 100, abi.PREC_FP32_SPLIT), _Q(512, 60, abi.PREC_FP32_STORAGE), _Q(256, 100, abi.PREC_FP32_SPLIT),
          _Q(512, 100, abi.PREC_FP32_MIXED)]
    rs = [_R(100, 1), _R(60, 1), _R(100, 1), _R(100, 1)]
    ms, n = 2.0, 100
    r = bench.pass_roofline("colpass_512_f64", ms, n, ps, rs)
    alg = 32.0 * 512 * 512 * (100 + 60)
    assert math.isclose(r["achieved"], round(alg / (ms * 1e-3) / 1e9, 1))
    assert math.isclose(r["frac"], round(alg / (ms * 1e-3) / 1e9 / bench.HBM_PEAK_GBS, 4))
    assert r["frac_at_minimum_bytes"] == r["frac"]  # the column pass moves exactly its minimum
    assert math.isclose(r["algorithmic_bytes_per_launch"], round(alg / n, 1))
    assert "32 B/pt" in r["algorithmic_bytes"]
