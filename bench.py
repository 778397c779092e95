#!/usr/bin/env python3
"""Benchmark of the SSFM propagation path (one MI355X).

Metric: grid-point * z-steps per second (sum over instances of nx*ny*steps
divided by wall time), FP32-mixed form, inputs resident on the device (the
pump is initialised on the GPU; nothing crosses PCIe in the timed region
except the scalar results).  A "step" of this benchmark is one complete
run of the workload (all thermo-optic passes and heat solves included).

Default workload (THERMAL, 256x256, 1000 z-steps, 50 mm, 10 W, 40 um,
heat grid 256x256x100, oven 40 C): see cushg_amd.configs.config2.

Prints one JSON line.  Extra keys: "roofline" of the dominant kernel
(algorithmic bytes per launch / HIP-event time of the launch on the
library's stream), "cpu_baseline" (the FP64 oracle on a bounded sample of
the same workload, host cores), "fp64_plain" (the same workload in the
plain FP64 GPU form) and the accuracy of the run against the oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# algorithmic HBM bytes per grid point of one launch
BYTES_PER_PT = {
    "rowpass": 32.0,          # read P,S + write P,S (FP32 complex)
    "rowpass_thermal": 60.0,  # + T (4) + dk, dk-base (16) + Q read/write (8)
    "colpass": 32.0,          # both fields: read + write
}


def workload(name: str):
    from cushg_amd import abi, configs
    if name == "cfg2":
        return [configs.config2()], "THERMAL 256x256, 1000 z-steps, 50 mm sPPLT, 10 W, 40 um, heat 256x256x100, oven 40 C"
    if name == "cfg2-optics":
        p = configs.config2()
        p.thermal = 0
        return [p], "256x256, 1000 z-steps, 50 mm sPPLT, 10 W, 40 um, THERMAL off"
    if name == "cfg5":
        return [configs.config5()], "THERMAL 512x512, 4000 z-steps, 50 mm PPLN, 30 W, 50 um, heat 512x512x400"
    if name == "cfg5-optics":
        p = configs.config5()
        p.thermal = 0
        return [p], "512x512, 4000 z-steps, 50 mm PPLN, 30 W, 50 um, THERMAL off"
    if name == "cfg3":
        return configs.config3(seed=0, n=64), "64-instance sweep, grids 128..512, 500-2000 z-steps"
    if name == "cfg1":
        return [configs.config1()], "64x64, 100 z-steps, 30 mm sPPLT, 1 W, 30 um"
    raise SystemExit(f"unknown workload {name}")


def pts_steps(ps, results):
    return float(sum(p.nx * p.ny * r.steps_done for p, r in zip(ps, results)))


def run_once(h, ps):
    rc, res, _ = h.run_batch(ps)
    if rc != 0:
        from cushg_amd import shg
        raise RuntimeError(f"run failed: {rc} {shg.last_error()}")
    return res


def kernel_roofline(h, ps):
    """Dominant kernel's achieved algorithmic bandwidth from HIP events."""
    h.set_profiling(True)
    res = run_once(h, ps)
    kt = h.kernel_times()
    h.set_profiling(False)
    if not kt:
        return None, kt
    name, (ms, n) = max(kt.items(), key=lambda kv: kv[1][0])
    base = "rowpass_thermal" if name.endswith("_thermal") else name.split("_")[0]
    bpp = BYTES_PER_PT.get(base)
    if bpp is None:
        return {"kernel": name, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None}, kt
    npts = sum(p.nx * p.ny for p in ps)
    bytes_per_launch = bpp * npts
    avg_ms = ms / n
    ach = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": round(avg_ms * 1e3, 3), "launches": n}, kt


def cpu_baseline(ps, budget_s: float):
    """FP64 oracle (OpenMP, all host threads) on a bounded sample of the
    workload: the first instance with its z-steps cut to fit ~budget_s."""
    from oracle import oracle
    p = ps[0].copy()
    threads = oracle.get_threads()
    # calibrate on a short march
    q = p.copy()
    q.nz = 8
    if q.thermal:
        q.nz_heat = min(q.nz_heat if q.nz_heat > 0 else q.nz, 8)
    t = time.perf_counter()
    r, _ = oracle.run(q, fields=False)
    dt = max(time.perf_counter() - t, 1e-3)
    per = dt / max(r.steps_done, 1)
    nz = int(max(8, min(p.nz, budget_s / per / max(r.outer_iters, 1))))
    q = p.copy()
    q.nz = nz
    if q.thermal:
        q.nz_heat = max(2, min(q.nz_heat if q.nz_heat > 0 else q.nz, nz))
    t = time.perf_counter()
    r, _ = oracle.run(q, fields=False)
    dt = time.perf_counter() - t
    return {"value": round(q.nx * q.ny * r.steps_done / dt, 1), "unit": "grid-pt*z-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{q.nx}x{q.ny}, {r.steps_done} z-steps ({r.outer_iters} thermo-optic passes), "
                      f"nz_heat={q.nz_heat if q.thermal else 0}, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the batch shards by "
              f"instance across GPUs by design, but no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)

    from cushg_amd import abi, shg

    ps, desc = workload(args.workload)
    h = shg.Handle(0)
    for _ in range(args.warmup):
        res = run_once(h, ps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = run_once(h, ps)
    dt = time.perf_counter() - t0
    work = pts_steps(ps, res)
    value = work * args.steps / dt

    roof, kt = kernel_roofline(h, ps)

    extra = {}
    if not args.no_fp64 and len(ps) == 1:
        q = ps[0].copy()
        q.precision = abi.PREC_FP64_PLAIN
        try:
            r64 = run_once(h, [q])
            t = time.perf_counter()
            r64 = run_once(h, [q])
            d64 = time.perf_counter() - t
            v64 = pts_steps([q], r64) / d64
            extra["fp64_plain"] = {"value": round(v64, 1), "ms_per_step": round(d64 * 1e3, 3),
                                   "eta": r64[0].eta}
            extra["speedup_vs_fp64_plain"] = round(value / v64, 2)
        except Exception as e:  # plain form unavailable: say so
            extra["fp64_plain"] = {"error": str(e)}

    line = {
        "metric": "grid-pt*z-steps/s (FP32-mixed form)",
        "value": round(value, 1),
        "unit": "grid-pt*z-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fields/FFT/RK4) + fp64 (phase, reductions, heat)",
        "data": "synthetic (Gaussian pump from parameters)",
        "config": {"workload": f"{args.workload}: {desc}", "instances": len(ps),
                   "z_steps_per_run": int(sum(r.steps_done for r in res)),
                   "eta": res[0].eta, "peak_dT_K": res[0].peak_dT_K, "outer_iters": res[0].outer_iters},
        "roofline": roof,
    }
    line.update(extra)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ps, args.cpu_budget)
    line["kernel_ms"] = {k: round(v[0], 3) for k, v in kt.items()}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
