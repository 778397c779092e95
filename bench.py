#!/usr/bin/env python3
"""Benchmark of the SSFM propagation path on one MI355X.

Metric: grid-point * z-steps per second = sum over instances of
nx*ny*(z-steps executed) / wall time of the fused, batched FP32-field form,
everything resident on the device (the pump is initialised on the GPU;
only scalar results cross PCIe in the timed region).  One benchmark "step"
is one complete run of the workload: every thermo-optic pass, heat solve
and z-march of every instance.  The timed form is SHG_PREC_FP32_SPLIT by
default (FP32 fields in HBM; the row pass in FP32 arithmetic, the column
pass in FP64 arithmetic); SHG_PREC_FP32_STORAGE (FP64 arithmetic in both
passes) is timed beside it.

Default workload: 16 THERMAL MgO:PPLN instances -- 512x512 grid, 4000
z-steps, 50 mm crystal, 30 W, heat grid 512x512x400 -- over seeded
(waist, focus, oven temperature) triples
(cushg_amd.configs.ppln512_thermal_batch).

The JSON line also carries:
  roofline        dominant kernel: algorithmic bytes per launch / average
                  launch time from HIP events that each launch's dispatch
                  records (hipExtLaunchKernel) on the library's stream;
                  `traffic` = HBM bytes per launch from the committed
                  rocprofv3 counter summary (profiles/pmc_traffic.json).
  roofline_other_pass  the same for the other fused pass (row or column)
                  of the same profiled run.
  fp64_plain      the plain FP64 GPU form (one instance at a time) over
                  every instance: its throughput, and
  eta_rel_err_vs_fp64  the eta error of the timed form against it (and
                  P_sh_rel_err_vs_fp64_max, peak_dT_rel_err_vs_fp64: the
                  SH power and peak temperature rise errors).
  other_fused_form     the other FP32-field form's value and eta error.
  adjoint         forward + adjoint (shg_adjoint) wall time over one forward
                  at 256x256 x 2000 slices.
  adjoint_thermal the same for the THERMAL 512x512x4000 case, and one
                  eta-maximisation step over 16 (waist, focus, oven) triples.
  cpu_baseline    the FP64 oracle (OpenMP, all host threads) on a bounded
                  sample of the first instance.
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
METRIC = "grid-pt\u00b7z-steps/s, FP32 batched vs FP64 plain, at \u03b7 rel-err \u22641e-5; %HBM peak"
UNIT = "grid-pt\u00b7z-steps/s"

# algorithmic HBM bytes per grid point of one launch (FP32 complex fields)
BYTES_PER_PT = {
    "rowpass": 32.0,  # read P,S + write P,S
    "colpass": 32.0,  # both fields: read + write (the propagator factors are O(nx + ny))
}


def rowpass_thermal_march_bytes(p) -> float:
    """Algorithmic HBM bytes of the nz + 1 THERMAL row-pass launches of one
    z-march of instance p: every launch reads and writes P,S (32 B/pt) and
    reads the (n_p, n_s), dk-base and dk values of its heat slab (8 + 8 + 8
    B/pt; launches that straddle two slabs read both); the nz launches that
    deposit the heat source read and write Q (8 B/pt), except on the first
    optical slice of a slab, which only writes it (4 B/pt)."""
    nzh = p.nz_heat if p.nz_heat > 0 else p.nz
    pts = p.nx * p.ny
    per = (p.nz + 1) * (32 + 24) + (nzh - 1) * 24 + p.nz * 8 - nzh * 4
    return float(per * pts)


def adjoint_cost(h) -> dict:
    """Cost of d eta / d(waist, focus, T, poling[2000]) by shg_adjoint
    relative to one forward run: the 256x256, 2000-slice poled MgO:sPPLT
    case (cushg_amd.configs.adjoint256), FP32 fields / FP64 arithmetic,
    checkpoints every 45 slices; wall times of the second of two calls."""
    from cushg_amd import abi, configs
    p = configs.adjoint256()
    p.precision = abi.PREC_FP32_STORAGE
    h.run(p, fields=False)
    t = time.perf_counter()
    h.run(p, fields=False)
    tf = time.perf_counter() - t
    h.adjoint(p, 45)
    t = time.perf_counter()
    h.adjoint(p, 45)
    ta = time.perf_counter() - t
    print(f"bench: adjoint {ta * 1e3:.1f} ms vs forward {tf * 1e3:.1f} ms", file=sys.stderr, flush=True)
    return {"workload": "256x256, 2000 slices, 50 mm MgO:sPPLT, seeded poling phases, checkpoints every 45",
            "forward_ms": round(tf * 1e3, 2), "forward_plus_adjoint_ms": round(ta * 1e3, 2),
            "ratio": round(ta / tf, 3), "target": 3.0}


# value form per workload when --precision is not given: the fastest form
# whose eta has stayed within 1e-5 of the FP64 plain form on every instance
# in every version measured.  On the 64-instance sweep the split form's
# worst (far-off-phase-matching) instance has ranged from 7.2e-6 to 1.15e-5
# across versions (rounding noise amplified where the SH left is a small
# residue), the storage form's stays below 5e-6: storage is its value form,
# and the split form is reported beside it.
DEFAULT_FORM = {"sweep64": "storage"}


# handles (one host thread each) that take the ascent step's gradients side by side
GRAD_HANDLES = 4


def adjoint_thermal_cost(h, batch: int) -> dict:
    """THERMAL gradient cost at the full coupled case: one MgO:PPLN
    512x512x4000 instance with a 512x512x400 heat grid (configs.ppln512_thermal),
    FP32 fields / FP64 arithmetic -- forward, then forward + adjoint
    (shg_adjoint at the converged steady state), wall times of the second of
    two calls; then, with batch > 0, one eta-maximisation step
    (cushg_amd.optimize.ascent_step) over that many seeded (waist, focus,
    oven) triples of the same case."""
    from cushg_amd import abi, configs, optimize, shg
    p = configs.ppln512_thermal()
    p.precision = abi.PREC_FP32_STORAGE
    h.run(p, fields=False)
    t = time.perf_counter()
    rf, _ = h.run(p, fields=False)
    tf = time.perf_counter() - t
    h.adjoint(p)
    t = time.perf_counter()
    h.adjoint(p)
    ta = time.perf_counter() - t
    print(f"bench: THERMAL adjoint {ta * 1e3:.1f} ms vs forward {tf * 1e3:.1f} ms", file=sys.stderr, flush=True)
    out = {"workload": "512x512, 4000 slices, 50 mm MgO:PPLN, 30 W, THERMAL, heat 512x512x400, FP32 fields / FP64 "
                       "arithmetic",
           "forward_ms": round(tf * 1e3, 1), "forward_outer_passes": rf.outer_iters,
           "forward_plus_adjoint_ms": round(ta * 1e3, 1), "ratio": round(ta / tf, 3),
           "adjoint_only_over_forward": round((ta - tf) / tf, 3), "target": 3.0}
    if batch > 0:
        ps = configs.ppln512_thermal_batch(batch)
        for q in ps:
            q.precision = abi.PREC_FP32_STORAGE
        st1 = optimize.ascent_step(h, ps)
        extra = [shg.Handle(0) for _ in range(GRAD_HANDLES - 1)]
        try:
            st = optimize.ascent_step(h, ps, extra_handles=extra)
        finally:
            for e in extra:
                e.close()
        t = time.perf_counter()
        for q in ps:
            h.run(q, fields=False)
        ts = time.perf_counter() - t
        rc, rb, _ = h.run_batch(ps)
        t = time.perf_counter()
        rc, rb, _ = h.run_batch(ps)
        tb = time.perf_counter() - t
        gains = [s.eta1 / s.eta0 - 1 for s in st.steps]
        out["ascent_step"] = {"instances": batch, "forward_plus_adjoint_s": round(st1.grad_wall_s, 2),
                              "forward_s": round(ts, 2), "ratio": round(st1.grad_wall_s / ts, 3),
                              "ratio_basis": "the instances' forward + adjoint calls over their forward calls, "
                                             "both one instance at a time on one handle",
                              "step_s_one_handle": round(st1.wall_s, 2),
                              "gradient_handles": GRAD_HANDLES,
                              "forward_plus_adjoint_s_handles": round(st.grad_wall_s, 2),
                              "step_s": round(st.wall_s, 2), "batched_forward_s": round(tb, 2),
                              "accepted": sum(s.accepted for s in st.steps),
                              "eta_gain_rel": {"min": min(gains), "max": max(gains)}}
        print(f"bench: ascent step over {batch}: {st1.wall_s:.1f} s on one handle (forward + adjoint "
              f"{st1.grad_wall_s:.1f} s, forward {ts:.1f} s, ratio {st1.grad_wall_s / ts:.2f}); {st.wall_s:.1f} s with "
              f"the gradients on {GRAD_HANDLES} handles ({st.grad_wall_s:.1f} s)", file=sys.stderr, flush=True)
    return out


def workload(name: str):
    from cushg_amd import configs
    if name == "ppln512t-b16":
        return configs.ppln512_thermal_batch(16), ("16 x THERMAL 512x512, 4000 z-steps, 50 mm MgO:PPLN, 30 W, "
                                           "heat 512x512x400, seeded waist/focus/oven")
    if name == "ppln512t":
        return [configs.ppln512_thermal()], "THERMAL 512x512, 4000 z-steps, 50 mm MgO:PPLN, 30 W, 50 um, heat 512x512x400"
    if name == "spplt256t":
        return [configs.spplt256_thermal()], "THERMAL 256x256, 1000 z-steps, 50 mm sPPLT, 10 W, 40 um, heat 256x256x100, oven 40 C"
    if name == "spplt256":
        p = configs.spplt256_thermal()
        p.thermal = 0
        return [p], "256x256, 1000 z-steps, 50 mm sPPLT, 10 W, 40 um, THERMAL off"
    if name == "sweep64":
        return configs.sweep_batch(seed=0, n=64), "64-instance sweep, grids 128..512, 500-2000 z-steps, THERMAL off"
    if name == "spplt64":
        return [configs.spplt64()], "64x64, 100 z-steps, 30 mm sPPLT, 1 W, 30 um"
    raise SystemExit(f"unknown workload {name}")


def pts_steps(ps, results):
    return float(sum(p.nx * p.ny * r.steps_done for p, r in zip(ps, results)))


def run_once(h, ps):
    from cushg_amd import shg
    rc, res, _ = h.run_batch(ps)
    if rc != 0:
        raise RuntimeError(f"run failed: {rc} {shg.last_error()}")
    return res


def committed_traffic(kernel: str):
    """Measured HBM bytes / algorithmic bytes of `kernel` from the committed
    rocprofv3 counter summary (profiles/pmc_traffic.json)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("kernels", {}).get(kernel)
    except (OSError, ValueError):
        return None


def kernel_roofline(h, ps):
    """Dominant kernel's achieved algorithmic bandwidth: its algorithmic
    bytes summed over every launch of a profiled run (each launch covers the
    instances still marching in it), divided by the kernel's total time from
    HIP events recorded on the library's stream.  Also returns the same
    figures for the other fused pass of the run (row or column) and the
    per-kernel times."""
    h.set_profiling(True)
    res = run_once(h, ps)
    kt = h.kernel_times()
    h.set_profiling(False)
    if not kt:
        return None, None, kt
    name, (ms, n) = max(kt.items(), key=lambda kv: kv[1][0])
    other = None
    fused = [k for k in kt if k.split("_")[0] in BYTES_PER_PT and k != name]
    if fused:
        o = max(fused, key=lambda k: kt[k][0])
        other = pass_roofline(o, kt[o][0], kt[o][1], ps, res)
    return pass_roofline(name, ms, n, ps, res), other, kt


def pass_roofline(name, ms, n, ps, res):
    """Roofline object of one fused-pass kernel (profiling name, total ms,
    launches) over the profiled run's instances and results."""
    from cushg_amd import abi
    parts = name.split("_")
    base, size = parts[0], int(parts[1])
    thermal = "_thermal" in name
    f64 = name.endswith("_f64")
    # arithmetic of each pass per form: FP64 in both (storage), FP64 in the
    # column pass only (split), FP32 in both (mixed)
    f64_forms = {"rowpass": (abi.PREC_FP32_STORAGE,),
                 "colpass": (abi.PREC_FP32_STORAGE, abi.PREC_FP32_SPLIT)}.get(base, ())
    if base not in BYTES_PER_PT:
        return {"kernel": name, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None}
    total = 0.0
    # the same launches priced at the minimum of the two-pass scheme: 32 B
    # per point and pass for the fields, plus 24 B per point and z-step for
    # THERMAL (T read, Q write, the FP64 running dk integral read + write),
    # all of it charged to the row pass that does the thermal work
    survey = 0.0
    for p, r in zip(ps, res):
        if (p.precision in f64_forms) != f64:
            continue
        pts = p.nx * p.ny
        if base == "rowpass" and p.nx == size and bool(p.thermal) == thermal:
            if thermal:
                total += r.outer_iters * rowpass_thermal_march_bytes(p)
                survey += r.steps_done * pts * (32.0 + 24.0)
            else:  # nz + 1 row passes per march
                total += (r.steps_done + r.outer_iters) * pts * BYTES_PER_PT["rowpass"]
                survey += r.steps_done * pts * 32.0
        elif base == "colpass" and p.ny == size:
            total += r.steps_done * pts * BYTES_PER_PT["colpass"]
            survey += r.steps_done * pts * 32.0
    ach = total / (ms * 1e-3) / 1e9
    ach_s = survey / (ms * 1e-3) / 1e9
    tr = committed_traffic(name)
    alg = total / n
    traffic = round(alg * tr["traffic_over_algorithmic"], 1) if tr else None
    layout = ("the layout's: fields 32 B/pt; THERMAL (n_p, n_s) float2, dk-base and dk doubles read per launch and Q "
              "read-modify-written (about 66 B/pt per launch)") if base == "rowpass" and thermal else \
        "the layout's: fields 32 B/pt (both fields read and written once)"
    return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": ("not measured in this run: this run's algorithmic bytes per launch x the measured "
                               "traffic / algorithmic ratio committed in profiles/pmc_traffic.json (rocprofv3 TCC "
                               "request counts and WRITE_SIZE, on a 200-step, 20-slab cut of the workload)")
                              if tr else None,
            "algorithmic_bytes_per_launch": round(total / n, 1),
            "algorithmic_bytes": layout,
            "achieved_at_minimum_bytes": round(ach_s, 1), "frac_at_minimum_bytes": round(ach_s / HBM_PEAK_GBS, 4),
            "minimum_bytes_per_launch": round(survey / n, 1),
            "minimum_bytes": "32 B/pt per pass, + 24 B/pt per z-step for THERMAL (T 4, Q 4, FP64 dk integral 16)",
            "avg_launch_us": round(ms / n * 1e3, 3),
            "launches": n, "timer": "HIP events recorded by each launch's own dispatch (hipExtLaunchKernel) on the "
                                    "library stream"}


def cpu_baseline(ps, budget_s: float):
    """FP64 oracle (OpenMP, all host threads) on a bounded sample of the
    workload: the first instance at its full transverse grid, one
    thermo-optic pass (heat solve + z-march), the heat grid cut to at most
    16 slices and the z-steps cut so that the run takes about budget_s."""
    from oracle import oracle
    p = ps[0].copy()
    p.max_outer = 1
    if p.thermal:
        p.nz_heat = min(p.nz_heat if p.nz_heat > 0 else p.nz, 16)
    threads = oracle.get_threads()

    def timed(nz):
        q = p.copy()
        q.nz = nz
        t = time.perf_counter()
        r, _ = oracle.run(q, fields=False)
        return q, r, time.perf_counter() - t

    # fixed cost (heat solve, set-up) and cost per z-step from two short runs
    _, _, d1 = timed(16)
    _, _, d2 = timed(48)
    per = max((d2 - d1) / 32, 1e-4)
    fixed = max(d1 - 16 * per, 0.0)
    nz = int(max(16, min(p.nz, (budget_s - fixed) / per)))
    q, r, dt = timed(nz)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(q.nx * q.ny * r.steps_done / dt, 1), "unit": UNIT,
            "cores": threads, "kind": "port", "cpu_model": model, "host_threads": os.cpu_count(),
            "sample": f"first instance, {q.nx}x{q.ny}, one thermo-optic pass of {r.steps_done} z-steps"
                      + (f" with a {q.nx}x{q.ny}x{q.nz_heat} heat solve" if q.thermal else "")
                      + f", {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ppln512t-b16")
    ap.add_argument("--nz", type=int, default=0, help="override z-steps, scaling the heat grid's z-slices with it (profiling runs)")
    ap.add_argument("--max-outer", type=int, default=0, help="override thermo-optic passes (profiling runs)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--precision", choices=("split", "storage", "mixed"), default=None,
                    help="fused form timed as the value: FP32 fields with the row pass in FP32 and the column "
                         "pass in FP64 arithmetic (split), FP64 arithmetic in both (storage) or FP32 in both "
                         "(mixed); storage is reported beside split, split beside the others.  Default: the "
                         "fastest form that has held every instance of the workload within the 1e-5 eta bar "
                         "in every version (storage for sweep64, whose far-off-phase-matching instances the split "
                         "form's FP32 row transforms have taken to between 7.2e-6 and 1.15e-5; split otherwise)")
    ap.add_argument("--no-alt", action="store_true", help="skip the other fused form")
    ap.add_argument("--autotune", action="store_true",
                    help="before timing, run shg_autotune for every grid size of the workload (batch = the "
                         "workload's instances of that size); the tuning time is not in the timed region")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint cost lines (256x256, 2000 slices; and THERMAL 512x512x4000)")
    ap.add_argument("--ascent-batch", type=int, default=16,
                    help="instances of the THERMAL eta-maximisation step timed in the adjoint line (0: none)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the THERMAL adjoint cost of one 512x512x4000 instance (profiling runs): prints its "
                         "JSON object and exits")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many MiB (PMC counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        # known-size streams for calibrating HBM byte counters: a runtime
        # buffer copy and an elementwise kernel, each reading and writing
        # copy_probe MiB per launch, three launches each
        import torch
        n = args.copy_probe * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        for _ in range(3):
            torch.mul(x, 2.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_MiB": args.copy_probe, "launches": {"copyBuffer": 3, "elementwise": 3}}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (asked for {args.gpus}); a batch shards by instance "
              f"across GPUs by design, but no multi-GPU figure is measured here.", file=sys.stderr)
        sys.exit(2)

    from cushg_amd import abi, shg

    if args.adjoint_only:
        with shg.Handle(0) as h:
            print(json.dumps(adjoint_thermal_cost(h, 0)))
        return
    ps, desc = workload(args.workload)
    forms = {"split": abi.PREC_FP32_SPLIT, "storage": abi.PREC_FP32_STORAGE, "mixed": abi.PREC_FP32_MIXED}
    names = {v: k for k, v in forms.items()}
    if args.precision is None:
        args.precision = DEFAULT_FORM.get(args.workload, "split")
    prec = forms[args.precision]
    alt_prec = abi.PREC_FP32_STORAGE if prec == abi.PREC_FP32_SPLIT else abi.PREC_FP32_SPLIT
    for p in ps:
        p.precision = prec
        if args.nz:
            if p.nz_heat > 0:  # keep the optical slices per heat slab
                p.nz_heat = min(args.nz, max(2, round(p.nz_heat * args.nz / p.nz)))
            p.nz = args.nz
        if args.max_outer:
            p.max_outer = args.max_outer
    print(f"bench: {args.workload} ({len(ps)} instances) warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)
    h = shg.Handle(0)
    plans = {}
    if args.autotune:
        from collections import Counter
        t = time.perf_counter()
        for n, cnt in sorted(Counter(p.nx for p in ps).items()):
            plans[n] = h.autotune(n, prec, batch=cnt, nz=64)
        print(f"bench: autotuned {plans} in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        res = run_once(h, ps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = run_once(h, ps)
    dt = time.perf_counter() - t0
    work = pts_steps(ps, res)
    value = work * args.steps / dt
    print(f"bench: timed {dt:.2f} s", file=sys.stderr, flush=True)

    roof, roof2, kt = (None, None, {}) if args.no_roofline else kernel_roofline(h, ps)

    extra = {}
    alt = None
    if not args.no_alt:
        qs = [p.copy() for p in ps]
        for q in qs:
            q.precision = alt_prec
        run_once(h, qs)
        t = time.perf_counter()
        ralt = run_once(h, qs)
        dalt = time.perf_counter() - t
        alt = {"precision": names[alt_prec],
               "value": round(pts_steps(qs, ralt) / dalt, 1), "unit": UNIT, "ms_per_step": round(dalt * 1e3, 3)}
        print(f"bench: other fused form {dalt:.2f} s", file=sys.stderr, flush=True)
    if not args.no_fp64:
        # the plain FP64 form on every instance (it runs them one at a time):
        # its throughput, and the eta error of the fused forms against it
        qs = [p.copy() for p in ps]
        for q in qs:
            q.precision = abi.PREC_FP64_PLAIN
        run_once(h, qs[:1])
        d64, r64 = 0.0, []
        for q in qs:
            t = time.perf_counter()
            r64 += run_once(h, [q])
            d64 += time.perf_counter() - t
        v64 = pts_steps(qs, r64) / d64
        errs = [abs(a.eta - b.eta) / abs(b.eta) for a, b in zip(res, r64)]
        extra["fp64_plain"] = {"value": round(v64, 1), "unit": UNIT,
                               "sample": f"all {len(qs)} instances, one at a time", "s_total": round(d64, 2),
                               "eta": [round(r.eta, 10) for r in r64[:4]]}
        extra["eta_rel_err_vs_fp64"] = {"max": max(errs), "mean": sum(errs) / len(errs), "bar": 1e-5,
                                        "instances": len(errs)}
        # P_SH = eta * P carries eta's relative error; the peak temperature
        # rise (THERMAL) against the FP64 form's, relative, over all instances
        dts = [abs(a.peak_dT_K - b.peak_dT_K) / abs(b.peak_dT_K) for a, b in zip(res, r64) if b.peak_dT_K]
        if dts:
            extra["peak_dT_rel_err_vs_fp64"] = {"max": max(dts), "mean": sum(dts) / len(dts)}
        pso = [abs(a.P_sh_W - b.P_sh_W) / abs(b.P_sh_W) for a, b in zip(res, r64) if b.P_sh_W]
        extra["P_sh_rel_err_vs_fp64_max"] = max(pso) if pso else None
        extra["speedup_vs_fp64_plain"] = round(value / v64, 2)
        if alt is not None:
            ae = [abs(a.eta - b.eta) / abs(b.eta) for a, b in zip(ralt, r64)]
            alt["eta_rel_err_vs_fp64_max"] = max(ae)
            adt = [abs(a.peak_dT_K - b.peak_dT_K) / abs(b.peak_dT_K) for a, b in zip(ralt, r64) if b.peak_dT_K]
            if adt:
                alt["peak_dT_rel_err_vs_fp64_max"] = max(adt)
            alt["speedup_vs_fp64_plain"] = round(alt["value"] / v64, 2)
        print(f"bench: fp64 plain {d64:.2f} s, eta rel err max {max(errs):.2e}", file=sys.stderr, flush=True)
    if alt is not None:
        extra["other_fused_form"] = alt
    if not args.no_adjoint:
        extra["adjoint"] = adjoint_cost(h)
        extra["adjoint_thermal"] = adjoint_thermal_cost(h, args.ascent_batch)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == abi.PREC_FP32_STORAGE else "fp32",
        "dtype_detail": {
            abi.PREC_FP32_STORAGE: "fields fp32 in HBM (8 B per complex point); FFTs, propagator, phase and RK4 in "
                                   "f64 registers; reductions and heat residual f64, heat corrections fp32 GEMMs",
            abi.PREC_FP32_SPLIT: "fields fp32 in HBM (8 B per complex point); row pass (row FFTs, RK4, heat "
                                 "source) in fp32 with the dk*z phase in f64; column pass (column FFTs, "
                                 "propagator) in f64 registers; reductions and heat residual f64, heat "
                                 "corrections fp32 GEMMs",
            abi.PREC_FP32_MIXED: "fields, FFT butterflies, RK4 in fp32; twiddle/propagator products, dk*z phase, "
                                 "reductions and heat residual in f64"}[prec],
        "data": "synthetic (Gaussian pump from parameters)",
        "config": {"workload": f"{args.workload}: {desc}", "instances": len(ps),
                   "plans": {str(n): list(v) for n, v in plans.items()} if plans else "built-in",
                   "form": "FP32 batched, " + {
                       abi.PREC_FP32_STORAGE: "FP32 fields / FP64 arithmetic (SHG_PREC_FP32_STORAGE)",
                       abi.PREC_FP32_SPLIT: "FP32 fields / FP32 row pass / FP64 column pass (SHG_PREC_FP32_SPLIT)",
                       abi.PREC_FP32_MIXED: "FP32 arithmetic (SHG_PREC_FP32_MIXED)"}[prec],
                   "z_steps_per_run": int(sum(r.steps_done for r in res)),
                   "eta": [round(r.eta, 9) for r in res[:4]],
                   "peak_dT_K": [round(r.peak_dT_K, 6) for r in res[:4]],
                   "outer_iters": [r.outer_iters for r in res]},
        "roofline": roof,
    }
    if roof2:
        line["roofline_other_pass"] = roof2
    line.update(extra)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ps, args.cpu_budget)
    line["kernel_ms"] = {k: round(v[0], 3) for k, v in kt.items()}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
