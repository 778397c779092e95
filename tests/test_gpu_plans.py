"""Launch plans, the autotuner and graph replay of the fused march.

A plan (rows per row-pass workgroup or the wide one-point-per-thread form,
columns per column-pass workgroup) only regroups independent rows /
columns, and graph replay only changes how the same launches are issued, so
every plan and every graph chunk must give bit-identical fields and eta to
the direct launches of the built-in plan -- for uneven batches (different
grids and slice counts), THERMAL runs and every fused form.  (The library
contracts multiply-adds into FMAs per source expression only, so the wide
and the standard row pass round alike.)
"""
import numpy as np
import pytest

from cushg_amd import abi, configs
from cushg_amd import shg as S

FUSED = [abi.PREC_FP32_MIXED, abi.PREC_FP32_STORAGE, abi.PREC_FP64_FUSED, abi.PREC_FP32_SPLIT]


def uneven_batch(precision, thermal=False):
    ps = []
    for i, (n, nz) in enumerate([(64, 40), (128, 33), (64, 17), (192, 25)]):
        p = configs.scaled(configs.spplt64(), n=n, nz=nz)
        p.precision = precision
        p.power_W = 5.0 + i
        p.waist_um = 20.0 + 3 * i
        if thermal:
            p.thermal = 1
            p.heat_solver = abi.HEAT_DIRECT
            p.nz_heat = 8
            p.max_outer = 3
            p.alpha_p, p.alpha_s = 2e-5, 2e-4
        ps.append(p)
    return ps


def run_all(handle, ps):
    rc, rs, outs = handle.run_batch(ps, fields=True)
    assert rc == 0, S.last_error()
    return rs, outs


def same(a, b):
    ra, oa = a
    rb, ob = b
    for x, y in zip(ra, rb):
        assert x.eta == y.eta and x.P_pump_out_W == y.P_pump_out_W and x.steps_done == y.steps_done
    for x, y in zip(oa, ob):
        assert np.array_equal(x["pump"], y["pump"]) and np.array_equal(x["signal"], y["signal"])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", FUSED)
@pytest.mark.parametrize("thermal", [False, True], ids=["optics", "thermal"])
def test_every_plan_gives_identical_results(precision, thermal):
    h = S.Handle(0)
    try:
        ps = uneven_batch(precision, thermal)
        print(f"case plans precision={precision} thermal={thermal}", flush=True)
        # the built-in plan (wide row pass at these lengths)
        default = run_all(h, ps)
        for n in (64, 128, 192):
            assert h.get_plan(n, precision)[0] == abi.PLAN_ROWS_WIDE
            for rows in (1, 2, 4, 8, abi.PLAN_ROWS_WIDE):
                for cols in (2, 4, 8, 16):
                    h.set_plan(n, rows, cols, precision)
                    assert h.get_plan(n, precision) == (rows, cols)
        for n in (64, 128, 192):
            h.set_plan(n, abi.PLAN_ROWS_WIDE, 2, precision)
        same(run_all(h, ps), default)
        for n in (64, 128, 192):
            h.set_plan(n, 1, 8, precision)
        ref = run_all(h, ps)
        same(default, ref)
        for rows, cols in ((1, 2), (8, 4), (2, 16), (4, 8)):
            for n in (64, 128, 192):
                h.set_plan(n, rows, cols, precision)
            same(run_all(h, ps), ref)
    finally:
        h.close()


@pytest.mark.gpu
def test_autotune_picks_a_supported_plan_and_keeps_results():
    h = S.Handle(0)
    try:
        for n, prec in ((128, abi.PREC_FP32_STORAGE), (256, abi.PREC_FP32_MIXED), (128, abi.PREC_FP32_SPLIT)):
            print(f"case autotune n={n} precision={prec}", flush=True)
            p = configs.scaled(configs.spplt64(), n=n, nz=24)
            p.precision = prec
            pt = configs.scaled(configs.spplt256_thermal(), n=n, nz=24, nz_heat=6)
            pt.precision = prec
            ref, reft = h.run(p), h.run(pt, temperature=True)
            rows, cols = h.autotune(n, prec, batch=2, nz=16)
            print(f"  tuned plan rows={rows} cols={cols}")
            assert rows in (1, 2, 4, 8, abi.PLAN_ROWS_WIDE) and cols in (2, 4, 8, 16)
            assert h.get_plan(n, prec) == (rows, cols)
            # every plan, tuned for optics or for THERMAL launches, gives the same bits
            r, out = h.run(p)
            same(([r], [out]), ([ref[0]], [ref[1]]))
            rt, outt = h.run(pt, temperature=True)
            same(([rt], [outt]), ([reft[0]], [reft[1]]))
        with pytest.raises(S.ShgError):
            h.autotune(100, abi.PREC_FP32_MIXED)
        with pytest.raises(S.ShgError):
            h.autotune(128, abi.PREC_FP64_PLAIN)
    finally:
        h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", FUSED)
@pytest.mark.parametrize("thermal", [False, True], ids=["optics", "thermal"])
def test_graph_replay_is_bit_identical(precision, thermal):
    h = S.Handle(0)
    try:
        ps = uneven_batch(precision, thermal)
        print(f"case graphs precision={precision} thermal={thermal}", flush=True)
        ref = run_all(h, ps)
        for steps in (1, 7, 32, 64):
            h.set_graph_steps(steps)
            same(run_all(h, ps), ref)
            same(run_all(h, ps), ref)  # the cached graph
        # a different batch on the same handle, then the first again
        other = uneven_batch(precision, thermal)[1:]
        h.set_graph_steps(0)
        ref2 = run_all(h, other)
        h.set_graph_steps(16)
        same(run_all(h, other), ref2)
        same(run_all(h, ps), ref)
        # a single run and the traced path (which launches directly)
        r0, o0 = h.run(ps[0])
        seen = []
        r1, o1 = h.run_traced(ps[0], 5, lambda s, k, a, b: seen.append(s), fields=True)
        assert r1.eta == r0.eta and np.array_equal(o1["signal"], o0["signal"]) and seen
    finally:
        h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [abi.PREC_FP32_SPLIT, abi.PREC_FP32_MIXED, abi.PREC_FP32_STORAGE])
@pytest.mark.parametrize("n", [256, 384, 512])
def test_thermal_row_pass_load_orders_agree(precision, n):
    """The THERMAL row pass reads a point's heat-grid values (n_p, n_s, the dk
    integral base and slope, the heat source) at different times in its two
    forms: the FP32-arithmetic one-row form one point ahead of the
    arithmetic (from 192 points), the wide form with the field loads, before
    the row IFFT (from 512 points in FP32 arithmetic, 384 in FP64), and both in
    place otherwise.  Every order must give the same bits, including at a heat
    slab boundary inside the march and on the first slice of a slab (where the
    heat source read ahead is overwritten unread)."""
    h = S.Handle(0)
    try:
        ps = []
        for i in range(2):
            p = configs.scaled(configs.ppln512_thermal(), n=n, nz=14 + 3 * i, nz_heat=4 + i)
            p.precision = precision
            p.max_outer = 3
            p.waist_um += 4.0 * i
            ps.append(p)
        print(f"case row pass load orders n={n} precision={precision}", flush=True)
        h.set_plan(n, abi.PLAN_ROWS_WIDE, 0, precision)
        wide = run_all(h, ps)
        assert all(r.peak_dT_K > 0 for r in wide[0])
        for rows in (1, 2):
            h.set_plan(n, rows, 0, precision)
            same(run_all(h, ps), wide)
    finally:
        h.close()
