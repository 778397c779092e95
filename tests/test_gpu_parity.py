"""GPU parity: the HIP library (through the C ABI) against the FP64 oracle,
for the four fused forms.

Tolerances: eta relative error <= 1e-5 against FP64 (the accuracy bar of
the FP32 batched form) for the FP32-field forms; their fields are stored in
FP32: relative L2 error <= 2e-5 with FP32 arithmetic in the row pass
(SHG_PREC_FP32_MIXED, SHG_PREC_FP32_SPLIT), <= 2e-6 with FP64 arithmetic
throughout (SHG_PREC_FP32_STORAGE, where only the stores round).  The
FP64-field form (SHG_PREC_FP64_FUSED): 1e-10 for both.
"""
import numpy as np
import pytest

from cushg_amd import abi, configs
from cushg_amd import shg as S
from oracle import oracle as O

ETA_TOL = 1e-5
FIELD_TOL = 2e-5


def rel(a, b):
    return abs(a - b) / abs(b)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


FORMS = [pytest.param(abi.PREC_FP32_MIXED, id="mixed"), pytest.param(abi.PREC_FP32_STORAGE, id="storage"),
         pytest.param(abi.PREC_FP64_FUSED, id="fp64fused"), pytest.param(abi.PREC_FP32_SPLIT, id="split")]
# (eta, field) relative tolerances per form: the FP64-field form differs
# from the oracle only in FP64 rounding of a different operation order
TOLS = {abi.PREC_FP32_MIXED: (ETA_TOL, FIELD_TOL), abi.PREC_FP32_STORAGE: (ETA_TOL, 2e-6),
        abi.PREC_FP64_FUSED: (1e-10, 1e-10), abi.PREC_FP32_SPLIT: (ETA_TOL, FIELD_TOL)}


@pytest.fixture(params=FORMS)
def form(request):
    return request.param


def _check(p, handle, form=abi.PREC_FP32_MIXED):
    p = p.copy()
    p.precision = form
    eta_tol, field_tol = TOLS[form]
    print(f"case nx={p.nx} ny={p.ny} nz={p.nz} scheme={p.scheme} thermal={p.thermal} precision={form}",
          flush=True)
    ro, o = O.run(p)
    rg, g = handle.run(p)
    assert rg.status == 0
    assert rg.steps_done >= p.nz
    e_eta = rel(rg.eta, ro.eta)
    e_p = rel_l2(g["pump"], o["pump"])
    e_s = rel_l2(g["signal"], o["signal"])
    print(f"  eta gpu={rg.eta:.12e} cpu={ro.eta:.12e} rel={e_eta:.2e}  pump {e_p:.2e}  sh {e_s:.2e}")
    assert e_eta <= eta_tol
    assert rel(rg.P_pump_in_W, ro.P_pump_in_W) <= min(1e-6, 10 * eta_tol)
    assert rel(rg.P_pump_out_W, ro.P_pump_out_W) <= min(1e-6, 10 * eta_tol)
    assert e_p <= field_tol
    assert e_s <= field_tol
    return ro, rg


@pytest.mark.gpu
def test_spplt64_parity(handle, form):
    """64x64, 100 slices, 30 mm MgO:sPPLT, 1 W, 30 um, 55.66 C."""
    _check(configs.spplt64(), handle, form)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 128, 192, 256, 384, 512])
def test_sizes_reference_defaults(handle, n, form):
    """Every transform length with the reference's absorption and 2PA on."""
    p = abi.default_params()
    p.nx = p.ny = n
    p.nz = 24
    p.lx_um = p.ly_um = 800.0
    p.power_W = 20.0
    _check(p, handle, form)


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", [(128, 64), (192, 256), (512, 128)])
def test_rectangular_grids(handle, nx, ny, form):
    p = configs.spplt64()
    p.nx, p.ny, p.nz = nx, ny, 16
    p.lx_um, p.ly_um = 1200.0, 700.0
    _check(p, handle, form)


@pytest.mark.gpu
def test_strang_scheme(handle, form):
    p = configs.spplt64()
    p.scheme = abi.SCHEME_STRANG
    p.alpha_p, p.alpha_s, p.beta_s = 0.17e-6, 1.57e-6, 5e-5
    _check(p, handle, form)


@pytest.mark.gpu
def test_walkoff_and_offcentre_focus(handle, form):
    p = configs.spplt64()
    p.rho_p, p.rho_s = 0.0, 0.01
    p.focal_um = 0.2 * p.lz_um
    p.nz = 40
    _check(p, handle, form)


@pytest.mark.gpu
def test_elliptical_pump(handle, form):
    """Astigmatic pump (waist_y_um set), off-centre focus, rectangular window."""
    p = configs.spplt64()
    p.nx, p.ny, p.nz = 128, 64, 30
    p.lx_um, p.ly_um = 900.0, 600.0
    p.waist_y_um = 1.6 * p.waist_um
    p.focal_um = 0.3 * p.lz_um
    p.power_W = 8.0
    _check(p, handle, form)


@pytest.mark.gpu
def test_elliptical_pump_plain_and_thermal(handle):
    p = configs.spplt64()
    p.waist_y_um = 0.7 * p.waist_um
    p.precision = abi.PREC_FP64_PLAIN
    ro, o = O.run(p)
    rg, g = handle.run(p)
    assert rel(rg.eta, ro.eta) < 1e-11 and rel_l2(g["signal"], o["signal"]) < 1e-11
    # waist_y_um == waist_um is the circular beam
    q = p.copy()
    q.waist_y_um = q.waist_um
    r1, _ = handle.run(q)
    q.waist_y_um = 0.0
    r0, _ = handle.run(q)
    assert rel(r1.eta, r0.eta) < 1e-12


@pytest.mark.gpu
def test_poling_phase(handle, form):
    p = configs.adjoint256(nz=60, n=64)
    _check(p, handle, form)


@pytest.mark.gpu
def test_batch_equals_sequential(handle):
    """Uneven batch (different grids, step counts and three fused forms) in
    one execution gives the same bits as running each instance alone."""
    ps = configs.sweep_batch(seed=3, n=6, scale_steps=0.02)
    forms = (abi.PREC_FP32_MIXED, abi.PREC_FP32_STORAGE, abi.PREC_FP32_SPLIT)
    for i, p in enumerate(ps):
        p.precision = forms[i % 3]
    rc, rb, ob = handle.run_batch(ps, fields=True)
    assert rc == 0
    for p, r, o in zip(ps, rb, ob):
        r1, o1 = handle.run(p)
        assert r.status == 0
        assert r.eta == r1.eta
        assert np.array_equal(o["signal"], o1["signal"])


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [-1, 0], ids=["graph", "direct"])
@pytest.mark.parametrize("thermal", [False, True], ids=["optics", "thermal"])
@pytest.mark.parametrize("prec", [abi.PREC_FP32_SPLIT, abi.PREC_FP32_STORAGE], ids=["split", "storage"])
def test_interleaved_halves_equal_sequential(handle, prec, thermal, graph):
    """Eight like 512x512 instances (2^21 grid points) march as two
    interleaved halves on two streams, one half's column pass of step g-1
    beside the other's row pass of step g, by graph replay (the default) or
    direct launches.  Step counts differ, so instances drop out of the
    launches at different steps.  Every instance gets the bits it gets alone."""
    ps = []
    for i, p in enumerate(configs.ppln512_thermal_batch(8, seed=11)):
        q = configs.scaled(p, nz=9 + 3 * i, nz_heat=4 if thermal else None)
        q.thermal = 1 if thermal else 0
        q.max_outer = 2
        q.precision = prec
        ps.append(q)
    print(f"case interleaved halves prec={prec} thermal={thermal} graph={graph}", flush=True)
    handle.set_graph_steps(graph)
    try:
        rc, rb, ob = handle.run_batch(ps, fields=True, temperature=thermal)
    finally:
        handle.set_graph_steps(-1)
    assert rc == 0
    for p, r, o in zip(ps, rb, ob):
        r1, o1 = handle.run(p, temperature=thermal)
        assert r.status == 0 and r.steps_done == r1.steps_done
        assert r.eta == r1.eta
        assert np.array_equal(o["pump"], o1["pump"])
        assert np.array_equal(o["signal"], o1["signal"])
        if thermal:
            assert r.outer_iters == r1.outer_iters
            assert np.array_equal(o["T"], o1["T"])


@pytest.mark.gpu
def test_invalid_size_is_einval(handle):
    p = configs.spplt64()
    p.nx = 100
    with pytest.raises(S.ShgError) as ei:
        handle.run(p)
    assert ei.value.status == abi.SHG_EINVAL
