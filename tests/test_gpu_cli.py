"""Per-slice trace (shg_run_traced) and the reference-compatible `shg`
program: slice states against the FP64 oracle's step-by-step march, files
in the reference's names and text format (src/headers/Files.h:69-90,
169-188), and the batched sweep against single runs."""
import os
import subprocess

import numpy as np
import pytest

from cushg_amd import abi, configs, shg
from oracle import oracle as O

SHG_EXE = os.path.join(os.path.dirname(shg.LIB_PATH), "shg")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def small(precision=abi.PREC_FP32_MIXED):
    p = configs.spplt64()
    p.nz = 24
    p.alpha_p, p.alpha_s, p.beta_s = 0.17e-6, 1.57e-6, 5e-5
    p.precision = precision
    return p


def oracle_slices(p):
    """State before the SSFM step of every slice (oracle, FP64)."""
    P = O.pump_init(p, p.T_C)
    S = np.zeros_like(P)
    out = [(P.copy(), S.copy())]
    for s in range(p.nz - 1):
        P, S = O.ssfm_step(p, s, P, S)
        out.append((P.copy(), S.copy()))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(abi.PREC_FP32_MIXED, 2e-5), (abi.PREC_FP32_STORAGE, 2e-6),
                                           (abi.PREC_FP32_SPLIT, 2e-5), (abi.PREC_FP64_PLAIN, 1e-11)])
def test_trace_slices_match_oracle_march(handle, precision, tol):
    p = small(precision)
    print(f"case trace precision={precision}", flush=True)
    got = {}
    r, out = handle.run_traced(p, 1, lambda s, ps, a, b: got.__setitem__(s, (a, b)), fields=True)
    assert sorted(got) == list(range(p.nz))
    ref = oracle_slices(p)
    for s in (0, 1, 5, p.nz - 1):
        assert rel_l2(got[s][0], ref[s][0]) < tol, s
        if s > 0:
            assert rel_l2(got[s][1], ref[s][1]) < tol, s
    assert not np.any(got[0][1])  # the SH starts at zero
    # the last traced slice is the saved output state (Pump_out/Signal_out)
    assert np.array_equal(got[p.nz - 1][0], out["pump"])
    assert np.array_equal(got[p.nz - 1][1], out["signal"])
    # the trace does not change the run
    r2, out2 = handle.run(p)
    assert r2.eta == r.eta and np.array_equal(out2["signal"], out["signal"])


@pytest.mark.gpu
def test_trace_stride_passes_and_cancel(handle):
    p = configs.scaled(configs.spplt256_thermal(), n=64, nz=40, nz_heat=10)
    p.lx_um = p.ly_um = 400.0
    p.alpha_p, p.alpha_s = 2e-6, 2e-5
    seen = []
    r, out = handle.run_traced(p, 7, lambda s, ps, a, b: seen.append((ps, s)), fields=True)
    assert r.outer_iters >= 2
    for k in range(r.outer_iters):
        assert [s for ps, s in seen if ps == k] == [0, 7, 14, 21, 28, 35]
    with pytest.raises(shg.ShgError) as e:
        handle.run_traced(small(), 1, lambda s, ps, a, b: s == 3)
    assert e.value.status == abi.SHG_EINVAL and "cancelled" in str(e.value)


def _run_cli(args, tmp):
    cmd = [SHG_EXE, *args, "--nx=64", "--ny=64", "--nz=40", "--lx=800", "--ly=800", "--lz=30000",
           f"--out={tmp}"]
    print("case cli", " ".join(cmd[1:]), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=120)


def cli_params(P, w, T, nz=40):
    p = shg.default_params()
    p.nx = p.ny = 64
    p.nz = nz
    p.lx_um = p.ly_um = 800.0
    p.lz_um = 30000.0
    p.power_W, p.waist_um, p.T_C, p.T_peltier_C = P, w, T, T
    p.focal_um = 0.5 * p.lz_um
    return p


def read_complex(d, name):
    re = np.loadtxt(os.path.join(d, name + "_r.dat"))
    im = np.loadtxt(os.path.join(d, name + "_i.dat"))
    return re + 1j * im


@pytest.mark.gpu
def test_cli_save_only_last(handle, tmp_path):
    res = _run_cli(["2", "30", "55.66", "1", "0"], tmp_path)
    assert res.returncode == 0, res.stderr
    assert "eta =" in res.stdout
    r, out = handle.run(cli_params(2.0, 30.0, 55.66))
    pump = read_complex(tmp_path, "Pump_out")
    sig = read_complex(tmp_path, "Signal_out")
    assert pump.shape == (64, 64)
    # 20 significant digits: the text holds the FP32 values exactly
    assert np.array_equal(pump, out["pump"]) and np.array_equal(sig, out["signal"])
    line = open(tmp_path / "Pump_out_r.dat").readline()
    assert line.endswith("\t\n") and line.count("\t") == 64


@pytest.mark.gpu
def test_cli_every_slice(tmp_path):
    res = _run_cli(["1", "30", "55.66", "0", "0"], tmp_path)
    assert res.returncode == 0, res.stderr
    names = sorted(f for f in os.listdir(tmp_path) if f.startswith("Pump_") and f.endswith("_r.dat"))
    assert len(names) == 40 and "Pump_out_r.dat" not in names
    assert os.path.exists(tmp_path / "Signal_39_i.dat")


@pytest.mark.gpu
def test_cli_sweep_is_batched_and_named_like_the_sweep_script(handle, tmp_path):
    res = _run_cli(["1,3", "25,35", "55.66", "1", "0"], tmp_path)
    assert res.returncode == 0, res.stderr
    for P in ("1", "3"):
        for w in ("25", "35"):
            d = tmp_path / f"sPPLT_P_{P}_W_T1_55.66_WAIST_{w}"
            assert d.is_dir()
            r, out = handle.run(cli_params(float(P), float(w), 55.66))
            assert np.array_equal(read_complex(d, "Signal_out"), out["signal"])


@pytest.mark.gpu
def test_cli_thermal_temperature_files(tmp_path):
    cmd = [SHG_EXE, "5", "40", "40", "1", "1", "--nx=64", "--ny=64", "--nz=30", "--nz-heat=6",
           "--lx=400", "--ly=400", "--lz=20000", "--thermal=1", "--heat=direct", "--alpha-p=2e-6",
           "--alpha-s=2e-5", f"--out={tmp_path}"]
    print("case cli thermal", flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("Tfinal_"))
    assert len(files) == 6
    T = np.loadtxt(tmp_path / "Tfinal_3.dat")
    assert T.shape == (64, 64) and T.min() >= 40.0 - 1e-6 and T.max() > 40.0
