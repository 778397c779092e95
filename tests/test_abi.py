"""The C-ABI boundary without a GPU: the library loads, exports every
function include/shg.h declares, its structs match the header byte for byte
(checked by compiling the header with gcc) and the Python mirror, its
defaults are the reference's compile-time constants, and parameter
validation rejects what the reference's compile-time setup cannot express.
No call here launches GPU work."""
import ctypes as C
import os
import re
import subprocess

import pytest

from cushg_amd import abi, shg
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shg.h")


def header_functions():
    text = open(HEADER).read()
    return re.findall(r"SHG_API\s+[\w\s\*]+?\b(shg_\w+)\s*\(", text)


def test_header_declares_what_the_binding_expects():
    assert sorted(header_functions()) == sorted(shg.EXPORTS)


def test_library_exports_every_declared_function():
    L = shg.lib()
    for name in header_functions():
        assert hasattr(L, name), name
    # only the declared API is visible
    out = subprocess.run(["nm", "-D", "--defined-only", shg.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln and ln.split()[-1].startswith("shg")}
    assert exported == set(header_functions())


def _c_layout(tmp_path, struct, fields):
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("{f} %zu\\n", offsetof({struct}, {f}));\n' for f in fields)
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"shg.h\"\nint main(void){\n"
                   f'  printf("sizeof %zu\\n", sizeof({struct}));\n{body}  return 0;\n}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {ln.split()[0]: int(ln.split()[1]) for ln in out.splitlines()}


def test_constants_match_header(tmp_path):
    """Every enumerator and SHG_PLAN_ROWS_WIDE as a C compiler sees them
    equal the binding's constants (status codes keep the SHG_ prefix in the
    binding, the others drop it)."""
    text = open(HEADER).read()
    names = sorted(set(re.findall(r"\b(SHG_[A-Z0-9_]+)\s*=", text)) | {"SHG_PLAN_ROWS_WIDE"})
    src = tmp_path / "consts.c"
    body = "".join(f'  printf("{n} %d\\n", (int)({n}));\n' for n in names)
    src.write_text(f'#include <stdio.h>\n#include "shg.h"\nint main(void){{\n{body}  return 0;\n}}\n')
    exe = tmp_path / "consts"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    seen = {ln.split()[0]: int(ln.split()[1]) for ln in out.splitlines()}
    assert len(seen) == len(names) and "SHG_PREC_FP32_SPLIT" in seen
    for n, v in seen.items():
        py = n if hasattr(abi, n) else n[len("SHG_"):]
        assert getattr(abi, py) == v, n


@pytest.mark.parametrize("struct,cls", [("shg_params", abi.ShgParams), ("shg_result", abi.ShgResult)])
def test_struct_layout_matches_header(tmp_path, struct, cls):
    names = [f for f, _ in cls._fields_]
    lay = _c_layout(tmp_path, struct, names)
    assert lay["sizeof"] == C.sizeof(cls)
    for f in names:
        assert lay[f] == getattr(cls, f).offset, f


def test_defaults_are_the_reference_constants():
    L = shg.lib()
    p = abi.ShgParams()
    L.shg_params_default(C.byref(p))
    q = abi.default_params()
    for f, _ in abi.ShgParams._fields_:
        if f == "poling_phase":
            assert not p.poling_phase
            continue
        assert getattr(p, f) == getattr(q, f), f
    # src/cuSHG.cu:11-13, src/headers/Crystal.h:26-36, src/cuSHG.sh:43-45
    assert (p.nx, p.ny, p.nz) == (128, 128, 100)
    assert (p.lx_um, p.ly_um, p.lz_um) == (2000.0, 1000.0, 30000.0)
    assert p.lambda_p_um == 1.064 and p.Lambda_um == 7.97 and p.deff_um_per_V == 11.0e-6
    assert p.focal_um == 0.5 * p.lz_um and p.T_inf_C == 25.0 and p.T_peltier_C == p.T_C
    assert p.max_outer == 20 and p.check_every == 5000 and p.jacobi_tol == 5e-4


@pytest.mark.parametrize("crystal", [abi.CRYSTAL_MGO_SPPLT, abi.CRYSTAL_MGO_PPLN])
def test_crystal_database_matches_mirror_and_oracle(crystal):
    L = shg.lib()
    p = abi.default_params()
    assert L.shg_params_set_crystal(C.byref(p), crystal) == abi.SHG_OK
    q = abi.default_params()
    abi.set_crystal(q, crystal)
    for f in ("deff_um_per_V", "Lambda_um", "k_th", "therm_exp_a", "therm_exp_b", "alpha_p", "alpha_s"):
        assert getattr(p, f) == getattr(q, f), f
    for lam in (1.064, 0.532, 1.55):
        for T in (25.0, 40.0, 56.9):
            a = L.shg_refractive_index(crystal, lam, T)
            assert a == pytest.approx(O.refractive_index(crystal, lam, T), rel=1e-15, abs=0)
            assert a == pytest.approx(abi.refractive_index(crystal, lam, T), rel=1e-14, abs=0)
    for T in (25.0, 40.0, 55.66):
        assert L.shg_delta_k(C.byref(p), T) == pytest.approx(O.delta_k(p, T), rel=1e-12, abs=1e-18)
    assert L.shg_params_set_crystal(C.byref(p), 7) == abi.SHG_EINVAL


def test_library_ppln_matches_decimal_fixture():
    """The library's own MgO:PPLN index and dk against the 40-digit decimal
    evaluation of Gayer et al. (tests/golden/sellmeier_ppln.json)."""
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sellmeier_ppln.json")))
    L = shg.lib()
    p = abi.default_params()
    assert L.shg_params_set_crystal(C.byref(p), abi.CRYSTAL_MGO_PPLN) == abi.SHG_OK
    assert (p.Lambda_um, p.therm_exp_a, p.therm_exp_b) == (g["Lambda_um"], g["therm_exp_a"], g["therm_exp_b"])
    for row in g["rows"]:
        assert L.shg_refractive_index(abi.CRYSTAL_MGO_PPLN, 1.064, row["T"]) == pytest.approx(row["n_p"], rel=2e-15)
        assert L.shg_refractive_index(abi.CRYSTAL_MGO_PPLN, 0.532, row["T"]) == pytest.approx(row["n_s"], rel=2e-15)
        assert L.shg_delta_k(C.byref(p), row["T"]) == pytest.approx(row["dk"], rel=1e-11, abs=3e-14)
    for row in g["index"]:
        assert L.shg_refractive_index(abi.CRYSTAL_MGO_PPLN, row["lambda_um"], row["T"]) == pytest.approx(
            row["n"], rel=2e-15)


def _bad(**kw):
    p = abi.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("kw,msg", [
    (dict(nx=100), "nx=100"),
    (dict(ny=1024), "ny=1024"),
    (dict(nz=1), "nz=1"),
    (dict(thermal=1, nz=100, nz_heat=1), "nz_heat=1"),
    (dict(thermal=1, nz=10, nz_heat=20), "nz_heat=20 must not exceed nz=10"),
    (dict(lz_um=0.0), "lengths"),
    (dict(waist_um=-1.0), ""),
    (dict(power_W=float("nan")), ""),
    (dict(crystal=5), "crystal"),
    (dict(scheme=9), "scheme"),
    (dict(precision=5), "precision"),
    (dict(thermal=1, heat_solver=4), "heat"),
    (dict(thermal=1, heat_bc=5), "boundary"),
    (dict(Lambda_um=0.0), "Lambda"),
])
def test_validate_rejects(kw, msg):
    L = shg.lib()
    p = _bad(**kw)
    assert L.shg_validate(C.byref(p)) == abi.SHG_EINVAL
    assert msg in shg.last_error()


def test_validate_accepts_every_supported_size():
    L = shg.lib()
    for n in abi.FFT_SIZES:
        assert L.shg_fft_size_supported(n) == 1
        p = _bad(nx=n, ny=abi.FFT_SIZES[-1])
        assert L.shg_validate(C.byref(p)) == abi.SHG_OK
    for n in (0, 32, 96, 100, 320, 1024):
        assert L.shg_fft_size_supported(n) == 0


def test_validate_accepts_every_precision():
    L = shg.lib()
    for prec in (abi.PREC_FP32_MIXED, abi.PREC_FP64_PLAIN, abi.PREC_FP32_STORAGE, abi.PREC_FP64_FUSED,
                 abi.PREC_FP32_SPLIT):
        p = _bad(precision=prec)
        assert L.shg_validate(C.byref(p)) == abi.SHG_OK


def test_null_arguments_are_einval_not_crashes():
    L = shg.lib()
    assert L.shg_validate(None) == abi.SHG_EINVAL
    assert L.shg_run(None, None, None) == abi.SHG_EINVAL
    assert L.shg_run_batch(None, None, 0, None) == abi.SHG_EINVAL
    assert L.shg_handle_destroy(None) == abi.SHG_OK
    assert L.shg_set_graph_steps(None, 4) == abi.SHG_EINVAL
    assert L.shg_get_plan(None, 64, abi.PREC_FP32_MIXED, None, None) == abi.SHG_EINVAL
    assert L.shg_autotune(None, 64, abi.PREC_FP32_MIXED, 1, 8) == abi.SHG_EINVAL
    assert L.shg_set_plan(None, 64, abi.PREC_FP32_MIXED, 1, 4) == abi.SHG_EINVAL


def test_plain_c_program_links_and_calls_the_library(tmp_path):
    """A C99 consumer needs only include/shg.h and -lshg_hip (no HIP headers,
    no C++): it links, and the host-only entry points answer as the Python
    mirror does."""
    src = tmp_path / "consumer.c"
    src.write_text(
        "#include <stdio.h>\n#include \"shg.h\"\n"
        "int main(void) {\n"
        "  shg_params p; shg_params_default(&p);\n"
        "  printf(\"%d %d %d %d\\n\", p.nx, p.nz, shg_validate(&p), shg_validate(0));\n"
        "  printf(\"%.17g\\n\", shg_refractive_index(SHG_CRYSTAL_MGO_SPPLT, 1.064, 25.0));\n"
        "  printf(\"%d %d\\n\", shg_fft_size_supported(384), shg_fft_size_supported(100));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "consumer"
    libdir = os.path.dirname(shg.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-L", libdir, "-lshg_hip",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out[:4] == ["128", "100", "0", str(abi.SHG_EINVAL)]
    assert float(out[4]) == pytest.approx(O.refractive_index(0, 1.064, 25.0), rel=1e-15)
    assert out[5:] == ["1", "0"]
