"""AddressSanitizer + UndefinedBehaviorSanitizer runs on the CPU.

  - the FP64 oracle with a driver that walks every code path on small grids
    (`make -C oracle sanitize`, oracle/sanity_main.c), leak detection on;
  - the host side of the HIP library (compiled with the sanitizers on its
    host pass only, `python cushg_amd/build.py --sanitize`) with the driver
    tests/native/host_sanity.cpp: parameter handling, validation, the
    crystal database, the CLI's argument checks and handle creation failing
    without a GPU.  On a machine with a GPU the same driver with argument `gpu` also
    runs every form, THERMAL solver and boundary set, batches, the trace,
    the adjoint, plans and the CLI under the sanitizers.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _env():
    e = dict(os.environ)
    e["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    e["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return e


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "oracle_sanity")], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, env=_env(), timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:]
    assert "sanity: ok" in r.stdout


def test_host_library_under_asan_ubsan():
    from cushg_amd import build
    if not os.path.exists(build.SAN_EXE):
        build.build_sanitized()
    e = _env()
    # the HIP runtime's own allocations outlive main(); the library's host
    # code is what is checked for leaks through the driver's calls
    e["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0"
    r = subprocess.run([build.SAN_EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=e,
                       timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:]
    assert "host_sanity: ok" in r.stdout
