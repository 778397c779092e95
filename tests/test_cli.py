"""The `shg` program's argument handling (no GPU work: every case here is
rejected before a device is touched)."""
import os
import subprocess

import pytest

from cushg_amd import abi, shg

SHG_EXE = os.path.join(os.path.dirname(shg.LIB_PATH), "shg")


def run(*args):
    return subprocess.run([SHG_EXE, *args], capture_output=True, text=True, timeout=60)


def test_program_is_built():
    assert os.access(SHG_EXE, os.X_OK)


def test_too_few_positional_arguments():
    r = run("30", "29", "55.66")
    assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.parametrize("args,msg", [
    (("x", "29", "55.66", "1", "0"), "not a number"),
    (("30", "29", "55.66", "1", "0", "--nx=100"), "nx=100"),
    (("30", "29", "55.66", "1", "0", "--lz=0"), "lengths"),
    (("30", "29", "55.66", "1", "0", "--waist-y=-3"), "waist_y_um"),
])
def test_invalid_inputs_fail_before_any_gpu_work(args, msg):
    r = run(*args)
    assert r.returncode == 1 and msg in r.stderr


@pytest.mark.parametrize("opt", ["--heat-bc=2", "--heat=gauss", "--scheme=euler", "--precision=fp16"])
def test_unknown_choice_is_rejected(opt):
    r = run("30", "29", "55.66", "1", "0", opt)
    assert r.returncode == 2 and "is not one of" in r.stderr


def test_unknown_option_and_crystal():
    assert run("30", "29", "55.66", "1", "0", "--bogus=1").returncode == 2
    r = run("30", "29", "55.66", "1", "0", "--crystal=quartz")
    assert r.returncode == 2 and "crystal" in r.stderr


def test_cli_main_usage_via_c_abi():
    import ctypes as C
    L = shg.lib()
    argv = (C.c_char_p * 2)(b"shg", b"30")
    assert L.shg_cli_main(2, argv, None, b".") == abi.SHG_EINVAL
    assert "usage" in shg.last_error()
