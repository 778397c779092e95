"""Builds the HIP library cushg_amd/lib/libshg_hip.so for gfx950 (in-tree).

Every source under cushg_amd/csrc is compiled with hipcc (HIP language,
gfx950 only) and linked into one shared library whose exported symbols are
exactly the C ABI of include/shg.h.  The CLI `shg` (the reference's program
interface) is linked against it.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libshg_hip.so")
CLI = os.path.join(LIBDIR, "shg")

ARCH = os.environ.get("SHG_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions (v_pk_*_f32) in any HIP code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hip", ".cpp")) and not f.startswith("cli_"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", src, "-o", obj] + COMMON
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp, *objs, f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "cli_main.cpp")
    if os.path.exists(cli_src):
        cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), cli_src, "-o", CLI,
               "-L", LIBDIR, "-lshg_hip", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI build failed:\n{r.stdout}")
    if verbose:
        print(f"built {LIB}")
    return LIB


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build (test
# infrastructure): every source compiled with the sanitizers on its host
# pass only (the GPU code objects are the normal ones), linked into a
# separate library with the driver tests/native/host_sanity.cpp.
SAN = os.path.join(HERE, "_san")
SAN_LIB = os.path.join(SAN, "libshg_hip_san.so")
SAN_EXE = os.path.join(SAN, "host_sanity")
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]


def _compile_san(src: str) -> str:
    obj = os.path.join(SAN, os.path.basename(src) + ".o")
    flags = [f for f in COMMON if f not in ("-O3",)] + ["-O1", "-g"]
    cmd = [HIPCC, "-x", "hip", "-c", src, "-o", obj] + flags + SAN_HOST
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc (sanitize) failed for {src}:\n{r.stdout}")
    return obj


def build_sanitized(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(SAN, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile_san, srcs))
    cmd = [HIPCC, "-shared", "-o", SAN_LIB, *objs, f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link (sanitize) failed:\n{r.stdout}")
    drv = os.path.join(ROOT, "tests", "native", "host_sanity.cpp")
    cmd = [HIPCC, "-O1", "-g", "-std=c++17", "-I", os.path.join(ROOT, "include"), drv, "-o", SAN_EXE,
           "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-L", SAN, "-lshg_hip_san",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sanitizer driver build failed:\n{r.stdout}")
    if verbose:
        print(f"built {SAN_EXE}")
    return SAN_EXE


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        build_sanitized(verbose=True)
    else:
        build(verbose=True)
    sys.exit(0)
