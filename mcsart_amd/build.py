"""Build libmcsart_hip.so in place with hipcc for gfx950.

Usage: ``python -m mcsart_amd.build`` (or ``build()`` from Python).  Each HIP source is compiled
to an object under ``mcsart_amd/_build/`` in parallel, then linked into
``mcsart_amd/libmcsart_hip.so`` next to this file.
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
LIB = os.path.join(HERE, "libmcsart_hip.so")

SOURCES = ["capi.hip", "driver.hip", "serve.hip", "params.hip", "k_project.hip",
           "k_backproject.hip", "k_elem.hip", "noalloc.hip"]

ARCH = os.environ.get("MCSART_OFFLOAD_ARCH", "gfx950")
CXXFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    # no packed FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *CXXFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_cli()
    _build_compat_check()
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


CLI = os.path.join(HERE, "mcsart_recon")


def _build_cli() -> None:
    """The file-driven driver (host code only) next to the library, found through $ORIGIN."""
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        raise RuntimeError("no host C++ compiler for mcsart_recon")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "recon_cli.cpp"), "-o", CLI + ".tmp", "-L", HERE, "-lmcsart_hip",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(CLI + ".tmp", CLI)


COMPAT_SRC = os.path.join(ROOT, "tests", "compat_check.cpp")
COMPAT_BIN = os.path.join(ROOT, "tests", "compat_check.bin")


def compat_check_command(out: str = COMPAT_BIN):
    """g++ line of the reference-side binding check (tests/compat_check.cpp): the reference's
    host_* calls compiled against include/mcsart_compat.h and linked with libmcsart_hip."""
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        raise RuntimeError("no host C++ compiler")
    return [cxx, "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I",
            os.path.join(ROOT, "include"), COMPAT_SRC, "-o", out, "-L", HERE, "-lmcsart_hip",
            "-Wl,-rpath,$ORIGIN/../mcsart_amd"]


def _build_compat_check() -> None:
    if not os.path.exists(COMPAT_SRC):
        return
    r = subprocess.run(compat_check_command(COMPAT_BIN + ".tmp"), capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compat check build failed\n{r.stdout}\n{r.stderr}")
    os.replace(COMPAT_BIN + ".tmp", COMPAT_BIN)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
