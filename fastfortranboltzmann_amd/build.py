"""Compile the HIP library libffb_hip.so for gfx950, in-tree.

    python -m fastfortranboltzmann_amd.build

Cross-compiles without a GPU.  Every HIP translation unit is built with
contraction off (the kernels keep the reference's separately rounded
operations) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libffb_hip.so")

SOURCES = ["ffb_api.hip", "ffb_batch.hip", "ffb_adjoint.hip"]
HEADERS = ["d2q9_kernels.h", "ffb_internal.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


FORTRAN = ["ffb_c.f90", "ffb_settings.f90", "ffb_host.f90"]
HOST_BIN = os.path.join(PKG, "bin", "ffb_host")


def flang() -> str | None:
    for cand in (os.environ.get("FLANG"), "/opt/rocm/llvm/bin/flang", "flang", "amdflang"):
        if not cand:
            continue
        if os.path.sep in cand and os.path.exists(cand):
            return cand
        if os.path.sep not in cand:
            from shutil import which
            if which(cand):
                return cand
    return None


def build_fortran_host(verbose: bool = False, defines: list[str] | None = None) -> str | None:
    """Compile the Fortran host (fortran/*.f90) against libffb_hip.so with flang.
    `defines` are FFB_* preprocessor settings, as in the reference's
    FFB_SETTINGS_DEFINITIONS.  Returns the binary path, or None without flang."""
    fc = flang()
    if fc is None:
        return None
    import tempfile

    os.makedirs(os.path.dirname(HOST_BIN), exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        objs = []
        for src in FORTRAN:
            obj = os.path.join(tmp, src.replace(".f90", ".o"))
            cmd = [fc, "-cpp", "-O2", "-module-dir", tmp, "-I", tmp,
                   *[f"-D{d}" for d in (defines or [])],
                   "-c", os.path.join(PKG, "fortran", src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True, capture_output=not verbose)
            objs.append(obj)
        cmd = [fc, "-o", HOST_BIN + ".tmp", *objs, "-L", PKG, "-lffb_hip", "-Wl,-rpath,$ORIGIN/.."]
        subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(HOST_BIN + ".tmp", HOST_BIN)
    return HOST_BIN


RUN_BIN = os.path.join(PKG, "bin", "ffb_run")


def build_cpp_driver(verbose: bool = False) -> str:
    """Compile the C++ driver (cpp/ffb_run.cpp) against libffb_hip.so with g++:
    host code over the C ABI only, no HIP headers."""
    os.makedirs(os.path.dirname(RUN_BIN), exist_ok=True)
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-Wextra",
           "-I", os.path.join(REPO, "include"), os.path.join(PKG, "cpp", "ffb_run.cpp"),
           "-o", RUN_BIN + ".tmp", "-L", PKG, "-lffb_hip", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(RUN_BIN + ".tmp", RUN_BIN)
    return RUN_BIN


def build(verbose: bool = False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(REPO, "include"), "-I", CSRC,
           "-o", LIB + ".tmp", *[os.path.join(CSRC, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libffb_hip.so")
    os.replace(LIB + ".tmp", LIB)
    build_fortran_host(verbose)
    build_cpp_driver(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(HOST_BIN if os.path.exists(HOST_BIN) else "(no flang: Fortran host not built)")
    print(RUN_BIN)
