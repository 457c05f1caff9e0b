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

SOURCES = ["ffb_api.hip", "ffb_batch.hip", "ffb_adjoint.hip", "ffb_info.hip"]
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


def git_commit() -> str:
    """Short hash of the checked-out commit, or "unknown" (the reference's FFB_GIT_COMMIT)."""
    try:
        r = subprocess.run(["git", "-C", REPO, "rev-parse", "--short", "HEAD"], capture_output=True, text=True,
                           timeout=10)
        return r.stdout.strip() if r.returncode == 0 and r.stdout.strip() else "unknown"
    except (OSError, subprocess.TimeoutExpired):
        return "unknown"


# definitions the reference sets through CMake options of their own, not through
# FFB_SETTINGS_DEFINITIONS (CMakeLists.txt:41, 106-110)
OPTION_DEFINES = ("FFB_FP64", "FFB_DENSITY_CHECKS")


def write_cmake_cache(build_dir: str, defines: list[str]) -> str:
    """CMakeCache.txt in the directory above bin/, where scripts/benchmark_run.py
    (read_sim_size, :186-200) looks for FFB_SETTINGS_DEFINITIONS to learn the
    grid size of a build; FFB_FP64 and FFB_DENSITY_CHECKS as the reference's
    BOOL options."""
    settings = [d for d in defines if d.split("=", 1)[0] not in OPTION_DEFINES]
    path = os.path.join(build_dir, "CMakeCache.txt")
    with open(path, "w") as f:
        f.write("# written by fastfortranboltzmann_amd/build.py for the Fortran host in bin/\n")
        f.write("FFB_SETTINGS_DEFINITIONS:STRING=" + ";".join(settings) + "\n")
        for opt in OPTION_DEFINES:
            f.write(f"{opt}:BOOL={'ON' if opt in defines else 'OFF'}\n")
    return path


def build_fortran_host(verbose: bool = False, defines: list[str] | None = None,
                       out_dir: str | None = None) -> str | None:
    """Compile the Fortran host (fortran/*.f90) against libffb_hip.so with flang.
    `defines` are FFB_* preprocessor settings, as in the reference's
    FFB_SETTINGS_DEFINITIONS (with FFB_USE_CMAKE_SETTINGS every setting must be
    given) plus FFB_FP64 / FFB_DENSITY_CHECKS.  The binary goes to
    <out_dir>/bin/ffb_host (default: the package) with a CMakeCache.txt in
    <out_dir> that records the definitions.  Returns the binary path, or None
    without flang."""
    fc = flang()
    if fc is None:
        return None
    import tempfile

    defines = list(defines or [])
    build_dir = out_dir or PKG
    host_bin = os.path.join(build_dir, "bin", "ffb_host")
    os.makedirs(os.path.dirname(host_bin), exist_ok=True)
    rpath = "-Wl,-rpath," + (os.path.dirname(LIB) if out_dir else "$ORIGIN/..")
    with tempfile.TemporaryDirectory() as tmp:
        objs = []
        for src in FORTRAN:
            obj = os.path.join(tmp, src.replace(".f90", ".o"))
            cmd = [fc, "-cpp", "-O2", "-module-dir", tmp, "-I", tmp,
                   *[f"-D{d}" for d in defines], f'-DFFB_GIT_COMMIT="{git_commit()}"',
                   "-c", os.path.join(PKG, "fortran", src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True, capture_output=not verbose)
            objs.append(obj)
        cmd = [fc, "-o", host_bin + ".tmp", *objs, "-L", PKG, "-lffb_hip", rpath]
        subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(host_bin + ".tmp", host_bin)
    write_cmake_cache(build_dir, defines)
    return host_bin


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
    """Compile every translation unit to an object in parallel (each holds its
    own device code; no symbol crosses units on the device), then link."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    inc = ["-I", os.path.join(REPO, "include"), "-I", CSRC]
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory() as tmp:
        jobs = []
        for s in SOURCES:
            obj = os.path.join(tmp, s.replace(".hip", ".o"))
            jobs.append(([hipcc(), *compile_flags, *inc, "-c", os.path.join(CSRC, s), "-o", obj], obj))

        def run(cmd):
            if verbose:
                print(" ".join(cmd), flush=True)
            return subprocess.run(cmd, capture_output=True, text=True)

        with ThreadPoolExecutor(max_workers=len(jobs)) as pool:
            results = list(pool.map(run, [c for c, _ in jobs]))
        for res in results:
            if res.returncode != 0:
                sys.stderr.write(res.stdout + res.stderr)
                raise RuntimeError("hipcc failed building libffb_hip.so")
        cmd = [hipcc(), *HIPCC_FLAGS, "-o", LIB + ".tmp", *[o for _, o in jobs]]
        res = run(cmd)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed linking libffb_hip.so")
    os.replace(LIB + ".tmp", LIB)
    build_fortran_host(verbose)
    build_cpp_driver(verbose)
    return LIB


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="build libffb_hip.so, the Fortran host and the C++ driver")
    ap.add_argument("--host-settings", default=None,
                    help="only rebuild the Fortran host, with these ';'-separated FFB_* definitions "
                         "(e.g. 'FFB_USE_CMAKE_SETTINGS;FFB_N_X=4096;...;FFB_FP64'), as the reference's "
                         "FFB_SETTINGS_DEFINITIONS; writes CMakeCache.txt beside bin/")
    args = ap.parse_args()
    if args.host_settings is not None:
        print(build_fortran_host(True, [d for d in args.host_settings.split(";") if d]))
    else:
        print(build(verbose=True))
        print(HOST_BIN if os.path.exists(HOST_BIN) else "(no flang: Fortran host not built)")
        print(RUN_BIN)
