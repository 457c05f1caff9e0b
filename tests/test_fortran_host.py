"""The Fortran host (fastfortranboltzmann_amd/fortran/ffb_host.f90) over the C ABI:
its stdout satisfies the reference harness's parsers and its exported fields
equal the golden vectors the reference's own kernels wrote."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "fastfortranboltzmann_amd", "bin", "ffb_host")
GOLDEN = os.path.join(REPO, "tests", "golden")

# the parsers of scripts/benchmark_run.py:17-26 (the stdout contract)
NUMBER = r"(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)"
STEP_RE = re.compile(rf"step time\s*[:=]\s*({NUMBER})\s*ms", re.IGNORECASE)
MLUPS_RE = re.compile(rf"MLUPS\s*[:=]\s*({NUMBER})", re.IGNORECASE)
TIMING_SPREAD_RE = re.compile(
    rf"^(kernel compute|halo sync|halo transfer|other|total)\s*"
    rf"\|\s*({NUMBER})\s*\((\d+)\)\s*"
    rf"\|\s*({NUMBER})\s*\((\d+)\)\s*$",
    re.IGNORECASE | re.MULTILINE,
)
LAUNCHED_RE = re.compile(r"^\[[0-9:]+\]\s+launched", re.MULTILINE)

needs_host = pytest.mark.skipif(not os.path.exists(HOST), reason="Fortran host not built (no flang)")


@needs_host
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU may be present")
def test_host_without_gpu_stops_with_message(tmp_path):
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True,
                       env=dict(os.environ, FFB_N_X="8", FFB_N_Y="8"), timeout=60)
    assert r.returncode != 0
    assert "ffb_create" in (r.stdout + r.stderr)


@needs_host
def test_host_rejects_multi_image_decomposition(tmp_path):
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True,
                       env=dict(os.environ, I_X="2", I_Y="1", FFB_N_X="8", FFB_N_Y="8"), timeout=60)
    assert r.returncode != 0
    assert "one domain" in (r.stdout + r.stderr)


@needs_host
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sl_fp64_48x40", "pf_fp64_40x24", "sw_fp32_32x32"])
def test_host_run_matches_golden_and_harness_contract(tmp_path, name):
    case = next(c for c in json.load(open(os.path.join(GOLDEN, "manifest.json")))["cases"]
                if c["name"] == name)
    print(f"[gpu] fortran host {name}", flush=True)
    env = dict(os.environ, FFB_N_X=str(case["n_x"]), FFB_N_Y=str(case["n_y"]),
               FFB_N_STEPS=str(case["n_steps"]), FFB_SIM_MODE=str(case["sim_mode"]),
               FFB_PRECISION=case["precision"], FFB_EXPORT_MACROS="true",
               FFB_EXPORT_ENDPOINT_STATES="true", FFB_EXPORT_INTERVAL="100000",
               FFB_EXPORT_NUM="t", FFB_PROGRESS_INTERVAL="50",
               I_X="1", I_Y="1", FOR_COARRAY_NUM_IMAGES="1")
    for k, v in case["literals"].items():
        if v.endswith("_FP"):
            env[k] = v
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    assert LAUNCHED_RE.search(out)
    assert STEP_RE.search(out) and MLUPS_RE.search(out)
    rows = {m.group(1) for m in TIMING_SPREAD_RE.finditer(out)}
    assert rows == {"kernel compute", "halo sync", "halo transfer", "other", "total"}

    g = np.load(os.path.join(GOLDEN, case["file"]))
    dt = np.float64 if case["precision"] == "fp64" else np.float32
    d = tmp_path / "output" / "t"
    meta = json.load(open(d / "config.json"))
    assert meta["N_X"] == case["n_x"] and meta["file_dtype"] == ("real64" if dt == np.float64 else "real32")
    step = f"{case['n_steps']:09d}"
    for field in ("density", "velocity_x", "velocity_y"):
        got = np.fromfile(d / f"{field}_{step}.bin", dtype=dt).reshape(case["n_y"], case["n_x"])
        want = g[{"density": "rho", "velocity_x": "u_x", "velocity_y": "u_y"}[field]]
        np.testing.assert_array_equal(got, want)
    # the report and metadata blocks of reporting.f90:46 / hardware_info.f90:138-144
    # and export.f90:170, 218-247
    for label in ("--- [ hardware ]", "cpu model", "logical threads/node", "compute nodes", "compiler",
                  "flags", "latest commit", "gpu"):
        assert label in out, label
    hw = meta["hardware"]
    assert hw["cpu_model"] and hw["logical_threads_per_node"] >= 1 and hw["compute_nodes"] == 1
    assert "flang" in hw["compiler"].lower() and hw["gpu_compute_units"] > 0
    assert meta["FFB_DENSITY_CHECKS"] is False
    assert meta["domain_decomposition"]["halo_cells_percent"] == 0.0
    mem = meta["memory_usage"]
    bytes_pop = 8 if dt == np.float64 else 4
    assert mem["dist_function_buffers_per_cell_B"] >= 2 * 9 * bytes_pop
    assert mem["total_GB"] >= mem["dist_function_buffers_total_GB"] + mem["macro_field_buffers_total_GB"] - 1e-9
    mag = np.fromfile(d / f"velocity_mag_{step}.bin", dtype=dt).reshape(case["n_y"], case["n_x"])
    np.testing.assert_allclose(mag, np.sqrt(g["u_x"] ** 2 + g["u_y"] ** 2), rtol=1e-6 if dt == np.float32 else 1e-14)
    # initial state exported at step 0
    assert (d / "density_000000000.bin").exists()


@needs_host
@pytest.mark.gpu
def test_host_deviation_form_takes_the_library_default_scale(tmp_path):
    """FFB_PRECISION=e3m12dev without FFB_STORAGE_SCALE_LOG2: the host asks the
    library for the default scale at the lid speed (s = 5 for u = 0.1), and its
    exported fields are bitwise the Python front end's at that scale."""
    from fastfortranboltzmann_amd import Lattice, Settings
    from fastfortranboltzmann_amd.precision import default_scale_log2

    nx, ny, steps = 40, 32, 60
    print("[gpu] fortran host e3m12dev default scale", flush=True)
    env = dict(os.environ, FFB_N_X=str(nx), FFB_N_Y=str(ny), FFB_N_STEPS=str(steps), FFB_SIM_MODE="4",
               FFB_OMEGA="1.3_FP", FFB_U_LID="0.1_FP", FFB_PRECISION="e3m12dev", FFB_EXPORT_MACROS="true",
               FFB_EXPORT_ENDPOINT_STATES="true", FFB_EXPORT_INTERVAL="100000", FFB_EXPORT_NUM="t",
               I_X="1", I_Y="1", FOR_COARRAY_NUM_IMAGES="1")
    env.pop("FFB_STORAGE_SCALE_LOG2", None)
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    s = default_scale_log2("e3m12dev", 0.1)
    assert s == 5
    with Lattice(Settings(N_X=nx, N_Y=ny, SIM_MODE=4, OMEGA=1.3, U_LID=0.1, precision="e3m12dev",
                          storage_scale_log2=s)) as lat:
        lat.init().step(steps, write_macro_last=True)
        rho, ux, uy = lat.macros()
    d = tmp_path / "output" / "t"
    step = f"{steps:09d}"
    for field, want in (("density", rho), ("velocity_x", ux), ("velocity_y", uy)):
        got = np.fromfile(d / f"{field}_{step}.bin", dtype=np.float32).reshape(ny, nx)
        np.testing.assert_array_equal(got, want)


# read_cmake_cache / read_cmake_settings_definitions of scripts/benchmark_run.py:160-200
CMAKE_CACHE_RE = re.compile(r"^([^#/\n][^:=\n]*):[^=\n]*=(.*)$", re.MULTILINE)


def _harness_sim_size(exe):
    from pathlib import Path

    p = Path(exe).resolve()
    build_dir = p.parent.parent if p.parent.name == "bin" else p.parent
    cache = {k: v.strip() for k, v in CMAKE_CACHE_RE.findall((build_dir / "CMakeCache.txt").read_text())}
    defs = dict(d.split("=", 1) if "=" in d else (d, "") for d in cache["FFB_SETTINGS_DEFINITIONS"].split(";") if d)
    if "FFB_USE_CMAKE_SETTINGS" not in defs:
        return None  # the harness then reads source/settings.f90 (2520 x 2520, the host's default too)
    return int(defs["FFB_N_X"]), int(defs["FFB_N_Y"])


@needs_host
def test_host_build_records_its_grid_for_the_harness(tmp_path):
    """A host built with FFB_SETTINGS_DEFINITIONS (every setting, as the
    reference's CMake runs pass them) carries a CMakeCache.txt beside bin/ from
    which the harness's read_sim_size gets that grid; the default build's cache
    has no settings definitions (the harness's fallback, 2520 x 2520, is the
    host's default)."""
    from fastfortranboltzmann_amd import build

    defs = ["FFB_USE_CMAKE_SETTINGS", "FFB_N_X=4096", "FFB_N_Y=2048", "FFB_N_STEPS=100", "FFB_SIM_MODE=4",
            "FFB_RHO_0=1.0_FP", "FFB_OMEGA=1.1_FP", "FFB_U_MAX=0.1_FP", "FFB_N_SIN=2.0_FP", "FFB_U_WALL=0.1_FP",
            "FFB_U_LID=0.1_FP", "FFB_RHO_IN=1.00005_FP", "FFB_RHO_OUT=0.99995_FP", "FFB_EXPORT_MACROS=.false.",
            "FFB_EXPORT_ENDPOINT_STATES=.true.", "FFB_EXPORT_INTERVAL=10000", "FFB_INTERACTIVE_PROGRESS=.false.",
            "FFB_PROGRESS_INTERVAL=50", "FFB_FP64", "FFB_DENSITY_CHECKS"]
    exe = build.build_fortran_host(defines=defs, out_dir=str(tmp_path))
    assert exe == str(tmp_path / "bin" / "ffb_host") and os.path.exists(exe)
    assert _harness_sim_size(exe) == (4096, 2048)
    text = (tmp_path / "CMakeCache.txt").read_text()
    assert "FFB_FP64:BOOL=ON" in text and "FFB_DENSITY_CHECKS:BOOL=ON" in text
    assert "FFB_FP64;" not in text and not text.rstrip().endswith(";FFB_FP64")
    assert _harness_sim_size(HOST) is None


@needs_host
@pytest.mark.gpu
def test_host_density_checks_setting_reaches_the_library_and_metadata(tmp_path):
    """FFB_DENSITY_CHECKS=true: the run (healthy state) completes, and
    config.json records the setting as the reference's does (export.f90:218-221)."""
    print("[gpu] fortran host density checks", flush=True)
    env = dict(os.environ, FFB_N_X="24", FFB_N_Y="16", FFB_N_STEPS="20", FFB_SIM_MODE="4",
               FFB_EXPORT_MACROS="true", FFB_EXPORT_INTERVAL="100000", FFB_EXPORT_NUM="t",
               FFB_DENSITY_CHECKS="true", I_X="1", I_Y="1", FOR_COARRAY_NUM_IMAGES="1")
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    meta = json.load(open(tmp_path / "output" / "t" / "config.json"))
    assert meta["FFB_DENSITY_CHECKS"] is True
