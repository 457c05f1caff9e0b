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
