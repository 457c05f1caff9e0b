"""The C++ driver of the C ABI (fastfortranboltzmann_amd/bin/ffb_run, built
from cpp/ffb_run.cpp by build()): argument handling on the CPU; on the GPU,
its run of a cavity is bitwise the Python front end's run of the same
configuration (same library, same kernels) and its FP64 mass sum is
ffb_moments'."""
import json
import os
import subprocess

import numpy as np
import pytest

from fastfortranboltzmann_amd import build as B


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(B.LIB):
        pytest.skip("libffb_hip.so not built")
    if not os.path.exists(B.RUN_BIN) or os.path.getmtime(B.RUN_BIN) < os.path.getmtime(
            os.path.join(B.PKG, "cpp", "ffb_run.cpp")):
        B.build_cpp_driver()
    return B.RUN_BIN


def test_help_and_bad_arguments(driver):
    r = subprocess.run([driver, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage: ffb_run" in r.stdout
    for args in (["--mode", "xx"], ["--precision", "fp8"], ["--nx"], ["--steps", "abc"],
                 ["--omega", "1.2", "--re", "100"], ["--bogus", "1"], ["--steps", "0"]):
        r = subprocess.run([driver, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.returncode, r.stderr)
        assert "ffb_run:" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("precision,mode", [("fp64", "sl"), ("fp32", "pf"), ("e3m12dev", "sl")])
def test_driver_matches_python_front_end(driver, tmp_path, precision, mode):
    from fastfortranboltzmann_amd import Lattice, Settings

    nx, ny, steps = 67, 45, 120
    print(f"[gpu] ffb_run {precision} {mode} {nx}x{ny}", flush=True)
    out = tmp_path / "macros.bin"
    args = [driver, "--nx", str(nx), "--ny", str(ny), "--mode", mode, "--precision", precision,
            "--omega", "1.3", "--u", "0.08", "--steps", str(steps), "--warmup", "0", "--macros", str(out)]
    if precision == "e3m12dev":
        args += ["--scale", "5"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    dt = np.float64 if precision == "fp64" else np.float32
    fields = np.fromfile(out, dtype=dt).reshape(3, ny, nx)
    s = Settings(N_X=nx, N_Y=ny, SIM_MODE={"sl": 4, "pf": 3}[mode], OMEGA=1.3, U_LID=0.08, U_WALL=0.08,
                 U_MAX=0.08, precision=precision, storage_scale_log2=5 if precision == "e3m12dev" else 0)
    with Lattice(s) as lat:
        lat.init().step(steps, write_macro_last=True)
        rho, ux, uy = lat.macros()
        mass = lat.moments()[0]
    np.testing.assert_array_equal(fields[0], rho)
    np.testing.assert_array_equal(fields[1], ux)
    np.testing.assert_array_equal(fields[2], uy)
    assert rec["mass"] == mass
    assert rec["steps"] == steps and rec["mlups"] > 0
