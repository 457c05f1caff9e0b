"""Write the golden vectors under tests/golden/ from the reference's own kernels.

For every case below, oracle/build_ref.sh compiles the UNMODIFIED reference
modules (source/settings.f90, source/layouts/xyd/{modes/*,initialization,
simulation}.f90, read in place under the reference tree) with flang, at -O0
and -ffp-contract=off, together with the project's single-domain driver
(oracle/ref_driver.f90), runs the case once, and this script stores the final
state as a small ``.npz`` fixture plus a line in ``manifest.json``.

Needs the reference tree and flang, so it runs in the build container only;
the fixtures it writes are committed and are all the tests read.

Fixture contents (arrays are numpy C order):
  f         (9, NY, NX)  populations of the owned cells after N_STEPS steps
  rho, u_x, u_y (NY, NX) macro fields written by the last step
  recv_macro (2, 3, NY)  Poiseuille inlet/outlet strips (zeros otherwise)
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
BUILD_REF = os.path.join(REPO, "oracle", "build_ref.sh")

# name: (precision, SIM_MODE, N_X, N_Y, N_STEPS, {other FFB_* literals})
CASES = {
    # every mode, both precisions, odd non-square grid
    "sw_fp64_13x9": ("fp64", 1, 13, 9, 25, {"FFB_OMEGA": "1.3_FP"}),
    "cf_fp64_13x9": ("fp64", 2, 13, 9, 25, {"FFB_OMEGA": "1.3_FP"}),
    "pf_fp64_13x9": ("fp64", 3, 13, 9, 25, {"FFB_OMEGA": "1.3_FP", "FFB_RHO_IN": "1.001_FP",
                                             "FFB_RHO_OUT": "0.999_FP"}),
    "sl_fp64_13x9": ("fp64", 4, 13, 9, 25, {"FFB_OMEGA": "1.3_FP"}),
    "sw_fp32_13x9": ("fp32", 1, 13, 9, 25, {"FFB_OMEGA": "1.3_FP"}),
    "cf_fp32_13x9": ("fp32", 2, 13, 9, 25, {"FFB_OMEGA": "1.3_FP"}),
    "pf_fp32_13x9": ("fp32", 3, 13, 9, 25, {"FFB_OMEGA": "1.3_FP", "FFB_RHO_IN": "1.001_FP",
                                             "FFB_RHO_OUT": "0.999_FP"}),
    "sl_fp32_13x9": ("fp32", 4, 13, 9, 25, {"FFB_OMEGA": "1.3_FP"}),
    # single steps: pin the Poiseuille pre-pass order and the first lid step
    "pf_fp64_13x9_1step": ("fp64", 3, 13, 9, 1, {"FFB_RHO_IN": "1.001_FP", "FFB_RHO_OUT": "0.999_FP"}),
    "pf_fp64_13x9_2step": ("fp64", 3, 13, 9, 2, {"FFB_RHO_IN": "1.001_FP", "FFB_RHO_OUT": "0.999_FP"}),
    "sl_fp64_13x9_1step": ("fp64", 4, 13, 9, 1, {}),
    # the loop-form kernel (simulation.f90:76-168) must agree with the unrolled one
    "sl_fp64_13x9_loopkernel": ("fp64", 4, 13, 9, 25, {"FFB_OMEGA": "1.3_FP",
                                                       "FFB_USE_UNROLLED_KERNELS": ".false."}),
    # defaults of source/settings.f90 (omega 1.1, U_LID 0.1, RHO_IN/OUT 1.00005/0.99995)
    "sl_fp64_16x16_defaults": ("fp64", 4, 16, 16, 10, {}),
    # medium grids, longer runs
    "sl_fp64_48x40": ("fp64", 4, 48, 40, 200, {"FFB_OMEGA": "1.7_FP"}),
    "pf_fp64_40x24": ("fp64", 3, 40, 24, 300, {"FFB_OMEGA": "0.8_FP", "FFB_RHO_IN": "1.001_FP",
                                               "FFB_RHO_OUT": "0.999_FP"}),
    "cf_fp64_24x20": ("fp64", 2, 24, 20, 400, {"FFB_OMEGA": "1.5_FP", "FFB_U_WALL": "0.05_FP"}),
    "sw_fp32_32x32": ("fp32", 1, 32, 32, 100, {"FFB_OMEGA": "1.0_FP", "FFB_N_SIN": "1.0_FP"}),
    "sw_fp64_32x24_nsin3": ("fp64", 1, 32, 24, 60, {"FFB_OMEGA": "0.7_FP", "FFB_N_SIN": "3.0_FP",
                                                     "FFB_U_MAX": "0.05_FP"}),
}

DEFAULTS = {"FFB_RHO_0": "1.0_FP", "FFB_OMEGA": "1.1_FP", "FFB_U_MAX": "0.1_FP",
            "FFB_N_SIN": "2.0_FP", "FFB_U_WALL": "0.1_FP", "FFB_U_LID": "0.1_FP",
            "FFB_RHO_IN": "1.00005_FP", "FFB_RHO_OUT": "0.99995_FP"}


def literal_value(lit: str) -> float:
    return float(lit.replace("_FP", ""))


def run_case(name: str) -> dict:
    prec, mode, nx, ny, steps, extra = CASES[name]
    lits = dict(DEFAULTS)
    lits.update(extra)
    defs = [f"FFB_SIM_MODE={mode}", f"FFB_N_X={nx}", f"FFB_N_Y={ny}", f"FFB_N_STEPS={steps}"]
    defs += [f"{k}={v}" for k, v in lits.items()]
    env = dict(os.environ, REF_OPT="-O0", REF_RUN="1", REF_KEEP="0")
    subprocess.run(["bash", BUILD_REF, name, prec, *defs], check=True,
                   stdout=subprocess.DEVNULL, env=env)
    dt = np.float64 if prec == "fp64" else np.float32
    raw = np.fromfile(os.path.join(REPO, "oracle", "_ref", name, "state.bin"), dtype=dt)
    n = 9 * (nx + 2) * (ny + 2)
    f = raw[:n].reshape(9, ny + 2, nx + 2)[:, 1:-1, 1:-1]
    macros = raw[n:n + 3 * nx * ny].reshape(3, ny, nx)
    strips = raw[n + 3 * nx * ny:].reshape(2, 3, ny)
    return {
        "name": name, "precision": prec, "sim_mode": mode, "n_x": nx, "n_y": ny,
        "n_steps": steps, "literals": lits,
        "params": {k[4:].lower(): literal_value(v) for k, v in lits.items() if v.endswith("_FP")},
        "unrolled": lits.get("FFB_USE_UNROLLED_KERNELS", ".true.") == ".true.",
        "arrays": {"f": np.ascontiguousarray(f), "rho": macros[0], "u_x": macros[1],
                   "u_y": macros[2], "recv_macro": strips},
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--build-only", action="store_true",
                    help="build and run oracle/_ref cases without rewriting fixtures")
    ap.add_argument("cases", nargs="*")
    args = ap.parse_args()
    names = args.cases or list(CASES)
    manifest = []
    for name in names:
        rec = run_case(name)
        if not args.build_only:
            np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **rec["arrays"])
        rec = {k: v for k, v in rec.items() if k != "arrays"}
        rec["file"] = f"{name}.npz"
        manifest.append(rec)
        print("golden:", name)
    if not args.build_only and not args.cases:
        with open(os.path.join(HERE, "manifest.json"), "w") as fh:
            json.dump({"generator": "tests/golden/make_golden.py",
                       "source": "reference kernels compiled by oracle/build_ref.sh (flang -O0 "
                                 "-ffp-contract=off) with oracle/ref_driver.f90",
                       "cases": manifest}, fh, indent=1)


if __name__ == "__main__":
    main()
