"""Deviation storage forms of the solver: error against FP64 and speed.

The populations are stored as D_i = (f_i - w_i) * 2^s in FP16, BF16 or FP32,
the step (boundary rules and collision) runs on D in FP32, and the
conservation sums (mass, momentum) are taken in FP64 (``Lattice.moments``).
These forms have no counterpart in the reference; their results are stated
against an FP64 run of the same initial state:

  * relative L2 error of the velocity field, ||u - u64|| / ||u64||
  * maximum absolute velocity error
  * density error: max |rho - rho64| and ||rho - rho64|| / ||rho64 - 1||
  * total-mass drift (m(t) - m(0)) / m(0), FP64 sums
  * the shear-wave viscosity fitted from the decay of a sine wave in the same
    form, against nu = (1/omega - 1/2)/3

    python -m fastfortranboltzmann_amd.precision --n 8192 --re 5000 --steps 50000
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from .lattice import Lattice, Settings, omega_for_reynolds

FP16_MAX = 65504.0


def default_scale_log2(precision: str, u_ref: float = 0.1) -> int:
    """Scale 2^s for the deviation store for flows with largest speed u_ref
    (ffb_default_scale_log2 of the library, host arithmetic only): FP16 puts
    the largest expected deviation (4/9)(0.05 + 3u + 6u^2) near 2^14, a factor
    4 below the FP16 maximum, so the small deviations of the resting bulk stay
    in the normal range; the 1-3-12 form puts the largest per-direction-class
    bound at 3/4 of its largest magnitude (s = 5 for u = 0.1); BF16 and FP32
    deviations need no scale (s = 0)."""
    from ._lib import lib
    from .lattice import PRECISIONS

    if precision not in PRECISIONS or not precision.endswith("dev"):
        raise ValueError(f"not a deviation form: {precision}")
    s = int(lib().ffb_default_scale_log2(PRECISIONS[precision], float(u_ref)))
    if s == -2 ** 31:
        raise ValueError(f"no default scale for {precision} at speed {u_ref}")
    return s


def cavity_settings(n: int, re: float, u_lid: float, precision: str, scale_log2: int = 0) -> Settings:
    return Settings(N_X=n, N_Y=n, SIM_MODE=4, OMEGA=omega_for_reynolds(re, u_lid, n), U_LID=u_lid,
                    RHO_0=1.0, precision=precision, storage_scale_log2=scale_log2)


def run_form(settings: Settings, steps: int, chunk: int = 5000):
    """Runs one form; returns (u_x, u_y) as float64 host arrays, timing and the
    mass history (FP64 sums)."""
    with Lattice(settings) as lat:
        lat.init()
        m0 = lat.moments()[0]
        masses = [m0]
        done, kernel_ms = 0, 0.0
        t0 = time.perf_counter()
        while done < steps:
            k = min(chunk, steps - done)
            lat.reset_timer()
            lat.step(k, write_macro_last=(done + k == steps))
            lat.sync()
            kernel_ms += lat.elapsed_ms()
            done += k
            masses.append(lat.moments()[0])
        wall = time.perf_counter() - t0
        rho, ux, uy = lat.macros()
    cells = settings.N_X * settings.N_Y
    return {
        "rho": rho.astype(np.float64), "u_x": ux.astype(np.float64), "u_y": uy.astype(np.float64),
        "mlups": cells * steps / (kernel_ms * 1e-3) / 1e6,
        "wall_s": wall,
        "mass_drift": (masses[-1] - m0) / m0,
        "mass_drift_max": max(abs(m - m0) for m in masses) / m0,
    }


def compare(ref: dict, run: dict) -> dict:
    du = np.sqrt((run["u_x"] - ref["u_x"]) ** 2 + (run["u_y"] - ref["u_y"]) ** 2)
    norm = math.sqrt(float(np.sum(ref["u_x"] ** 2 + ref["u_y"] ** 2)))
    drho = run["rho"] - ref["rho"]
    dev = math.sqrt(float(np.sum((ref["rho"] - 1.0) ** 2)))
    return {"rel_l2_u": float(math.sqrt(float(np.sum(du ** 2))) / norm),
            "max_abs_u": float(du.max()),
            "u_max_ref": float(np.sqrt(ref["u_x"] ** 2 + ref["u_y"] ** 2).max()),
            "max_abs_rho": float(np.abs(drho).max()),
            "rel_l2_rho_dev": float(math.sqrt(float(np.sum(drho ** 2))) / dev) if dev > 0 else 0.0}


def shear_wave_viscosity(precision: str, n: int = 256, omega: float = 1.0, u_max: float = 0.05,
                         steps: int = 2000, chunk: int = 200, scale_log2: int | None = None) -> dict:
    """Shear-wave decay u_x = u_max sin(2 pi y / n) in one storage form: the
    amplitude every `chunk` steps, the viscosity fitted from its decay, and its
    relative error against nu = (1/omega - 1/2)/3."""
    from .analysis import decay_amplitude, fit_viscosity, lattice_viscosity

    if scale_log2 is None:
        scale_log2 = default_scale_log2(precision, u_max) if precision.endswith("dev") else 0
    s = Settings(N_X=n, N_Y=n, SIM_MODE=1, OMEGA=omega, U_MAX=u_max, N_SIN=1.0, RHO_0=1.0,
                 precision=precision, storage_scale_log2=scale_log2)
    amps, times = [], []
    with Lattice(s) as lat:
        lat.init()
        for t in range(chunk, steps + 1, chunk):
            lat.step(chunk, write_macro_last=True)
            _, ux, _ = lat.macros()
            amps.append(decay_amplitude(ux.astype(np.float64), 1.0))
            times.append(t)
    nu = lattice_viscosity(omega)
    nu_fit = fit_viscosity(np.array(times), np.array(amps), n, 1.0)
    return {"nu_fit": nu_fit, "nu": nu, "nu_rel_err": abs(nu_fit - nu) / nu, "n": n, "omega": omega,
            "steps": steps}


def precision_study(n: int, re: float, steps: int, u_lid: float = 0.1,
                    forms=("fp16dev", "bf16dev", "fp32", "fp32dev", "e3m12dev"), scales: dict | None = None,
                    log=print) -> dict:
    out = {"n": n, "re": re, "u_lid": u_lid, "steps": steps,
           "omega": omega_for_reynolds(re, u_lid, n), "forms": {}}
    log(f"precision: fp64 reference {n}^2, {steps} steps")
    ref = run_form(cavity_settings(n, re, u_lid, "fp64"), steps)
    out["forms"]["fp64"] = {"mlups": ref["mlups"], "mass_drift": ref["mass_drift"],
                            "mass_drift_max": ref["mass_drift_max"]}
    for form in forms:
        s = (scales or {}).get(form)
        if s is None:
            s = default_scale_log2(form, u_lid) if form.endswith("dev") else 0
        log(f"precision: {form} (s = {s})")
        r = run_form(cavity_settings(n, re, u_lid, form, s), steps)
        rec = {"scale_log2": s, "mlups": r["mlups"], "mass_drift": r["mass_drift"],
               "mass_drift_max": r["mass_drift_max"], **compare(ref, r),
               "shear_wave": shear_wave_viscosity(form)}
        out["forms"][form] = rec
        log(f"precision: {form}: " + json.dumps({k: v for k, v in rec.items()}))
    if "fp32" in out["forms"]:
        for form in forms:
            if form.endswith("dev"):
                out["forms"][form]["speedup_vs_fp32"] = out["forms"][form]["mlups"] / out["forms"]["fp32"]["mlups"]
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="deviation storage forms: error and speed against FP64")
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--re", type=float, default=5000.0)
    ap.add_argument("--u-lid", type=float, default=0.1)
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--forms", default="fp16dev,bf16dev,fp32,fp32dev,e3m12dev")
    ap.add_argument("--scale", action="append", default=[], help="form=s, e.g. fp16dev=16")
    args = ap.parse_args(argv)
    scales = {k: int(v) for k, v in (x.split("=") for x in args.scale)}
    res = precision_study(args.n, args.re, args.steps, args.u_lid, tuple(args.forms.split(",")),
                          scales, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
