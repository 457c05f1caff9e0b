"""Benchmark of the fused D2Q9 pull-stream-collide step on one MI355X.

Workload (default): the plain lid-driven cavity, 4096 x 4096 cells, halfway
bounce-back on three walls and a moving-wall lid at u_lid = 0.1, omega from
Re = 1000 (omega = 1/(3*0.1*4096/1000 + 1/2)), rho_0 = 1, fluid at rest,
FP32 populations.  One "step" is one fused pull-stream-collide update of the
whole domain; the library runs every form but the 1-3-12 layout as step
pairs, two updates per kernel launch (temporal blocking, bitwise the same
result), and the line also reports one launch per step ("single_pass").  Inputs are
resident in HBM before timing.

    python bench.py [--steps K] [--warmup W] [--precision fp32|fp64]

Prints ONE JSON line: MLUPS (= N_X*N_Y*steps/s/1e6, source/layouts/xyd/main.f90:326),
the HBM roofline of the step kernel (algorithmic 18*s bytes per cell update,
HIP events on the library's stream), an FP64 run of the same workload, a
copy-kernel ceiling, and the CPU baseline: the reference's own kernel
modules compiled by flang -O3 (oracle/_ref) on one host core, or, where that
binary is absent, the CPU oracle on all host cores.  Unless --no-extras, the
line also carries the deviation storage forms (8192^2 speed, 1024^2 error
against FP64), the 512-lattice batched form, the adjoint's backward/forward
cost and the 256^2 shear-wave case with its fitted viscosity.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
BYTES_PER_POP = {"fp64": 8, "fp32": 4, "fp16dev": 2, "bf16dev": 2, "fp32dev": 4, "e3m12dev": 2}


def cavity_omega(re: float = 1000.0, u: float = 0.1, n: int = 4096) -> float:
    return 1.0 / (3.0 * u * n / re + 0.5)


class pass_mode:
    """Lattices created inside run step pairs in one launch (pairs=True), one
    launch per step (False), or the library's default (None)."""

    def __init__(self, pairs: bool | None):
        self.pairs = pairs

    def __enter__(self):
        self.saved = os.environ.get("FFB_TWO_STEP")
        if self.pairs is not None:
            os.environ["FFB_TWO_STEP"] = "1" if self.pairs else "0"
        return self

    def __exit__(self, *exc):
        if self.saved is None:
            os.environ.pop("FFB_TWO_STEP", None)
        else:
            os.environ["FFB_TWO_STEP"] = self.saved


def uses_pairs(precision: str) -> bool:
    """Whether the library runs step pairs for this form (FFB_TWO_STEP or its default)."""
    env = os.environ.get("FFB_TWO_STEP")
    if env is not None:
        try:
            return int(env) != 0
        except ValueError:
            return False
    return precision != "e3m12dev"


def batch_uses_pairs(precision: str) -> bool:
    """Whether a batch runs step pairs (FFB_BATCH_PAIRS or its default: FP32,
    FP64 and FP32 deviations); the autotuner then tunes the single-step kernel
    that runs odd remainders."""
    env = os.environ.get("FFB_BATCH_PAIRS")
    if env:
        try:
            return int(env) != 0 and precision in ("fp32", "fp64", "fp32dev")
        except ValueError:
            pass
    return precision in ("fp32", "fp64", "fp32dev")


def run_gpu(n: int, precision: str, steps: int, warmup: int, scale_log2: int = 0,
            pairs: bool | None = None) -> dict:
    from fastfortranboltzmann_amd import Lattice, Settings

    s = Settings(N_X=n, N_Y=n, SIM_MODE=4, OMEGA=cavity_omega(n=n), U_LID=0.1, RHO_0=1.0,
                 precision=precision, storage_scale_log2=scale_log2)
    with pass_mode(pairs):
        paired = uses_pairs(precision)
        lat = Lattice(s)
    with lat:
        lat.init()
        lat.step(warmup)
        lat.sync()
        lat.reset_timer()
        t0 = time.perf_counter()
        lat.step(steps)
        lat.sync()
        wall = time.perf_counter() - t0
        ev_ms = lat.elapsed_ms()
        mass = lat.moments()[0]
    cells = n * n
    # launches: step pairs (and a single step for an odd count), or one per step
    launches = steps // 2 + steps % 2 if paired else steps
    per_launch = 2 if paired and steps % 2 == 0 else 1
    kernel_ms = ev_ms / launches  # the kernel's average launch duration
    alg_bytes = 18 * BYTES_PER_POP[precision] * cells  # per cell update (step)
    achieved = alg_bytes * steps / (ev_ms * 1e-3) / 1e9
    return {
        "mlups": cells * steps / wall / 1e6,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms": kernel_ms,
        "steps_per_launch": per_launch,
        "alg_bytes_per_launch": alg_bytes * per_launch,
        "achieved_gbs": achieved,
        "mass_per_cell": mass / cells,
    }


def reduced_storage(n: int, steps: int, err_n: int = 1024, err_steps: int = 2000) -> dict:
    """Deviation storage forms: MLUPS at n^2 (cavity, Re 5000) against FP32 at
    the same size, and the relative L2 velocity error and mass drift against
    FP64 on a smaller cavity run (err_n^2, Re 1000, err_steps steps), for the
    16-bit forms (FP16, BF16, the 1-3-12 layout), FP32 deviations and plain
    FP32; for each form also the density
    error and the relative error of the shear-wave viscosity fitted in that form
    (256^2, omega 1, 2000 steps) against nu = (1/omega - 1/2)/3."""
    from fastfortranboltzmann_amd.precision import (cavity_settings, compare, default_scale_log2,
                                                    run_form, shear_wave_viscosity)

    forms = ("fp32", "fp32dev", "fp16dev", "bf16dev", "e3m12dev")
    out = {"size": n, "steps": steps, "error_case": f"cavity {err_n}^2 Re 1000, {err_steps} steps"}
    rates = {}
    for form in forms:
        s = default_scale_log2(form) if form.endswith("dev") else 0
        rates[form] = run_form(cavity_settings(n, 5000.0, 0.1, form, s), steps, chunk=steps)["mlups"]
    with pass_mode(False):  # FP32 with one launch per step, as the 16-bit forms run
        fp32_single = run_form(cavity_settings(n, 5000.0, 0.1, "fp32", 0), steps, chunk=steps)["mlups"]
    ref = run_form(cavity_settings(err_n, 1000.0, 0.1, "fp64"), err_steps, chunk=err_steps)
    for form in forms:
        s = default_scale_log2(form) if form.endswith("dev") else 0
        r = run_form(cavity_settings(err_n, 1000.0, 0.1, form, s), err_steps, chunk=err_steps)
        c = compare(ref, r)
        sw = shear_wave_viscosity(form)
        out[form] = {"mlups": round(rates[form], 1), "speedup_vs_fp32": round(rates[form] / rates["fp32"], 3),
                     "speedup_vs_fp32_single_pass": round(rates[form] / fp32_single, 3),
                     "step_pairs": uses_pairs(form),
                     "achieved_gbs": round(rates[form] * 1e6 * 18 * BYTES_PER_POP[form] / 1e9, 1),
                     "scale_log2": s, "rel_l2_u_vs_fp64": c["rel_l2_u"], "max_abs_rho_vs_fp64": c["max_abs_rho"],
                     "rel_l2_rho_dev_vs_fp64": c["rel_l2_rho_dev"], "mass_drift": r["mass_drift"],
                     "shear_wave_nu_rel_err": sw["nu_rel_err"],
                     "rounding": "stochastic" if form in ("fp16dev", "bf16dev", "e3m12dev") else "nearest"}
    out["fp64_mass_drift"] = ref["mass_drift"]
    out["fp32_mlups"] = round(rates["fp32"], 1)
    out["fp32_single_pass_mlups"] = round(fp32_single, 1)
    return out


def shear_wave_256(steps: int = 2000, chunk: int = 200) -> dict:
    """The reference's CPU-runnable case on the GPU: shear-wave decay 256^2,
    FP64, omega 1, u_max 0.05, n_sin 1, 2000 steps.  MLUPS (cache-resident at
    this size, not graded against HBM) and the decay-rate viscosity fitted
    from the amplitude every `chunk` steps against nu = (1/omega - 1/2)/3."""
    import numpy as np

    from fastfortranboltzmann_amd import Lattice, Settings
    from fastfortranboltzmann_amd.analysis import decay_amplitude, fit_viscosity, lattice_viscosity

    n = 256
    s = Settings(N_X=n, N_Y=n, SIM_MODE=1, OMEGA=1.0, U_MAX=0.05, N_SIN=1.0, RHO_0=1.0, precision="fp64")
    amps, times, kernel_ms = [], [], 0.0
    with Lattice(s) as lat:
        lat.init()
        for t in range(chunk, steps + 1, chunk):
            lat.reset_timer()
            lat.step(chunk, write_macro_last=True)
            lat.sync()
            kernel_ms += lat.elapsed_ms()
            _, ux, _ = lat.macros()
            amps.append(decay_amplitude(ux, 1.0))
            times.append(t)
    nu = lattice_viscosity(1.0)
    nu_fit = fit_viscosity(np.array(times), np.array(amps), n, 1.0)
    return {"size": n, "steps": steps, "mlups": round(n * n * steps / (kernel_ms * 1e-3) / 1e6, 1),
            "nu_fit": nu_fit, "nu": nu, "nu_rel_err": abs(nu_fit - nu) / nu,
            "amplitude_ratio": amps[-1] / 0.05}


def batched(single_fp32_mlups: float, single_pass_fp32_mlups: float, steps: int = 200) -> dict:
    """512 mixed lattices (sides 64..1024, omega 0.6..1.9, shear wave /
    Couette / body-force channel), FP32, autotuned persistent launch; the same
    batch in the 1-3-12 deviation form (each lattice at its default scale,
    stochastic rounding) beside it."""
    from fastfortranboltzmann_amd.batch import Batch, mixed_batch

    lats = mixed_batch(0, 512)
    out = {}
    for form in ("fp32", "e3m12dev"):
        with Batch(lats, form) as b:
            tun = b.autotune(20)
            info = b.info()
            b.init()
            b.step(2)
            b.sync()
            b.reset_timer()
            b.step(steps)
            b.sync()
            ms = b.elapsed_ms()
        out[form] = (info["cells"] * steps / (ms * 1e-3) / 1e6, tun)
    mlups, tun = out["fp32"]
    return {"lattices": len(lats), "cells": info["cells"], "steps": steps, "mlups": round(mlups, 1),
            "vs_single_4096_fp32": round(mlups / single_fp32_mlups, 3),
            "vs_single_pass_4096_fp32": round(mlups / single_pass_fp32_mlups, 3),
            "step_pairs": batch_uses_pairs("fp32"),
            "tuning": {"tile_w": tun["tile_w"], "grid_blocks": tun["grid_blocks"]},
            "e3m12dev": {"mlups": round(out["e3m12dev"][0], 1),
                         "vs_batch_fp32": round(out["e3m12dev"][0] / mlups, 3),
                         "tuning": {"tile_w": out["e3m12dev"][1]["tile_w"],
                                    "grid_blocks": out["e3m12dev"][1]["grid_blocks"]}}}


def adjoint_cost(n: int = 512, steps: int = 2000) -> dict:
    """Body-force channel n^2, FP32: forward (with checkpoints every 100 steps)
    and reverse sweep (recomputation + adjoint) device times."""
    import numpy as np

    from fastfortranboltzmann_amd.adjoint import Adjoint, channel_settings

    with Adjoint(channel_settings(n, 1.2, 1e-6, "fp32"), steps, 100) as a:
        for _ in range(2):
            _, ux, uy = a.forward()
            a.backward(np.ones_like(ux), np.zeros_like(uy))
        f_ms, b_ms = a.timing()
    return {"size": n, "steps": steps, "forward_ms": round(f_ms, 2), "backward_ms": round(b_ms, 2),
            "backward_over_forward": round(b_ms / f_ms, 3),
            "forward_mlups": round(n * n * steps / (f_ms * 1e-3) / 1e6, 1)}


def copy_ceiling() -> float | None:
    from fastfortranboltzmann_amd import _lib
    import ctypes

    L = _lib.lib()
    if not hasattr(L, "ffb_bench_copy"):
        return None
    gbs = ctypes.c_double()
    if L.ffb_bench_copy(0, 1 << 30, 20, ctypes.byref(gbs)) != 0:
        return None
    return gbs.value


def traffic_from_profiles(workload: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    as recorded in profiles/traffic.txt ("<workload> <fetch_bytes> <write_bytes>")."""
    path = os.path.join(REPO, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 3 and parts[0] == workload:
            try:
                return float(parts[1]) + float(parts[2])
            except ValueError:
                return None
    return None


def cpu_baseline(precision: str) -> dict | None:
    exe = os.path.join(REPO, "oracle", "_ref", f"bench_sl4096_{precision}", "ref_driver")
    if os.path.exists(exe):
        try:
            out = subprocess.run([exe, "-"], capture_output=True, text=True, timeout=300).stdout
            m = re.search(r"steps=(\d+).*MLUPS=\s*([0-9.eE+-]+)", out)
            if m:
                return {"value": float(m.group(2)), "unit": "MLUPS", "cores": 1, "kind": "reference",
                        "sample": f"{m.group(1)} steps of the same 4096^2 cavity ({precision}); the "
                                  "reference's kernel modules (simulation.f90, modes/*.f90) "
                                  "compiled by flang -O3, one thread = one coarray image"}
        except (subprocess.TimeoutExpired, OSError):
            pass
    return cpu_oracle_threads(precision, 10.0)


def host_cpu() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_oracle_threads(precision: str, seconds: float = 6.0) -> dict:
    """The CPU oracle (the reference restated in C, same operation order) on up
    to 16 host threads (OpenMP over rows), timed for about `seconds` on the
    same 4096^2 cavity."""
    import numpy as np

    from oracle.oracle import OracleLattice, Params, max_threads, set_threads

    threads = min(16, os.cpu_count() or 1)
    set_threads(threads)
    o = OracleLattice(4, 4096, 4096, Params(omega=cavity_omega(), u_lid=0.1),
                      np.float64 if precision == "fp64" else np.float32).init()
    o.run(1)
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.run(1)
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": round(4096 * 4096 * steps / dt / 1e6, 1), "unit": "MLUPS", "cores": max_threads(),
            "kind": "port", "host": host_cpu(),
            "sample": f"{steps} steps of the same 4096^2 cavity ({precision}), CPU oracle, OpenMP"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--precision", default="fp32", choices=list(BYTES_PER_POP))
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 run of the workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the reduced-storage, batched and adjoint measurements")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the domain is one device-resident "
              "lattice and no multi-GPU path is built (run with --gpus 1)", file=sys.stderr)
        return 2

    n = args.size
    print(f"bench: cavity {n}^2 {args.precision}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    main_run = run_gpu(n, args.precision, args.steps, args.warmup)
    print("bench: one launch per step", file=sys.stderr, flush=True)
    single = run_gpu(n, args.precision, args.steps, args.warmup, pairs=False)
    extra = {"steps_per_launch": main_run["steps_per_launch"],
             "single_pass": {"mlups": round(single["mlups"], 1), "kernel_ms": round(single["kernel_ms"], 5),
                             "achieved_gbs": round(single["achieved_gbs"], 1),
                             "frac": round(single["achieved_gbs"] / HBM_PEAK_GBS, 4),
                             "traffic": traffic_from_profiles(f"sl{n}_{args.precision}_single_pass")}}
    if not args.no_fp64 and args.precision != "fp64":
        print("bench: fp64 run", file=sys.stderr, flush=True)
        r64 = run_gpu(n, "fp64", max(2, args.steps // 2), args.warmup)
        s64 = run_gpu(n, "fp64", max(2, args.steps // 2), args.warmup, pairs=False)
        extra["fp64"] = {"mlups": round(r64["mlups"], 1), "achieved_gbs": round(r64["achieved_gbs"], 1),
                         "frac": round(r64["achieved_gbs"] / HBM_PEAK_GBS, 4),
                         "kernel_ms": round(r64["kernel_ms"], 5), "steps_per_launch": r64["steps_per_launch"],
                         "single_pass": {"mlups": round(s64["mlups"], 1),
                                         "achieved_gbs": round(s64["achieved_gbs"], 1),
                                         "frac": round(s64["achieved_gbs"] / HBM_PEAK_GBS, 4),
                                         "kernel_ms": round(s64["kernel_ms"], 5)}}
    if not args.no_extras and args.precision == "fp32":
        print("bench: reduced storage", file=sys.stderr, flush=True)
        extra["reduced_storage"] = reduced_storage(2 * n if n <= 4096 else n, max(50, args.steps // 10))
        print("bench: batched", file=sys.stderr, flush=True)
        extra["batched"] = batched(main_run["mlups"], single["mlups"])
        print("bench: adjoint", file=sys.stderr, flush=True)
        extra["adjoint"] = adjoint_cost()
        print("bench: shear wave 256^2", file=sys.stderr, flush=True)
        extra["shear_wave_256_fp64"] = shear_wave_256()
    cc = copy_ceiling()
    workload = f"sl{n}_{args.precision}"
    traffic = traffic_from_profiles(workload)
    cpu = None
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.precision)
        if cpu is not None:
            cpu["host"] = host_cpu()
    line = {
        "metric": "MLUPS + % HBM roofline (plain FP64/FP32 vs reduced/batched/adjoint) at stated error",
        "value": round(main_run["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp64": "f64", "fp16dev": "f16-storage/f32", "bf16dev": "bf16-storage/f32",
                  "fp32dev": "f32-deviation", "e3m12dev": "e3m12-storage/f32"}[args.precision],
        "data": "synthetic (fluid at rest, deterministic initial state)",
        "config": {"workload": f"lid-driven cavity {n}x{n}, bounce-back walls + moving lid u=0.1, "
                               f"Re=1000 (omega={cavity_omega(n=n):.6f}), {args.precision} populations",
                   "n_x": n, "n_y": n, "sim_mode": "sliding_lid", "precision": args.precision},
        "roofline": {"bound": "hbm", "achieved": round(main_run["achieved_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(main_run["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "kernel_ms": round(main_run["kernel_ms"], 5),
        "copy_ceiling_gbs": round(cc, 1) if cc else None,
        **extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
