"""Benchmark of the fused D2Q9 pull-stream-collide step on one MI355X.

Workload (default): the plain lid-driven cavity, 4096 x 4096 cells, halfway
bounce-back on three walls and a moving-wall lid at u_lid = 0.1, omega from
Re = 1000 (omega = 1/(3*0.1*4096/1000 + 1/2)), rho_0 = 1, fluid at rest,
FP32 populations.  One "step" is one fused pull-stream-collide update of the
whole domain; the library runs consecutive steps five per kernel launch
(FP32; four for FP64; temporal blocking, bitwise the same result; fewer for
a remainder), and the line also reports step pairs ("pairs") and one launch
per step ("single_pass").  Inputs are resident in
HBM before timing.

    python bench.py [--steps K] [--warmup W] [--precision fp32|fp64|...]

Prints ONE JSON line: MLUPS (= N_X*N_Y*steps/s/1e6, source/layouts/xyd/main.f90:326)
and the HBM roofline of the kernel that runs the workload.  `roofline.achieved`
is the bytes one launch must move -- one read and one write of the stored
state, 18*s bytes per cell, whatever the number of updates the launch
performs -- over the launch's average duration (HIP events on the library's
stream, over whole launches); `effective_gbs` is the per-update rate
18*s*cells*steps/time, above the peak for multi-step launches, and not a
roofline figure.  `roofline.traffic` is the
HBM traffic per launch measured by rocprofv3 (profiles/traffic.txt), reported
only when that file was written from the kernel sources being benched.  The
line also carries an FP64 run of the workload, a copy-kernel ceiling, and the
CPU baseline: the reference's own kernel modules compiled by flang
(oracle/_ref) on the host's cores.  Unless --no-extras, it also carries the
deviation storage forms (8192^2 speed, error against FP64), the 512-lattice
batched form, the adjoint's backward/forward cost and the 256^2 shear-wave
case with its fitted viscosity.

Profiling (repo root, GPU host):
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir_f> -o run -- python3 bench.py --profile-run
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir_w> -o run -- python3 bench.py --profile-run
    python3 bench.py --record-traffic <dir_f>/run_counter_collection.csv <dir_w>/run_counter_collection.csv
writes profiles/traffic.txt for the current kernel sources.
"""
from __future__ import annotations

import argparse
import ctypes
import csv
import hashlib
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
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic.txt")
MARK_START, MARK_END = 1 << 20, 2 << 20  # bytes of the marker copies of --profile-run
KERNEL_SOURCES = ("d2q9_kernels.h", "ffb_internal.h", "ffb_api.hip", "ffb_batch.hip", "ffb_adjoint.hip")
CPU_THREADS = 16  # the host share of one GPU on the MI355X boxes


def cavity_omega(re: float = 1000.0, u: float = 0.1, n: int = 4096) -> float:
    return 1.0 / (3.0 * u * n / re + 0.5)


def kernel_source_hash() -> str:
    """Hash of the HIP sources libffb_hip.so is built from (stamps profiles/traffic.txt)."""
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(REPO, "fastfortranboltzmann_amd", "csrc", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


class pass_mode:
    """Lattices created inside run k steps per launch (1 = one launch per
    step, 2 = step pairs, up to 5), or the library's default (None)."""

    KEYS = ("FFB_STEPS_PER_PASS", "FFB_TWO_STEP")

    def __init__(self, k: int | None):
        self.k = k

    def __enter__(self):
        self.saved = {key: os.environ.get(key) for key in self.KEYS}
        if self.k is not None:
            os.environ["FFB_STEPS_PER_PASS"] = str(int(self.k))
            os.environ["FFB_TWO_STEP"] = "1" if self.k >= 2 else "0"
        return self

    def __exit__(self, *exc):
        for key, val in self.saved.items():
            if val is None:
                os.environ.pop(key, None)
            else:
                os.environ[key] = val


def default_steps_per_pass(precision: str) -> int:
    """The library's default steps per launch (FFB_STEPS_PER_PASS unset): five
    for FP32 and FP32 deviations, four for FP64, two for the 1-3-12
    deviations, three for FP16 and BF16 deviations."""
    return {"fp32": 5, "fp32dev": 5, "fp64": 4, "e3m12dev": 2}.get(precision, 3)


def steps_per_pass(precision: str) -> int:
    """Steps the library computes per launch for this form: FFB_STEPS_PER_PASS
    (1..5; five only for the 4-byte forms, four only for those with a four-step
    kernel), 1 where FFB_TWO_STEP=0, else the library's default."""
    env = os.environ.get("FFB_TWO_STEP")
    if env is not None:
        try:
            if int(env) == 0:
                return 1
        except ValueError:
            return 1
    env = os.environ.get("FFB_STEPS_PER_PASS")
    if env is not None:
        try:
            k = max(1, min(5, int(env)))
        except ValueError:
            return 1
        four = precision in ("fp32", "fp64", "fp32dev", "fp16dev")  # forms with a four-step kernel
        five = precision in ("fp32", "fp32dev")  # forms with a five-step kernel
        if k == 5 and not five:
            k = 4
        return k if (k < 4 or four) else 3
    return default_steps_per_pass(precision) if precision in BYTES_PER_POP else 1


def batch_steps_per_pass(precision: str) -> int:
    """Steps a batch computes per launch, as ffb_batch_create chooses them:
    multi-step launches for every form unless FFB_BATCH_PAIRS says
    otherwise, five steps each for the 4-byte forms and three for the others
    unless FFB_BATCH_STEPS_PER_PASS (1..5; four and five only for the 4-byte forms)
    says otherwise (which alone, without FFB_BATCH_PAIRS, also selects
    multi-step launches for any form); 1 = one launch per step."""
    pairs_env = os.environ.get("FFB_BATCH_PAIRS")
    k_env = os.environ.get("FFB_BATCH_STEPS_PER_PASS")
    pairs = True
    if pairs_env:
        try:
            pairs = int(pairs_env) != 0
        except ValueError:
            pairs = False
    four = BYTES_PER_POP.get(precision, 0) == 4
    k = (5 if four else 3) if pairs else 1
    if k_env:
        try:
            k = int(k_env)
        except ValueError:
            pass
        if not pairs_env:
            pairs = k >= 2
    return max(2, min(5 if four else 3, k)) if pairs else 1


def hbm_rates(mlups: float, precision: str, per_launch: int) -> dict:
    """For a run at `mlups` with `per_launch` updates per launch: the rate at
    which its launches move the bytes they must move (one read and one write
    of the state per launch), that rate over the 8 TB/s peak, and the
    per-update 'effective' rate (18*s bytes per cell update)."""
    eff = mlups * 1e6 * 18 * BYTES_PER_POP[precision] / 1e9
    return {"min_bytes_gbs": round(eff / per_launch, 1), "frac": round(eff / per_launch / HBM_PEAK_GBS, 4),
            "effective_gbs": round(eff, 1)}


def run_gpu(n: int, precision: str, steps: int, warmup: int, scale_log2: int = 0,
            per_pass: int | None = None) -> dict:
    """The cavity n^2 for `steps` steps after `warmup`: MLUPS from the wall
    clock around the timed steps (inputs resident, synchronised on both
    sides), and the kernel's average launch duration from HIP events on the
    library stream over a run of whole launches (a multiple of the steps per
    launch, so that no remainder launch of another kernel is mixed in)."""
    from fastfortranboltzmann_amd import Lattice, Settings

    s = Settings(N_X=n, N_Y=n, SIM_MODE=4, OMEGA=cavity_omega(n=n), U_LID=0.1, RHO_0=1.0,
                 precision=precision, storage_scale_log2=scale_log2)
    with pass_mode(per_pass):
        k = steps_per_pass(precision)
        lat = Lattice(s)
    with lat:
        lat.init()
        lat.step(warmup)
        lat.sync()
        t0 = time.perf_counter()
        lat.step(steps)
        lat.sync()
        wall = time.perf_counter() - t0
        ev_steps = max(k, (steps // k) * k)
        lat.reset_timer()
        lat.step(ev_steps)
        lat.sync()
        ev_ms = lat.elapsed_ms()
    cells = n * n
    launches = ev_steps // k
    kernel_ms = ev_ms / launches  # the kernel's average launch duration
    must_move = 18 * BYTES_PER_POP[precision] * cells  # one read + one write of the state
    return {
        "mlups": cells * steps / wall / 1e6,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms": kernel_ms,
        "steps_per_launch": k,
        "bytes_per_launch": must_move,
        "achieved_gbs": must_move / (kernel_ms * 1e-3) / 1e9,
        "effective_gbs": must_move * ev_steps / (ev_ms * 1e-3) / 1e9,
    }


def roofline(run: dict, workload: str) -> dict:
    """Roofline block of a run_gpu result: achieved = bytes the launch must
    move / launch duration; traffic = measured HBM bytes per launch of the
    same kernel (null unless profiles/traffic.txt matches these sources)."""
    tr = traffic_from_profiles(workload)
    out = {"bound": "hbm", "achieved": round(run["achieved_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(run["achieved_gbs"] / HBM_PEAK_GBS, 4),
           "traffic": tr["bytes"] if tr else None,
           "bytes_per_launch": run["bytes_per_launch"], "steps_per_launch": run["steps_per_launch"]}
    if tr:
        out["traffic_over_min"] = round(tr["bytes"] / run["bytes_per_launch"], 4)
        out["traffic_gbs"] = round(tr["bytes"] / (run["kernel_ms"] * 1e-3) / 1e9, 1)
        out["kernel"] = tr["kernel"]
    else:
        out["traffic_note"] = traffic_status()
    va = valu_from_profiles(workload)
    if va:
        # share of the SIMDs' cycles the launch's VALU instructions take (a
        # wave64 instruction takes a SIMD 4 cycles): the bound of launches
        # that trade HBM traffic for arithmetic
        out["valu"] = {"wave_insts_per_launch": va["wave_insts_per_launch"], "busy": va["busy"]}
        # what holds the launch back: the vector ALUs when their instructions fill
        # the SIMDs' cycles, HBM when the measured traffic runs at 0.7 of the
        # peak or more (the copy ceiling is about 0.84), else neither
        if va["busy"] >= 0.9:
            out["limiter"] = "valu"
        elif tr and tr["bytes"] / (run["kernel_ms"] * 1e-3) / 1e9 >= 0.7 * HBM_PEAK_GBS:
            out["limiter"] = "hbm"
        else:
            out["limiter"] = "neither"
    return out


def reduced_storage(n: int, steps: int, err_n: int = 1024, err_steps: int = 2000, big_err_steps: int = 5000) -> dict:
    """Deviation storage forms: MLUPS at n^2 (cavity, Re 5000) against FP32 at
    the same size (both in step pairs, the default; and against FP32 with one
    launch per step), with the bytes their launches must move; the relative
    L2 velocity error, density error and mass drift against FP64 on a 1024^2
    cavity (Re 1000, err_steps steps) and on the n^2 Re 5000 cavity itself
    after big_err_steps steps; and the relative error of the shear-wave
    viscosity fitted in each form (256^2, omega 1, 2000 steps) against
    nu = (1/omega - 1/2)/3."""
    from fastfortranboltzmann_amd.precision import (cavity_settings, compare, default_scale_log2,
                                                    run_form, shear_wave_viscosity)

    forms = ("fp32", "fp32dev", "fp16dev", "bf16dev", "e3m12dev")
    out = {"size": n, "steps": steps, "error_case": f"cavity {err_n}^2 Re 1000, {err_steps} steps",
           "error_case_big": f"cavity {n}^2 Re 5000, {big_err_steps} steps"}
    scale = {f: (default_scale_log2(f) if f.endswith("dev") else 0) for f in forms}
    rates = {f: run_form(cavity_settings(n, 5000.0, 0.1, f, scale[f]), steps, chunk=steps)["mlups"] for f in forms}
    with pass_mode(1):  # FP32 with one launch per step
        fp32_single = run_form(cavity_settings(n, 5000.0, 0.1, "fp32", 0), steps, chunk=steps)["mlups"]
    with pass_mode(2):  # FP32 in step pairs
        fp32_pair = run_form(cavity_settings(n, 5000.0, 0.1, "fp32", 0), steps, chunk=steps)["mlups"]
    ref = run_form(cavity_settings(err_n, 1000.0, 0.1, "fp64"), err_steps, chunk=err_steps)
    ref_big = run_form(cavity_settings(n, 5000.0, 0.1, "fp64"), big_err_steps, chunk=big_err_steps)
    for form in forms:
        r = run_form(cavity_settings(err_n, 1000.0, 0.1, form, scale[form]), err_steps, chunk=err_steps)
        c = compare(ref, r)
        rb = run_form(cavity_settings(n, 5000.0, 0.1, form, scale[form]), big_err_steps, chunk=big_err_steps)
        cb = compare(ref_big, rb)
        sw = shear_wave_viscosity(form)
        per_launch = steps_per_pass(form)
        out[form] = {"mlups": round(rates[form], 1), "speedup_vs_fp32": round(rates[form] / rates["fp32"], 3),
                     "speedup_vs_fp32_pairs": round(rates[form] / fp32_pair, 3),
                     "speedup_vs_fp32_single_pass": round(rates[form] / fp32_single, 3),
                     "steps_per_launch": per_launch, **hbm_rates(rates[form], form, per_launch),
                     "scale_log2": scale[form], "rel_l2_u_vs_fp64": c["rel_l2_u"],
                     "max_abs_rho_vs_fp64": c["max_abs_rho"], "rel_l2_rho_dev_vs_fp64": c["rel_l2_rho_dev"],
                     "mass_drift": r["mass_drift"],
                     "big": {"rel_l2_u_vs_fp64": cb["rel_l2_u"], "max_abs_u_vs_fp64": cb["max_abs_u"],
                             "max_abs_rho_vs_fp64": cb["max_abs_rho"], "mass_drift": rb["mass_drift"]},
                     "shear_wave_nu_rel_err": sw["nu_rel_err"],
                     "rounding": "stochastic" if form in ("fp16dev", "bf16dev", "e3m12dev") else "nearest"}
    out["fp64_mass_drift"] = ref["mass_drift"]
    out["fp32_mlups"] = round(rates["fp32"], 1)
    out["fp32_steps_per_launch"] = steps_per_pass("fp32")
    out["fp32_pairs_mlups"] = round(fp32_pair, 1)
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


def batched(single_fp32_mlups: float, single_pass_fp32_mlups: float, steps: int = 202) -> dict:
    """512 mixed lattices (sides 64..1024, omega 0.6..1.9, shear wave /
    Couette / body-force channel), autotuned: FP32 (five steps per launch),
    and the 1-3-12 and FP16 deviation forms (three per launch), each with the
    kernel and launch shape that ran, and every lattice's relative L2
    velocity error against an FP64 batch of the same lattices after the same
    steps (median / p90 / max)."""
    from fastfortranboltzmann_amd.batch import batch_study, mixed_batch

    lats = mixed_batch(0, 512)
    forms = ("fp32", "e3m12dev", "fp16dev")
    res = batch_study(lats, steps, autotune=True, trial_steps=30, forms=forms, log=lambda m: None)
    out = {"lattices": len(lats), "cells": res["cells"], "steps": res["steps"]}
    for form in forms:
        rec = res["forms"][form]
        k = batch_steps_per_pass(form)
        tun = rec["tuning"]
        if k > 1:
            launch = {"kernel": "batch_pair_kernel", "pair_shape": tun["pair_shape"],
                      "pair_group": tun["pair_group"], "pair_grid": tun["pair_grid"]}
        else:
            launch = {"kernel": "batch_step_kernel", "tile_w": tun["tile_w"], "grid_blocks": tun["grid_blocks"]}
        err = rec["rel_l2_u_vs_fp64"]
        out[form] = {"mlups": round(rec["mlups"], 1), "steps_per_launch": k,
                     **hbm_rates(rec["mlups"], form, k), "launch": launch,
                     "rel_l2_u_vs_fp64": {k: float(f"{v:.3e}") for k, v in err.items()}}
    out["mlups"] = out["fp32"]["mlups"]
    out["vs_single_4096_fp32"] = round(out["fp32"]["mlups"] / single_fp32_mlups, 3)
    out["vs_single_pass_4096_fp32"] = round(out["fp32"]["mlups"] / single_pass_fp32_mlups, 3)
    mix = large_grid_mix(lats)
    out["large_grids_same_kinds"] = mix
    out["vs_large_grids_same_kinds"] = round(out["fp32"]["mlups"] / mix["mlups"], 3)
    return out


def large_grid_mix(lats, n: int = 4096, steps: int = 240, warmup: int = 24) -> dict:
    """The per-cell rate of large single grids of the batch's own lattice kinds:
    FP32 n^2 shear wave, Couette flow and body-force channel (each with the
    library's default steps per launch, HIP events), combined as the batch
    combines them -- time per cell weighted by each kind's share of the
    batch's cells."""
    from fastfortranboltzmann_amd import Lattice, Settings
    from fastfortranboltzmann_amd.batch import BODY_FORCE_CHANNEL, COUETTE_FLOW, SHEAR_WAVE

    total = sum(L.n_x * L.n_y for L in lats)
    rates, share = {}, {}
    for kind, name, kw in ((SHEAR_WAVE, "shear_wave", dict(U_MAX=0.01, N_SIN=1.0)),
                           (COUETTE_FLOW, "couette", dict(U_WALL=0.05)),
                           (BODY_FORCE_CHANNEL, "body_force_channel", dict(BODY_FORCE=1e-6))):
        share[name] = sum(L.n_x * L.n_y for L in lats if L.kind == kind) / total
        with Lattice(Settings(N_X=n, N_Y=n, SIM_MODE=kind, OMEGA=1.25, precision="fp32", **kw)) as lat:
            lat.init()
            lat.step(warmup)
            lat.sync()
            lat.reset_timer()
            lat.step(steps)
            lat.sync()
            rates[name] = n * n * steps / (lat.elapsed_ms() * 1e-3) / 1e6
    per_cell = sum(share[k] / rates[k] for k in rates)
    return {"size": n, "mlups": round(1.0 / per_cell, 1), "by_kind": {k: round(v, 1) for k, v in rates.items()},
            "cell_share": {k: round(v, 4) for k, v in share.items()}}


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


def copy_ceiling(nbytes: int = 1 << 30, iters: int = 20) -> float | None:
    from fastfortranboltzmann_amd import _lib

    L = _lib.lib()
    gbs = ctypes.c_double()
    if L.ffb_bench_copy(0, nbytes, iters, ctypes.byref(gbs)) != 0:
        return None
    return gbs.value


# ---------------------------------------------------------------------------
# measured HBM traffic (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)
# ---------------------------------------------------------------------------

def _profile_lattice(n: int, precision: str, per_pass: int | None, launches: int, scale: int = 0):
    from fastfortranboltzmann_amd import Lattice, Settings

    s = Settings(N_X=n, N_Y=n, SIM_MODE=4, OMEGA=cavity_omega(re=1000.0 if n == 4096 else 5000.0, n=n),
                 U_LID=0.1, RHO_0=1.0, precision=precision, storage_scale_log2=scale)
    with pass_mode(per_pass):
        k = steps_per_pass(precision)
        lat = Lattice(s)
    with lat:
        lat.init()
        lat.sync()
        copy_ceiling(MARK_START, 1)  # the dispatches up to the end marker are this workload's
        lat.step(launches * k)
        lat.sync()
        copy_ceiling(MARK_END, 1)


def _profile_batch(precision: str, launches: int):
    from fastfortranboltzmann_amd.batch import Batch, mixed_batch

    # the default tuning (an autotuned shape may differ between the two
    # counter passes, whose segments must run the same kernel)
    with Batch(mixed_batch(0, 512), precision) as b:
        b.init()
        b.sync()
        copy_ceiling(MARK_START, 1)
        b.step(launches * batch_steps_per_pass(precision))
        b.sync()
        copy_ceiling(MARK_END, 1)


def _profile_workloads():
    """(workload, run) in the order --profile-run executes them."""
    from fastfortranboltzmann_amd.precision import default_scale_log2

    k = 6
    out = [("sl4096_fp32", lambda: _profile_lattice(4096, "fp32", None, k)),
           ("sl4096_fp32_pair", lambda: _profile_lattice(4096, "fp32", 2, k)),
           ("sl4096_fp32_single_pass", lambda: _profile_lattice(4096, "fp32", 1, k)),
           ("sl4096_fp64", lambda: _profile_lattice(4096, "fp64", None, k)),
           ("sl4096_fp64_pair", lambda: _profile_lattice(4096, "fp64", 2, k)),
           ("sl4096_fp64_single_pass", lambda: _profile_lattice(4096, "fp64", 1, k))]
    for form in ("fp32dev", "fp16dev", "bf16dev", "e3m12dev"):
        s = default_scale_log2(form)
        out.append((f"sl8192_{form}", lambda form=form, s=s: _profile_lattice(8192, form, None, k, s)))
    out.append(("sl8192_fp32", lambda: _profile_lattice(8192, "fp32", None, k)))
    for form in ("fp32", "e3m12dev", "fp16dev"):
        out.append((f"batch512_{form}", lambda form=form: _profile_batch(form, k)))
    return out


def profile_run() -> int:
    """The workloads of profiles/traffic.txt, each between two marker copies,
    then a 1 GiB copy that calibrates the counters."""
    for name, fn in _profile_workloads():
        print(f"profile-run: {name}", file=sys.stderr, flush=True)
        fn()
    copy_ceiling(1 << 30, 1)
    print("profile-run: done", file=sys.stderr, flush=True)
    return 0


def _counter_dispatches(path: str):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return [(r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"])) for r in rows]


def _segments(disp):
    """The dispatches between each start marker (copy kernels over MARK_START
    bytes: one thread per 16 bytes) and the following end marker (MARK_END
    bytes); per segment the kernel with the most dispatches (fills, step
    counters and initial states left out): (kernel, mean counter value, count)."""
    start, end = MARK_START // 16, MARK_END // 16
    segs, cur = [], None
    for name, grid, val in disp:
        if "copy_kernel" in name:
            if grid == start:
                cur = []
            elif grid == end and cur is not None:
                segs.append(cur)
                cur = None
            continue
        if cur is None or "fill" in name.lower() or "count" in name or "init" in name:
            continue
        cur.append((name, val))
    out = []
    for seg in segs:
        by = {}
        for name, val in seg:
            by.setdefault(name, []).append(val)
        name = max(by, key=lambda k: len(by[k]))
        out.append((name, sum(by[name]) / len(by[name]), len(by[name])))
    return out


def record_traffic(fetch_csv: str, write_csv: str) -> int:
    fetch, write = _counter_dispatches(fetch_csv), _counter_dispatches(write_csv)
    # calibration: the 1 GiB nontemporal copy (1048576 KiB read, 1048576 KiB written)
    def calib(disp):
        vals = [v for n, g, v in disp if "copy_kernel<3>" in n and g > (1 << 20)]
        if not vals:
            raise SystemExit("record-traffic: no 1 GiB copy in the counter file")
        return 1048576.0 / (sum(vals) / len(vals))
    fk, wk = calib(fetch), calib(write)
    fs, ws = _segments(fetch), _segments(write)
    names = [n for n, _ in _profile_workloads()]
    if not (len(fs) == len(ws) == len(names)):
        raise SystemExit(f"record-traffic: {len(fs)}/{len(ws)} segments for {len(names)} workloads")
    lines = [f"# kernel sources {kernel_source_hash()}",
             "# HBM bytes per launch of the kernel that runs each workload, from two rocprofv3 --pmc",
             "# passes over `python3 bench.py --profile-run` (FETCH_SIZE, then WRITE_SIZE; KiB per",
             "# dispatch, averaged over the workload's dispatches), scaled by what the same counters",
             f"# read for a 1 GiB copy in the same runs: fetch x{fk:.5f}, write x{wk:.5f}.",
             "# Commands (repo root, GPU host):",
             "#   rocprofv3 --pmc FETCH_SIZE --output-format csv -d <f> -o run -- python3 bench.py --profile-run",
             "#   rocprofv3 --pmc WRITE_SIZE --output-format csv -d <w> -o run -- python3 bench.py --profile-run",
             "#   python3 bench.py --record-traffic <f>/run_counter_collection.csv <w>/run_counter_collection.csv",
             "# workload fetch_bytes write_bytes dispatches kernel"]
    for name, (kf, vf, cf), (kw, vw, cw) in zip(names, fs, ws):
        if kf != kw:
            raise SystemExit(f"record-traffic: {name}: kernels differ ({kf} / {kw})")
        lines.append(f"{name} {vf * 1024 * fk:.0f} {vw * 1024 * wk:.0f} {min(cf, cw)} {kf}")
    with open(TRAFFIC_FILE, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


VALU_FILE = os.path.join(REPO, "profiles", "valu.txt")
VALU_COUNTERS = ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES", "SQ_WAVES", "GRBM_GUI_ACTIVE")
N_SIMD = 256 * 4  # MI355X: 256 CUs of 4 SIMDs
VALU_CYCLES = 4   # cycles a SIMD takes per wave64 VALU instruction (16 lanes per cycle, FP32/FP64 unpacked)


def _multi_counter_dispatches(path: str):
    """(kernel, grid, {counter: value}) per dispatch of a multi-counter rocprofv3 csv."""
    rows = {}
    for r in csv.DictReader(open(path)):
        d = rows.setdefault(int(r["Dispatch_Id"]), (r["Kernel_Name"], int(r["Grid_Size"]), {}))
        d[2][r["Counter_Name"]] = d[2].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return [rows[k] for k in sorted(rows)]


def record_valu(csv_path: str) -> int:
    """profiles/valu.txt: VALU instructions per launch (SQ_INSTS_VALU, wave-level,
    summed over the device) of the kernel that runs each workload of
    --profile-run, the GPU cycles of its dispatches (GRBM_GUI_ACTIVE, counted
    per XCD: / 8) and the share of the SIMDs' cycles the instructions take:
    busy = insts * 4 / (1024 SIMDs * cycles); busy_active, the same from
    SQ_ACTIVE_INST_VALU (cycles in units of 4), as a cross-check."""
    disp = _multi_counter_dispatches(csv_path)
    names = [n for n, _ in _profile_workloads()]
    per = {c: _segments([(k, g, v.get(c, 0.0)) for k, g, v in disp]) for c in VALU_COUNTERS}
    if any(len(v) != len(names) for v in per.values()):
        raise SystemExit(f"record-valu: {len(per['SQ_INSTS_VALU'])} segments for {len(names)} workloads")
    lines = [f"# kernel sources {kernel_source_hash()}",
             "# VALU instructions per launch of the kernel that runs each workload (SQ_INSTS_VALU: wave-level",
             "# instructions summed over the device; mean over the workload's dispatches) from",
             "#   rocprofv3 --pmc " + " ".join(VALU_COUNTERS) + " --output-format csv -d <d> -o run -- \\",
             "#       python3 bench.py --profile-run",
             "#   python3 bench.py --record-valu <d>/run_counter_collection.csv",
             "# cycles = GRBM_GUI_ACTIVE / 8 XCDs (GPU cycles of a dispatch);",
             f"# busy = SQ_INSTS_VALU * {VALU_CYCLES} / ({N_SIMD} SIMDs * cycles);",
             f"# busy_active = SQ_ACTIVE_INST_VALU * 4 / ({N_SIMD} SIMDs * cycles).",
             "# workload valu_insts waves cycles busy busy_active kernel"]
    for i, name in enumerate(names):
        kern, insts, _ = per["SQ_INSTS_VALU"][i]
        waves = per["SQ_WAVES"][i][1]
        act = per["SQ_ACTIVE_INST_VALU"][i][1]
        gui = per["GRBM_GUI_ACTIVE"][i][1] / 8.0
        busy = insts * VALU_CYCLES / (N_SIMD * gui) if gui > 0 else float("nan")
        busy2 = act * 4 / (N_SIMD * gui) if gui > 0 else float("nan")
        lines.append(f"{name} {insts:.0f} {waves:.0f} {gui:.0f} {busy:.3f} {busy2:.3f} {kern}")
    with open(VALU_FILE, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


def valu_from_profiles(workload: str) -> dict | None:
    """VALU record of `workload` from profiles/valu.txt when it was written from
    these kernel sources, else None."""
    if not os.path.exists(VALU_FILE):
        return None
    stamp, out = None, None
    for line in open(VALU_FILE):
        if line.startswith("# kernel sources "):
            stamp = line.split()[-1]
        if line.startswith("#") or not line.strip():
            continue
        parts = line.split(None, 6)
        if parts[0] == workload and len(parts) == 7:
            out = {"wave_insts_per_launch": float(parts[1]), "busy": float(parts[4]),
                   "cycles": float(parts[3]), "kernel": parts[6].strip()}
    return out if stamp == kernel_source_hash() else None


def _traffic_table():
    if not os.path.exists(TRAFFIC_FILE):
        return None, {}
    stamp, table = None, {}
    for line in open(TRAFFIC_FILE):
        if line.startswith("# kernel sources "):
            stamp = line.split()[-1]
        if line.startswith("#") or not line.strip():
            continue
        parts = line.split(None, 4)
        if len(parts) == 5:
            table[parts[0]] = {"bytes": float(parts[1]) + float(parts[2]), "fetch": float(parts[1]),
                               "write": float(parts[2]), "kernel": parts[4].strip()}
    return stamp, table


def traffic_status() -> str:
    stamp, _ = _traffic_table()
    if stamp is None:
        return "profiles/traffic.txt missing or unstamped"
    if stamp != kernel_source_hash():
        return f"profiles/traffic.txt is from kernel sources {stamp}, these are {kernel_source_hash()}: not used"
    return "workload not in profiles/traffic.txt"


def traffic_from_profiles(workload: str) -> dict | None:
    """Measured HBM bytes per launch of `workload` from profiles/traffic.txt,
    or None when the file is stamped with other kernel sources (stale)."""
    stamp, table = _traffic_table()
    if stamp != kernel_source_hash():
        return None
    return table.get(workload)


# ---------------------------------------------------------------------------
# CPU baseline: the reference's own kernels (oracle/_ref, flang)
# ---------------------------------------------------------------------------

REF_FLAGS = "flang -O3 -march=x86-64-v3 -ffp-contract=off"


def _ref_exe(name: str) -> str:
    return os.path.join(REPO, "oracle", "_ref", name, "ref_driver")


def _parse_ref(out: str):
    m = re.search(r"steps=(\d+)\s+cells=(\d+)\s+seconds=\s*([0-9.eE+-]+)", out)
    return (int(m.group(1)), int(m.group(2)), float(m.group(3))) if m else None


def cpu_reference(precision: str) -> dict | None:
    """The reference's kernel modules (simulation.f90, modes/*.f90) compiled
    by flang, timed on the host: one process on the full 4096^2 cavity (one
    core, one coarray image), and CPU_THREADS concurrent processes each on a
    4096 x (4096 / CPU_THREADS) strip -- the reference's 1 x I_Y decomposition
    on all cores without its halo exchange, MLUPS = all cells x steps / the
    slowest process's loop time (main.f90:287-326)."""
    one = _ref_exe(f"bench_sl4096_{precision}")
    strip = _ref_exe(f"bench_sl4096x256_{precision}")
    if not os.path.exists(one):
        return None
    try:
        r1 = _parse_ref(subprocess.run([one, "-"], capture_output=True, text=True, timeout=300).stdout)
        res = None
        if os.path.exists(strip):
            env = dict(os.environ, OMP_NUM_THREADS="1")
            procs = [subprocess.Popen([strip, "-"], stdout=subprocess.PIPE, text=True, env=env)
                     for _ in range(CPU_THREADS)]
            outs = [p.communicate(timeout=300)[0] for p in procs]
            parsed = [_parse_ref(o) for o in outs]
            if all(parsed):
                steps, cells = parsed[0][0], sum(p[1] for p in parsed)
                slowest = max(p[2] for p in parsed)
                res = {"value": round(cells * steps / slowest / 1e6, 1), "cores": CPU_THREADS,
                       "sample": f"{CPU_THREADS} concurrent processes x {steps} steps, each on a "
                                 f"4096x{4096 // CPU_THREADS} strip of the cavity ({precision}); "
                                 f"reference kernels ({REF_FLAGS}), one process = one coarray image, "
                                 "no halo exchange"}
    except (subprocess.TimeoutExpired, OSError):
        return None
    if r1 is None:
        return None
    single = {"value": round(r1[1] * r1[0] / r1[2] / 1e6, 1), "cores": 1,
              "sample": f"{r1[0]} steps of the same 4096^2 cavity ({precision}); reference kernels "
                        f"({REF_FLAGS}), one process = one coarray image"}
    if res is None:
        return {**single, "unit": "MLUPS", "kind": "reference"}
    return {**res, "unit": "MLUPS", "kind": "reference", "single_core": single}


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
    to CPU_THREADS host threads (OpenMP over rows), timed for about `seconds`
    on the same 4096^2 cavity; the baseline where the reference build is absent."""
    import numpy as np

    from oracle.oracle import OracleLattice, Params, max_threads, set_threads

    set_threads(min(CPU_THREADS, os.cpu_count() or 1))
    o = OracleLattice(4, 4096, 4096, Params(omega=cavity_omega(), u_lid=0.1),
                      np.float64 if precision == "fp64" else np.float32).init()
    o.run(1)
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.run(1)
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": round(4096 * 4096 * steps / dt / 1e6, 1), "unit": "MLUPS", "cores": max_threads(),
            "kind": "port", "sample": f"{steps} steps of the same 4096^2 cavity ({precision}), CPU oracle, OpenMP"}


def cpu_baseline(precision: str) -> dict:
    base = cpu_reference(precision) or cpu_oracle_threads(precision)
    base["host"] = host_cpu()
    if precision != "fp64":
        b64 = cpu_reference("fp64")
        if b64:
            base["fp64"] = {k: v for k, v in b64.items() if k not in ("unit", "kind")}
    return base


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
    ap.add_argument("--profile-run", action="store_true", help="run the profiled workloads (rocprofv3 --pmc)")
    ap.add_argument("--record-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.txt from the two counter files of --profile-run")
    ap.add_argument("--record-valu", metavar="SQ_CSV",
                    help="write profiles/valu.txt from the VALU counter file of --profile-run")
    args = ap.parse_args()
    if args.record_traffic:
        return record_traffic(*args.record_traffic)
    if args.record_valu:
        return record_valu(args.record_valu)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the domain is one device-resident "
              "lattice and no multi-GPU path is built (run with --gpus 1)", file=sys.stderr)
        return 2
    if args.profile_run:
        return profile_run()

    n = args.size
    print(f"bench: cavity {n}^2 {args.precision}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    main_run = run_gpu(n, args.precision, args.steps, args.warmup)
    print("bench: one launch per step", file=sys.stderr, flush=True)
    single = run_gpu(n, args.precision, args.steps, args.warmup, per_pass=1)
    workload = f"sl{n}_{args.precision}"
    extra = {"single_pass": {"mlups": round(single["mlups"], 1), "kernel_ms": round(single["kernel_ms"], 5),
                             "roofline": roofline(single, workload + "_single_pass")}}
    if main_run["steps_per_launch"] != 2:
        print("bench: step pairs", file=sys.stderr, flush=True)
        pair = run_gpu(n, args.precision, args.steps, args.warmup, per_pass=2)
        extra["pairs"] = {"mlups": round(pair["mlups"], 1), "kernel_ms": round(pair["kernel_ms"], 5),
                          "roofline": roofline(pair, workload + "_pair")}
    if not args.no_fp64 and args.precision != "fp64":
        print("bench: fp64 run", file=sys.stderr, flush=True)
        r64 = run_gpu(n, "fp64", max(2, args.steps // 2), args.warmup)
        s64 = run_gpu(n, "fp64", max(2, args.steps // 2), args.warmup, per_pass=1)
        p64 = run_gpu(n, "fp64", max(2, args.steps // 2), args.warmup, per_pass=2)
        extra["fp64"] = {"mlups": round(r64["mlups"], 1), "kernel_ms": round(r64["kernel_ms"], 5),
                         "effective_gbs": round(r64["effective_gbs"], 1),
                         "roofline": roofline(r64, f"sl{n}_fp64"),
                         "pairs": {"mlups": round(p64["mlups"], 1), "kernel_ms": round(p64["kernel_ms"], 5),
                                   "roofline": roofline(p64, f"sl{n}_fp64_pair")},
                         "single_pass": {"mlups": round(s64["mlups"], 1), "kernel_ms": round(s64["kernel_ms"], 5),
                                         "roofline": roofline(s64, f"sl{n}_fp64_single_pass")}}
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
    cpu = None
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.precision)
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
        "roofline": roofline(main_run, workload),
        "cpu_baseline": cpu,
        "kernel_ms": round(main_run["kernel_ms"], 5),
        "effective_gbs": round(main_run["effective_gbs"], 1),
        "copy_ceiling_gbs": round(cc, 1) if cc else None,
        "kernel_sources": kernel_source_hash(),
        **extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
