"""Batched form: hundreds of independent lattices of uneven size, each with
its own omega and boundary kind, advanced two steps per launch of one
persistent kernel (hipGraph-replayed step pairs) in libffb_hip.so, in any
storage form (the deviation forms with a scale per lattice); see ffb_batch_*
in include/ffb.h.  No reference counterpart.

The mixed workload (``mixed_batch``): side lengths uniform on [64, 1024]
(x and y drawn separately, so non-square), omega uniform on [0.6, 1.9], kinds
shear wave / Couette (top wall u = 0.05) / body-force channel (force density
1e-6 along x, periodic x) in equal shares, shuffled, from numpy's PCG64
generator seeded with ``seed``.  Shear waves start from u_x = 0.01 sin(2 pi y / n_y).
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass

import numpy as np

from ._lib import (BODY_FORCE_CHANNEL, COUETTE_FLOW, SHEAR_WAVE, SLIDING_LID, FfbBatchLattice, FfbBatchTuning,
                   check, lib)
from .lattice import PRECISIONS


@dataclass
class BatchLattice:
    n_x: int
    n_y: int
    kind: int
    omega: float
    rho_0: float = 1.0
    u_wall: float = 0.0
    body_force: float = 0.0
    u_max: float = 0.0
    n_sin: float = 1.0
    storage_scale_log2: int | None = None  # deviation forms; None: the default at speed()

    def speed(self) -> float:
        """Largest expected speed: the wave amplitude, the wall speed, or for a
        body-force channel its steady centre-line speed F n_y^2 / (8 nu),
        capped at 0.1."""
        if self.kind == SHEAR_WAVE:
            return abs(self.u_max)
        if self.kind in (COUETTE_FLOW, SLIDING_LID):
            return abs(self.u_wall)
        nu = (1.0 / self.omega - 0.5) / 3.0
        return min(0.1, abs(self.body_force) * self.n_y ** 2 / (8.0 * nu))

    def scale_for(self, precision: str) -> int:
        if self.storage_scale_log2 is not None:
            return int(self.storage_scale_log2)
        if not precision.endswith("dev"):
            return 0
        from .precision import default_scale_log2
        return default_scale_log2(precision, self.speed())

    def c(self, precision: str = "fp32") -> FfbBatchLattice:
        return FfbBatchLattice(n_x=self.n_x, n_y=self.n_y, kind=self.kind,
                               storage_scale_log2=self.scale_for(precision),
                               omega=self.omega, rho_0=self.rho_0, u_wall=self.u_wall,
                               body_force=self.body_force, u_max=self.u_max, n_sin=self.n_sin)


def mixed_batch(seed: int = 0, count: int = 512, side=(64, 1024), omega=(0.6, 1.9),
                u_wall: float = 0.05, body_force: float = 1e-6, sw_u_max: float = 0.01):
    rng = np.random.Generator(np.random.PCG64(seed))
    kinds = np.array([SHEAR_WAVE, COUETTE_FLOW, BODY_FORCE_CHANNEL] * (count // 3 + 1))[:count]
    rng.shuffle(kinds)
    out = []
    for k in kinds:
        nx, ny = (int(v) for v in rng.integers(side[0], side[1] + 1, size=2))
        om = float(rng.uniform(omega[0], omega[1]))
        if k == SHEAR_WAVE:
            out.append(BatchLattice(nx, ny, int(k), om, u_max=sw_u_max, n_sin=1.0))
        elif k == COUETTE_FLOW:
            out.append(BatchLattice(nx, ny, int(k), om, u_wall=u_wall))
        else:
            out.append(BatchLattice(nx, ny, int(k), om, body_force=body_force))
    return out


class Batch:
    def __init__(self, lattices, precision: str = "fp32", device: int = 0):
        if precision not in PRECISIONS:
            raise ValueError(f"unknown precision {precision!r}")
        self.lattices = list(lattices)
        self.precision = precision
        self.scales = [l.scale_for(precision) for l in self.lattices]
        self.dtype = np.dtype(np.float64 if precision == "fp64" else np.float32)
        arr = (FfbBatchLattice * len(self.lattices))(*[l.c(precision) for l in self.lattices])
        h = ctypes.c_void_p()
        check(lib().ffb_batch_create(arr, len(self.lattices), PRECISIONS[precision], device, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ffb_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init(self):
        check(lib().ffb_batch_init(self._h))
        return self

    def step(self, n: int, write_macro_last: bool = False):
        check(lib().ffb_batch_step(self._h, int(n), 1 if write_macro_last else 0))
        return self

    def sync(self):
        check(lib().ffb_batch_sync(self._h))

    def elapsed_ms(self) -> float:
        ms = ctypes.c_double()
        check(lib().ffb_batch_elapsed_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def reset_timer(self):
        check(lib().ffb_batch_reset_timer(self._h))

    @staticmethod
    def _tuning_dict(t) -> dict:
        return {k: getattr(t, k) for k in ("tile_w", "grid_blocks", "pair_shape", "pair_group", "pair_grid")}

    def info(self):
        cells, tiles, t = ctypes.c_int64(), ctypes.c_int32(), FfbBatchTuning()
        check(lib().ffb_batch_info(self._h, ctypes.byref(cells), ctypes.byref(tiles), ctypes.byref(t)))
        return {"cells": cells.value, "tiles": tiles.value, "tuning": self._tuning_dict(t)}

    def set_tuning(self, tile_w: int | None = None, grid_blocks: int | None = None, **pair):
        """Changes the launch shape; arguments left out keep their value.
        pair: pair_shape, pair_group, pair_grid (step pairs)."""
        cur = self.info()["tuning"]
        if tile_w is not None:
            cur["tile_w"] = tile_w
        if grid_blocks is not None:
            cur["grid_blocks"] = grid_blocks
        for k, v in pair.items():
            if k not in ("pair_shape", "pair_group", "pair_grid"):
                raise TypeError(f"unknown tuning field {k!r}")
            cur[k] = v
        t = FfbBatchTuning(**cur)
        check(lib().ffb_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, trial_steps: int = 20):
        t, ms = FfbBatchTuning(), ctypes.c_double()
        check(lib().ffb_batch_autotune(self._h, int(trial_steps), ctypes.byref(t), ctypes.byref(ms)))
        return {**self._tuning_dict(t), "ms_per_step": ms.value}

    def macros(self, i: int):
        L = self.lattices[i]
        rho = np.empty((L.n_y, L.n_x), self.dtype)
        ux, uy = np.empty_like(rho), np.empty_like(rho)
        check(lib().ffb_batch_read_macros(self._h, i, ctypes.c_void_p(rho.ctypes.data),
                                          ctypes.c_void_p(ux.ctypes.data), ctypes.c_void_p(uy.ctypes.data)))
        return rho, ux, uy

    def populations(self, i: int) -> np.ndarray:
        L = self.lattices[i]
        f = np.empty((9, L.n_y + 2, L.n_x + 2), self.dtype)
        check(lib().ffb_batch_read_populations(self._h, i, ctypes.c_void_p(f.ctypes.data)))
        return f


def batch_study(lattices, steps: int, autotune: bool = True, trial_steps: int = 20,
                compare_fp64: bool = True, log=print, forms=("fp32",)) -> dict:
    """Aggregate MLUPS of the batch in each storage form of `forms` (any of
    "fp32", "fp32dev", "fp16dev", "bf16dev", "e3m12dev"; the deviation forms at
    each lattice's default scale; events around `steps` steps) and the
    per-instance relative L2 error of u against an FP64 batch of the same
    lattices.  The first form's figures are also the top-level keys."""
    res = {"forms": {}}
    fields = {}
    for form in forms:
        rec = {}
        with Batch(lattices, form) as b:
            if autotune:
                rec["tuning"] = b.autotune(trial_steps)
                log(f"batch {form}: autotuned {rec['tuning']}")
            info = b.info()
            res.update(cells=info["cells"], tiles=info["tiles"], count=len(lattices))
            b.init()
            b.step(2)
            b.sync()
            b.reset_timer()
            t0 = time.perf_counter()
            b.step(steps, write_macro_last=True)
            b.sync()
            wall = time.perf_counter() - t0
            ms = b.elapsed_ms()
            rec["mlups"] = info["cells"] * steps / (ms * 1e-3) / 1e6
            rec["mlups_wall"] = info["cells"] * steps / wall / 1e6
            res["steps"] = steps + 2
            if compare_fp64:
                fields[form] = [b.macros(i)[1:] for i in range(len(lattices))]
        res["forms"][form] = rec
    if compare_fp64:
        errs = {form: [] for form in forms}
        with Batch(lattices, "fp64") as b64:
            b64.init()
            b64.step(steps + 2, write_macro_last=True)
            for i in range(len(lattices)):
                _, ux, uy = b64.macros(i)
                den = math.sqrt(float(np.sum(ux ** 2 + uy ** 2)))
                for form in forms:
                    ux32, uy32 = fields[form][i]
                    num = math.sqrt(float(np.sum((ux32 - ux) ** 2 + (uy32 - uy) ** 2)))
                    errs[form].append(num / den if den > 0 else 0.0)
        for form in forms:
            e = np.array(errs[form])
            res["forms"][form]["rel_l2_u_vs_fp64"] = {"max": float(e.max()), "median": float(np.median(e)),
                                                      "p90": float(np.quantile(e, 0.9))}
            res["forms"][form]["per_instance_rel_l2_u"] = [float(x) for x in e]
    res.update({k: v for k, v in res["forms"][forms[0]].items()})
    return res


def main(argv=None) -> int:
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description="batched form: aggregate MLUPS and error vs FP64")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--trial-steps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--forms", default="fp32,fp32dev,e3m12dev,fp16dev", help="storage forms, comma-separated")
    args = ap.parse_args(argv)
    lats = mixed_batch(args.seed, args.count)
    res = batch_study(lats, args.steps, trial_steps=args.trial_steps, compare_fp64=not args.no_fp64,
                      log=lambda m: print(m, file=sys.stderr, flush=True), forms=tuple(args.forms.split(",")))
    if args.count > 64:
        res.pop("per_instance_rel_l2_u", None)
        for rec in res["forms"].values():
            rec.pop("per_instance_rel_l2_u", None)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
