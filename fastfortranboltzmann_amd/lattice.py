"""Host-side mirror of the reference's single-domain solver interface.

``Lattice`` owns one device-resident D2Q9 domain in libffb_hip.so and exposes
what the reference's driver does with its arrays:

  reference (source/layouts/xyd/...)              here
  ----------------------------------------------  --------------------------------
  allocate(f_a, f_b, rho, u_x, u_y) main.f90:96    Lattice(...)
  apply_condition_*_local initialization.f90       Lattice.init()
  loop body main.f90:194-262 (exchange_halos,      Lattice.step(n, write_macro_last)
    PF pre-pass, execute_local_sim_step)
  rho, u_x, u_y (NX, NY)                           Lattice.macros() -> (ny, nx) arrays
  f(0:NX+1, 0:NY+1, 9)                             Lattice.populations() -> (9, ny+2, nx+2)
  recv_macro_left/right (NY, 3)                    Lattice.pf_strips()

Arrays handed back are numpy, in the reference's Fortran element order viewed
from C (x fastest).  The element type is float64 for FP64 storage and float32
otherwise.  Parameter names follow source/settings.f90.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (BODY_FORCE_CHANNEL, COUETTE_FLOW, POISEUILLE_FLOW, PREC_BF16_DEV, PREC_E3M12_DEV,
                   PREC_FP16_DEV, PREC_FP32, PREC_FP32_DEV, PREC_FP64, SHEAR_WAVE, SLIDING_LID, FfbConfig, check,
                   lib)

PRECISIONS = {"fp32": PREC_FP32, "fp64": PREC_FP64, "fp16dev": PREC_FP16_DEV,
              "bf16dev": PREC_BF16_DEV, "fp32dev": PREC_FP32_DEV, "e3m12dev": PREC_E3M12_DEV}
ROUNDINGS = {"nearest": _lib.ROUND_NEAREST, "stochastic": _lib.ROUND_STOCHASTIC}
MODES = {"shear_wave": SHEAR_WAVE, "couette_flow": COUETTE_FLOW,
         "poiseuille_flow": POISEUILLE_FLOW, "sliding_lid": SLIDING_LID,
         "body_force_channel": BODY_FORCE_CHANNEL}


@dataclass
class Settings:
    """Run-time form of the reference's compile-time settings
    (source/settings.f90:21-93; defaults are the reference's)."""

    N_X: int = 2520
    N_Y: int = 2520
    N_STEPS: int = 1000
    SIM_MODE: int = SLIDING_LID
    OMEGA: float = 1.1
    RHO_0: float = 1.0
    U_MAX: float = 0.1
    N_SIN: float = 2.0
    U_WALL: float = 0.1
    U_LID: float = 0.1
    RHO_IN: float = 1.00005
    RHO_OUT: float = 0.99995
    precision: str = "fp32"          # FFB_FP64 off -> "fp32", on -> "fp64"
    storage_scale_log2: int = 0      # s of the deviation forms
    device: int = 0
    BODY_FORCE: float = 0.0          # body-force channel (mode 5) only
    storage_rounding: str = "stochastic"  # 16-bit forms: "stochastic" or "nearest"
    DENSITY_CHECKS: bool = False     # FFB_DENSITY_CHECKS: sync reports rho <= 0 (settings.f90, simulation.f90:229)
    extra: dict = field(default_factory=dict)

    def config(self) -> FfbConfig:
        if self.precision not in PRECISIONS:
            raise ValueError(f"unknown precision {self.precision!r}")
        if self.storage_rounding not in ROUNDINGS:
            raise ValueError(f"unknown storage rounding {self.storage_rounding!r}")
        return FfbConfig(
            n_x=int(self.N_X), n_y=int(self.N_Y), sim_mode=int(self.SIM_MODE),
            precision=PRECISIONS[self.precision], omega=float(self.OMEGA),
            rho_0=float(self.RHO_0), u_max=float(self.U_MAX), n_sin=float(self.N_SIN),
            u_wall=float(self.U_WALL), u_lid=float(self.U_LID), rho_in=float(self.RHO_IN),
            rho_out=float(self.RHO_OUT), storage_scale_log2=int(self.storage_scale_log2),
            device=int(self.device), body_force=float(self.BODY_FORCE),
            storage_rounding=ROUNDINGS[self.storage_rounding],
            density_checks=1 if self.DENSITY_CHECKS else 0)


def reynolds(settings: Settings) -> float:
    """calculate_reynolds (source/settings.f90:154-176)."""
    nu = (1.0 / settings.OMEGA - 0.5) / 3.0
    if settings.SIM_MODE == COUETTE_FLOW:
        return settings.U_WALL * settings.N_Y / nu
    if settings.SIM_MODE == SLIDING_LID:
        return settings.U_LID * settings.N_Y / nu
    return 0.0


def omega_for_reynolds(re: float, u: float, length: int) -> float:
    """Inverse of calculate_reynolds: omega with nu = u*length/re."""
    nu = u * length / re
    return 1.0 / (3.0 * nu + 0.5)


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


class Lattice:
    """One device-resident domain (see module docstring)."""

    def __init__(self, settings: Settings | None = None, **kw):
        if settings is None:
            settings = Settings(**kw)
        elif kw:
            raise TypeError("pass either a Settings or keyword settings, not both")
        self.settings = settings
        self.nx, self.ny = int(settings.N_X), int(settings.N_Y)
        self.dtype = np.dtype(np.float64 if settings.precision == "fp64" else np.float32)
        cfg = settings.config()
        h = ctypes.c_void_p()
        check(lib().ffb_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ffb_destroy(self._h)
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

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError("lattice is closed")
        return self._h

    # -- the reference's operations ----------------------------------------
    def init(self) -> "Lattice":
        check(lib().ffb_init(self.handle))
        return self

    def step(self, n_steps: int = 1, write_macro_last: bool = False) -> "Lattice":
        check(lib().ffb_step(self.handle, int(n_steps), 1 if write_macro_last else 0))
        return self

    def sync(self) -> None:
        check(lib().ffb_sync(self.handle))

    def macros(self):
        rho = np.empty((self.ny, self.nx), self.dtype)
        ux, uy = np.empty_like(rho), np.empty_like(rho)
        check(lib().ffb_read_macros(self.handle, _ptr(rho), _ptr(ux), _ptr(uy)))
        return rho, ux, uy

    def populations(self) -> np.ndarray:
        f = np.empty((9, self.ny + 2, self.nx + 2), self.dtype)
        check(lib().ffb_read_populations(self.handle, _ptr(f)))
        return f

    def interior(self) -> np.ndarray:
        return self.populations()[:, 1:-1, 1:-1]

    def set_populations(self, f_halo: np.ndarray) -> None:
        f = np.ascontiguousarray(f_halo, dtype=self.dtype)
        if f.shape != (9, self.ny + 2, self.nx + 2):
            raise ValueError(f"expected shape {(9, self.ny + 2, self.nx + 2)}, got {f.shape}")
        check(lib().ffb_write_populations(self.handle, _ptr(f)))

    def pf_strips(self):
        left = np.empty((3, self.ny), self.dtype)
        right = np.empty_like(left)
        check(lib().ffb_read_pf_strips(self.handle, _ptr(left), _ptr(right)))
        return left, right

    def moments(self):
        m, mx, my = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().ffb_moments(self.handle, ctypes.byref(m), ctypes.byref(mx), ctypes.byref(my)))
        return m.value, mx.value, my.value

    def deviation_extent(self):
        """(max_dev, limit): the largest |f - w| of the current state per
        direction class (rest, axis, diagonal) and the largest the storage form
        holds (ffb_deviation_extent); max_dev == limit shows clamping in the
        1-3-12 form."""
        dev = (ctypes.c_double * 3)()
        lim = (ctypes.c_double * 3)()
        check(lib().ffb_deviation_extent(self.handle, dev, lim))
        return np.array(dev[:]), np.array(lim[:])

    def elapsed_ms(self) -> float:
        ms = ctypes.c_double()
        check(lib().ffb_elapsed_ms(self.handle, ctypes.byref(ms)))
        return ms.value

    def reset_timer(self) -> None:
        check(lib().ffb_reset_timer(self.handle))

    @property
    def steps(self) -> int:
        n = ctypes.c_int64()
        check(lib().ffb_step_count(self.handle, ctypes.byref(n)))
        return n.value

    def device_info(self):
        cur, stream = ctypes.c_void_p(), ctypes.c_void_p()
        pitch, plane = ctypes.c_int64(), ctypes.c_int64()
        check(lib().ffb_device_info(self.handle, ctypes.byref(cur), ctypes.byref(stream),
                                    ctypes.byref(pitch), ctypes.byref(plane)))
        return cur.value, stream.value, pitch.value, plane.value


def version() -> str:
    return lib().ffb_version().decode()


__all__ = ["Lattice", "Settings", "reynolds", "omega_for_reynolds", "version", "MODES",
           "PRECISIONS", "_lib"]
