"""ctypes binding of libffb_hip.so (the C ABI declared in include/ffb.h).

The library is the product: if it is missing this module raises, it never
falls back to any CPU implementation.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libffb_hip.so")

SHEAR_WAVE, COUETTE_FLOW, POISEUILLE_FLOW, SLIDING_LID, BODY_FORCE_CHANNEL = 1, 2, 3, 4, 5
PREC_FP32, PREC_FP64, PREC_FP16_DEV, PREC_BF16_DEV, PREC_FP32_DEV, PREC_E3M12_DEV = 0, 1, 2, 3, 4, 5
ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1
E_INVALID, E_HIP, E_STATE, E_DENSITY = 1, 2, 3, 4

# every entry point include/ffb.h declares
EXPORTS = (
    "ffb_config_default", "ffb_create", "ffb_destroy", "ffb_init", "ffb_step", "ffb_sync",
    "ffb_read_macros", "ffb_read_populations", "ffb_write_populations", "ffb_read_pf_strips",
    "ffb_elapsed_ms", "ffb_reset_timer", "ffb_step_count", "ffb_moments", "ffb_moments_noalloc",
    "ffb_deviation_extent", "ffb_device_properties", "ffb_memory_usage",
    "ffb_default_scale_log2",
    "ffb_device_info",
    "ffb_bench_copy", "ffb_last_error", "ffb_version",
    "ffb_batch_create", "ffb_batch_destroy", "ffb_batch_init", "ffb_batch_step", "ffb_batch_sync",
    "ffb_batch_elapsed_ms", "ffb_batch_reset_timer", "ffb_batch_info", "ffb_batch_set_tuning",
    "ffb_batch_autotune", "ffb_batch_read_macros", "ffb_batch_read_populations",
    "ffb_set_parameters", "ffb_adjoint_create", "ffb_adjoint_destroy", "ffb_adjoint_forward",
    "ffb_adjoint_backward", "ffb_adjoint_timing",
)


class FfbBatchLattice(ctypes.Structure):
    """Mirror of ffb_batch_lattice_t (include/ffb.h)."""

    _fields_ = [
        ("n_x", ctypes.c_int32), ("n_y", ctypes.c_int32), ("kind", ctypes.c_int32),
        ("storage_scale_log2", ctypes.c_int32), ("omega", ctypes.c_double), ("rho_0", ctypes.c_double),
        ("u_wall", ctypes.c_double), ("body_force", ctypes.c_double),
        ("u_max", ctypes.c_double), ("n_sin", ctypes.c_double),
    ]


class FfbBatchTuning(ctypes.Structure):
    """Mirror of ffb_batch_tuning_t (include/ffb.h)."""

    _fields_ = [("tile_w", ctypes.c_int32), ("grid_blocks", ctypes.c_int32), ("pair_shape", ctypes.c_int32),
                ("pair_group", ctypes.c_int32), ("pair_grid", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class FfbConfig(ctypes.Structure):
    """Mirror of ffb_config_t (include/ffb.h)."""

    _fields_ = [
        ("n_x", ctypes.c_int32), ("n_y", ctypes.c_int32),
        ("sim_mode", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("omega", ctypes.c_double), ("rho_0", ctypes.c_double),
        ("u_max", ctypes.c_double), ("n_sin", ctypes.c_double),
        ("u_wall", ctypes.c_double), ("u_lid", ctypes.c_double),
        ("rho_in", ctypes.c_double), ("rho_out", ctypes.c_double),
        ("storage_scale_log2", ctypes.c_int32), ("device", ctypes.c_int32),
        ("body_force", ctypes.c_double),
        ("storage_rounding", ctypes.c_int32), ("density_checks", ctypes.c_int32),
    ]


class FfbError(RuntimeError):
    """A non-zero status of the C ABI, with ffb_last_error()'s message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"ffb error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """Load libffb_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m fastfortranboltzmann_amd.build` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)
    sig = {
        "ffb_config_default": ([ctypes.POINTER(FfbConfig)], None),
        "ffb_create": ([ctypes.POINTER(FfbConfig), ctypes.POINTER(vp)], ctypes.c_int),
        "ffb_destroy": ([vp], ctypes.c_int),
        "ffb_init": ([vp], ctypes.c_int),
        "ffb_step": ([vp, i32, i32], ctypes.c_int),
        "ffb_sync": ([vp], ctypes.c_int),
        "ffb_read_macros": ([vp, vp, vp, vp], ctypes.c_int),
        "ffb_read_populations": ([vp, vp], ctypes.c_int),
        "ffb_write_populations": ([vp, vp], ctypes.c_int),
        "ffb_read_pf_strips": ([vp, vp, vp], ctypes.c_int),
        "ffb_elapsed_ms": ([vp, dp], ctypes.c_int),
        "ffb_reset_timer": ([vp], ctypes.c_int),
        "ffb_step_count": ([vp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "ffb_moments": ([vp, dp, dp, dp], ctypes.c_int),
        "ffb_device_properties": ([i32, ctypes.c_char_p, i32, ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "ffb_memory_usage": ([vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                              ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "ffb_moments_noalloc": ([ctypes.POINTER(FfbConfig), vp, ctypes.c_int64, ctypes.c_int64, vp, vp],
                                ctypes.c_int),
        "ffb_deviation_extent": ([vp, dp, dp], ctypes.c_int),
        "ffb_default_scale_log2": ([i32, ctypes.c_double], ctypes.c_int32),
        "ffb_device_info": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
                             ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
                            ctypes.c_int),
        "ffb_bench_copy": ([i32, ctypes.c_int64, i32, dp], ctypes.c_int),
        "ffb_batch_create": ([ctypes.POINTER(FfbBatchLattice), i32, i32, i32, ctypes.POINTER(vp)],
                             ctypes.c_int),
        "ffb_batch_destroy": ([vp], ctypes.c_int),
        "ffb_batch_init": ([vp], ctypes.c_int),
        "ffb_batch_step": ([vp, i32, i32], ctypes.c_int),
        "ffb_batch_sync": ([vp], ctypes.c_int),
        "ffb_batch_elapsed_ms": ([vp, dp], ctypes.c_int),
        "ffb_batch_reset_timer": ([vp], ctypes.c_int),
        "ffb_batch_info": ([vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(i32),
                            ctypes.POINTER(FfbBatchTuning)], ctypes.c_int),
        "ffb_batch_set_tuning": ([vp, ctypes.POINTER(FfbBatchTuning)], ctypes.c_int),
        "ffb_batch_autotune": ([vp, i32, ctypes.POINTER(FfbBatchTuning), dp], ctypes.c_int),
        "ffb_batch_read_macros": ([vp, i32, vp, vp, vp], ctypes.c_int),
        "ffb_batch_read_populations": ([vp, i32, vp], ctypes.c_int),
        "ffb_set_parameters": ([vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double], ctypes.c_int),
        "ffb_adjoint_create": ([vp, i32, i32, ctypes.POINTER(vp)], ctypes.c_int),
        "ffb_adjoint_destroy": ([vp], ctypes.c_int),
        "ffb_adjoint_forward": ([vp, vp, vp, vp], ctypes.c_int),
        "ffb_adjoint_backward": ([vp, vp, vp, dp, dp, dp], ctypes.c_int),
        "ffb_adjoint_timing": ([vp, dp, dp], ctypes.c_int),
        "ffb_last_error": ([], ctypes.c_char_p),
        "ffb_version": ([], ctypes.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        raise FfbError(rc, lib().ffb_last_error().decode(errors="replace"))
