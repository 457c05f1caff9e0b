"""ctypes front end of the CPU oracle (oracle/libffb_oracle.so).

TEST INFRASTRUCTURE ONLY.  tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py use this module as the checker.  The product
path (fastfortranboltzmann_amd) never imports it.

Arrays use the reference's "xyd" layout with the one-cell halo:
``f[i-1, y, x]`` for x = 0..nx+1, y = 0..ny+1 (numpy shape ``(9, ny+2, nx+2)``,
C order), i.e. ``f(0:NX+1, 0:NY+1, 1:9)`` of
source/layouts/xyd/simulation.f90:27 viewed from C.  Macro fields are
``(ny, nx)`` (x fastest), the Fortran ``rho(NX, NY)``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libffb_oracle.so")

SHEAR_WAVE, COUETTE_FLOW, POISEUILLE_FLOW, SLIDING_LID, BODY_FORCE_CHANNEL = 1, 2, 3, 4, 5

_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "libffb_oracle.so"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        for p in ("d", "s"):
            real = ctypes.c_double if p == "d" else ctypes.c_float
            ptr = ctypes.c_void_p
            i = ctypes.c_int
            f = getattr(_lib, f"ffbo_{p}_init")
            f.argtypes = [i, i, i, i, i, real, real, real, ptr, ptr, ptr, ptr]
            f.restype = None
            f = getattr(_lib, f"ffbo_{p}_run")
            f.argtypes = [i, i, i, i, real, real, real, real, real, real,
                          ptr, ptr, ptr, i, ptr, ptr, ptr, real]
            f.restype = None
            f = getattr(_lib, f"ffbo_{p}_pf_seed")
            f.argtypes = [i, real, real, ptr]
            f.restype = None
        f = _lib.ffbo_q_run
        f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int] * 4 + [ctypes.c_float] * 6 + \
            [ctypes.c_void_p] * 3 + [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_float]
        f.restype = None
        _lib.ffbo_q_encode.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 2
        _lib.ffbo_q_encode.restype = None
        _lib.ffbo_q_decode.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2
        _lib.ffbo_q_decode.restype = None
        _lib.ffbo_float_to_half.argtypes = [ctypes.c_float]
        _lib.ffbo_float_to_half.restype = ctypes.c_ushort
        _lib.ffbo_half_to_float.argtypes = [ctypes.c_ushort]
        _lib.ffbo_half_to_float.restype = ctypes.c_float
        _lib.ffbo_float_to_bf16.argtypes = [ctypes.c_float]
        _lib.ffbo_float_to_bf16.restype = ctypes.c_ushort
        _lib.ffbo_bf16_to_float.argtypes = [ctypes.c_ushort]
        _lib.ffbo_bf16_to_float.restype = ctypes.c_float
        _lib.ffbo_round_sr.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_uint]
        _lib.ffbo_round_sr.restype = ctypes.c_float
        _lib.ffbo_e3m12_to_float.argtypes = [ctypes.c_ushort]
        _lib.ffbo_e3m12_to_float.restype = ctypes.c_float
        _lib.ffbo_float_to_e3m12.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_uint]
        _lib.ffbo_float_to_e3m12.restype = ctypes.c_ushort
        _lib.ffbo_set_threads.argtypes = [ctypes.c_int]
        _lib.ffbo_max_threads.restype = ctypes.c_int
    return _lib


@dataclass
class Params:
    """Physical parameters of one run (names of source/settings.f90:35-53)."""

    omega: float = 1.1
    rho_0: float = 1.0
    u_max: float = 0.1
    n_sin: float = 2.0
    u_wall: float = 0.1
    u_lid: float = 0.1
    rho_in: float = 1.00005
    rho_out: float = 0.99995
    force: float = 0.0  # body-force channel (mode 5): force density along x


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


STORAGE_KINDS = {"fp32dev": 0, "fp16dev": 1, "bf16dev": 2, "e3m12dev": 3}


class OracleLattice:
    """One single-domain run of the reference loop on the CPU.

    ``storage`` = "fp32dev" / "fp16dev" / "bf16dev" / "e3m12dev" restates the
    GPU's deviation forms (no reference counterpart): the state is D = (f - w)
    * 2^scale_log2 stored as FP32 / FP16 / BF16 / the 1-3-12 layout, every
    step works on D in FP32 (deviation-space halo rules and collision,
    ffbo_q_step; the 1-3-12 form on D * 2^-120), and ``f`` holds the
    populations decoded as the GPU returns them, w + D * 2^-scale_log2 in
    FP32.  ``rounding`` ("stochastic", the GPU default, or "nearest") applies
    to the 16-bit forms."""

    def __init__(self, mode: int, nx: int, ny: int, params: Params | None = None,
                 dtype=np.float64, storage: str | None = None, scale_log2: int = 0,
                 rounding: str = "stochastic"):
        self.mode, self.nx, self.ny = int(mode), int(nx), int(ny)
        self.params = params or Params()
        if storage is not None:
            if storage not in STORAGE_KINDS:
                raise ValueError(f"unknown storage {storage!r}")
            dtype = np.float32
        self.storage, self.scale_log2 = storage, int(scale_log2)
        # exponent of the scale the step works in (the 1-3-12 form: D * 2^-120)
        self.work_log2 = self.scale_log2 - (120 if storage == "e3m12dev" else 0)
        if rounding not in ("stochastic", "nearest"):
            raise ValueError(f"unknown rounding {rounding!r}")
        self.stochastic = 1 if rounding == "stochastic" else 0
        self.dtype = np.dtype(dtype)
        self.p = "d" if self.dtype == np.float64 else "s"
        self.f = np.zeros((9, ny + 2, nx + 2), self.dtype)
        self.f_b = np.zeros_like(self.f)
        self.rho = np.zeros((ny, nx), self.dtype)
        self.u_x = np.zeros_like(self.rho)
        self.u_y = np.zeros_like(self.rho)
        self.pf_state = np.zeros((4, 3, ny), self.dtype)
        if storage is not None:
            self.d = np.zeros_like(self.f)
            self.d_b = np.zeros_like(self.f)
        self.steps = 0

    def init(self) -> "OracleLattice":
        P = self.params
        getattr(lib(), f"ffbo_{self.p}_init")(
            self.mode, self.nx, self.ny, 1, self.ny, P.rho_0, P.u_max, P.n_sin,
            _p(self.f), _p(self.rho), _p(self.u_x), _p(self.u_y))
        getattr(lib(), f"ffbo_{self.p}_pf_seed")(self.ny, P.rho_in, P.rho_out, _p(self.pf_state))
        if self.storage is not None:
            lib().ffbo_q_encode(STORAGE_KINDS[self.storage], self.work_log2, self.nx, self.ny,
                                _p(self.f), _p(self.d))
            self._decode()
        self.steps = 0
        return self

    def run(self, n_steps: int, write_macro_last: bool = True) -> "OracleLattice":
        P = self.params
        if self.storage is not None:
            lib().ffbo_q_run(
                STORAGE_KINDS[self.storage], self.work_log2, self.stochastic, self.steps, self.mode,
                self.nx, self.ny, int(n_steps), P.omega, P.rho_0, P.u_wall, P.u_lid, P.rho_in, P.rho_out,
                _p(self.d), _p(self.d_b), _p(self.pf_state), 1 if write_macro_last else 0,
                _p(self.rho), _p(self.u_x), _p(self.u_y), P.force)
            self._decode()
            self.steps += int(n_steps)
            return self
        getattr(lib(), f"ffbo_{self.p}_run")(
            self.mode, self.nx, self.ny, int(n_steps), P.omega, P.rho_0, P.u_wall, P.u_lid,
            P.rho_in, P.rho_out, _p(self.f), _p(self.f_b), _p(self.pf_state),
            1 if write_macro_last else 0, _p(self.rho), _p(self.u_x), _p(self.u_y), P.force)
        self.steps += int(n_steps)
        return self

    def set_populations(self, f_halo: np.ndarray) -> "OracleLattice":
        """Replaces the state by f_halo (shape (9, ny+2, nx+2)); a deviation form
        stores it encoded (D = round((f - w) 2^s)), as ffb_write_populations does."""
        self.f[:] = f_halo.astype(self.dtype)
        if self.storage is not None:
            lib().ffbo_q_encode(STORAGE_KINDS[self.storage], self.work_log2, self.nx, self.ny,
                                _p(self.f), _p(self.d))
            self._decode()
        return self

    def _decode(self) -> None:
        lib().ffbo_q_decode(self.work_log2, self.nx, self.ny, _p(self.d), _p(self.f))

    @property
    def interior(self) -> np.ndarray:
        """Populations of the owned cells, shape (9, ny, nx)."""
        return self.f[:, 1:-1, 1:-1]

    @property
    def recv_macro_left(self) -> np.ndarray:
        return self.pf_state[0]

    @property
    def recv_macro_right(self) -> np.ndarray:
        return self.pf_state[1]


def half_bits(x: float) -> int:
    return int(lib().ffbo_float_to_half(float(x)))


def half_value(bits: int) -> float:
    return float(lib().ffbo_half_to_float(int(bits)))


def bf16_bits(x: float) -> int:
    return int(lib().ffbo_float_to_bf16(float(x)))


def bf16_value(bits: int) -> float:
    return float(lib().ffbo_bf16_to_float(int(bits)))


def round_sr(kind: str, x: float, r: int) -> float:
    """Stochastic rounding of x to "fp16", "bf16" or "e3m12" with random bits r (0..127)."""
    return float(lib().ffbo_round_sr({"fp16": 1, "bf16": 2, "e3m12": 3}[kind], float(x), int(r)))


def e3m12_bits(v: float, r: int | None = None) -> int:
    """Stored bits of the FP32 value v (= D * 2^-120) in the 1-3-12 layout:
    nearest even, or stochastic with random bits r."""
    return int(lib().ffbo_float_to_e3m12(float(v), 0 if r is None else 1, 0 if r is None else int(r)))


def e3m12_value(bits: int) -> float:
    return float(lib().ffbo_e3m12_to_float(int(bits)))


def set_threads(n: int) -> None:
    lib().ffbo_set_threads(int(n))


def max_threads() -> int:
    return int(lib().ffbo_max_threads())
