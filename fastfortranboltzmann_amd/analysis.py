"""Shear-wave decay analysis: the mode amplitude of the reference's decay plot
(visual/plot_decay.py:127-150) plus a least-squares viscosity fit the
reference does not have.

The field is (N_Y, N_X) with x fastest, the layout of the exported
``velocity_x_*.bin`` files (export.f90:500-515, visual/plot_A.py:78-81).
"""
from __future__ import annotations

import numpy as np


def wave_number(n_y: int, n_sin: float) -> float:
    """k = 2 pi n_sin / N_Y (plot_decay.py:127-128)."""
    return 2.0 * np.pi * n_sin / n_y


def lattice_viscosity(omega: float) -> float:
    """nu = (1/omega - 1/2) / 3 (settings.f90:166, plot_decay.py:131-132)."""
    return (1.0 / omega - 0.5) / 3.0


def decay_amplitude(u_x: np.ndarray, n_sin: float) -> float:
    """Projection of the x-mean of u_x(y) on sin(k y), y = 0..N_Y-1
    (plot_decay.py:135-144)."""
    u_x = np.asarray(u_x, dtype=np.float64)
    n_y = u_x.shape[0]
    profile = u_x.mean(axis=1)
    mode = np.sin(wave_number(n_y, n_sin) * np.arange(n_y, dtype=np.float64))
    norm = float(mode @ mode)
    if norm == 0.0:
        raise ValueError("shear wave mode has zero normalization")
    return float(profile @ mode / norm)


def analytical_amplitudes(steps, u_max: float, omega: float, n_y: int, n_sin: float):
    """A(t) = u_max exp(-nu k^2 t) (plot_decay.py:147-150)."""
    k = wave_number(n_y, n_sin)
    return u_max * np.exp(-lattice_viscosity(omega) * k * k * np.asarray(steps, dtype=np.float64))


def fit_viscosity(steps, amplitudes, n_y: int, n_sin: float) -> float:
    """nu from a least-squares line through ln A(t) = ln A0 - nu k^2 t."""
    t = np.asarray(steps, dtype=np.float64)
    a = np.asarray(amplitudes, dtype=np.float64)
    if t.size < 2:
        raise ValueError("need at least two samples")
    slope = np.polyfit(t, np.log(a), 1)[0]
    k = wave_number(n_y, n_sin)
    return float(-slope / (k * k))
