"""CPU-side checks of the C ABI library: it loads, exports every symbol that
include/ffb.h declares, fills the reference defaults, and reports errors
cleanly when no GPU is present (no compute calls here)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(REPO, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(REPO, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ffb_\w+)\s*\(", src, re.M))
    return names


def test_library_exports_every_declared_symbol():
    from fastfortranboltzmann_amd import _lib

    L = _lib.lib()
    declared = _declared_symbols()
    assert declared >= set(_lib.EXPORTS)
    for name in sorted(declared):
        assert hasattr(L, name), name


def test_config_defaults_are_reference_settings():
    from fastfortranboltzmann_amd import _lib

    c = _lib.FfbConfig()
    _lib.lib().ffb_config_default(ctypes.byref(c))
    # source/settings.f90:27-63; FFB_DENSITY_CHECKS is off unless defined
    assert (c.n_x, c.n_y, c.sim_mode) == (2520, 2520, 4)
    assert c.density_checks == 0
    assert (c.omega, c.rho_0, c.u_max, c.n_sin) == (1.1, 1.0, 0.1, 2.0)
    assert (c.u_wall, c.u_lid, c.rho_in, c.rho_out) == (0.1, 0.1, 1.00005, 0.99995)


def test_invalid_config_is_rejected_with_message():
    from fastfortranboltzmann_amd import _lib

    L = _lib.lib()
    c = _lib.FfbConfig()
    L.ffb_config_default(ctypes.byref(c))
    h = ctypes.c_void_p()
    c.sim_mode = 9
    assert L.ffb_create(ctypes.byref(c), ctypes.byref(h)) == 1
    assert b"sim mode" in L.ffb_last_error()
    c.sim_mode = 4
    c.n_x = 0
    assert L.ffb_create(ctypes.byref(c), ctypes.byref(h)) == 1
    assert h.value is None
    # a row of more than 2^24 cells: its 9 direction rows and byte offsets
    # would overflow 32 bits, so it is refused before any device call
    c.n_x, c.n_y = (1 << 24) + 1, 1
    assert L.ffb_create(ctypes.byref(c), ctypes.byref(h)) == 1
    assert b"2^24" in L.ffb_last_error() and h.value is None
    c.n_x = 1 << 24
    assert L.ffb_create(ctypes.byref(c), ctypes.byref(h)) != 0  # no GPU here: fails later, at the device
    assert b"2^24" not in L.ffb_last_error()


def test_invalid_batch_is_rejected_before_any_gpu_call():
    """ffb_batch_create checks every lattice before it looks for a device:
    sizes (n_x <= 2^24, n_x * n_y <= 2^30), kind, precision, scale."""
    from fastfortranboltzmann_amd import _lib

    L = _lib.lib()

    def create(prec=_lib.PREC_FP16_DEV, **kw):
        d = _lib.FfbBatchLattice(n_x=64, n_y=64, kind=_lib.SHEAR_WAVE, storage_scale_log2=8, omega=1.0,
                                 rho_0=1.0, u_wall=0.0, body_force=0.0, u_max=0.01, n_sin=1.0)
        for k, v in kw.items():
            setattr(d, k, v)
        h = ctypes.c_void_p()
        rc = L.ffb_batch_create(ctypes.byref(d), 1, prec, 0, ctypes.byref(h))
        return rc, L.ffb_last_error(), h

    for kw, msg in (({"n_x": (1 << 24) + 1, "n_y": 1}, b"bad size"), ({"n_x": 0}, b"bad size"),
                    ({"n_x": 1 << 16, "n_y": (1 << 14) + 1}, b"bad size"), ({"kind": 3}, b"kind"),
                    ({"storage_scale_log2": 31}, b"storage_scale_log2")):
        rc, err, h = create(**kw)
        assert rc == 1 and msg in err and h.value is None, (kw, err)
    rc, err, _ = create(prec=9)
    assert rc == 1 and b"precision" in err


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU may be present")
def test_create_without_gpu_fails_loudly():
    from fastfortranboltzmann_amd import FfbError, Lattice

    with pytest.raises(FfbError):
        Lattice(N_X=8, N_Y=8)


def test_python_mirror_settings():
    from fastfortranboltzmann_amd import Settings, omega_for_reynolds, reynolds

    s = Settings(N_Y=4096, SIM_MODE=4, U_LID=0.1, OMEGA=omega_for_reynolds(1000, 0.1, 4096))
    assert abs(s.OMEGA - 0.578436) < 1e-6
    assert abs(reynolds(s) - 1000) < 1e-9


def test_default_scale_log2_contract():
    """ffb_default_scale_log2 (host arithmetic, no GPU): for each speed, the
    largest s that keeps the documented bound inside the form's headroom --
    FP16: (4/9)(0.05 + 3u + 6u^2) 2^s <= 2^14; 1-3-12: every class bound
    w (0.05 + 3|c|u + 4.5|c|^2 u^2 + 1.5 u^2) 2^off 2^s <= 3/4 of 2^1 --
    so s + 1 breaks it; 0 for the forms that need no scale; INT32_MIN for an
    unknown precision or a speed outside [0, 1]."""
    import math

    from fastfortranboltzmann_amd import _lib

    f = _lib.lib().ffb_default_scale_log2
    for u in (0.0, 0.01, 0.05, 0.1, 0.137, 0.2, 0.3, 0.5, 1.0):
        b16 = (4 / 9) * (0.05 + 3 * u + 6 * u * u)
        s = f(_lib.PREC_FP16_DEV, u)
        assert b16 * 2.0 ** s <= 2.0 ** 14 < b16 * 2.0 ** (s + 1), (u, s)
        bounds = [(4 / 9) * (0.05 + 1.5 * u * u) * 2.0 ** -1,                       # rest, off -1
                  (1 / 9) * (0.05 + 3 * u + 6 * u * u),                             # axis, off 0
                  (1 / 36) * (0.05 + 3 * math.sqrt(2) * u + 10.5 * u * u) * 2.0]    # diagonal, off +1
        s = f(_lib.PREC_E3M12_DEV, u)
        assert max(bounds) * 2.0 ** s <= 1.5 < max(bounds) * 2.0 ** (s + 1), (u, s)
        for p in (_lib.PREC_FP32, _lib.PREC_FP64, _lib.PREC_BF16_DEV, _lib.PREC_FP32_DEV):
            assert f(p, u) == 0
    assert f(_lib.PREC_E3M12_DEV, 0.1) == 5 and f(_lib.PREC_FP16_DEV, 0.1) == 16
    for p, u in ((_lib.PREC_E3M12_DEV, -0.1), (_lib.PREC_E3M12_DEV, 1.5), (_lib.PREC_FP16_DEV, float("nan")),
                 (9, 0.1), (-1, 0.1)):
        assert f(p, u) == -2 ** 31
