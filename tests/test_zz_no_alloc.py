"""ffb_moments_noalloc (the entry point that allocates nothing) against
ffb_moments on the same populations: bitwise equal FP64 sums, through the C
ABI, on two calls in a row whose second has the smaller domain (the partial
sums of the first stay in the library's scratch array and must not leak into
the second), in every storage form; plus the argument checks, which fail
before anything is launched."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu

FORMS = [("fp32", 0), ("fp64", 0), ("fp32dev", 0), ("fp16dev", 12), ("bf16dev", 0), ("e3m12dev", 5)]


def _sums_noalloc(lat, out):
    from fastfortranboltzmann_amd._lib import check, lib

    cur, stream, pitch, plane = lat.device_info()
    cfg = lat.settings.config()
    check(lib().ffb_moments_noalloc(ctypes.byref(cfg), ctypes.c_void_p(cur), pitch, plane,
                                    ctypes.c_void_p(stream), ctypes.c_void_p(out.data_ptr())))
    lat.sync()
    return tuple(float(v) for v in out.cpu())


@pytest.mark.parametrize("form", FORMS, ids=[f[0] for f in FORMS])
def test_noalloc_moments_equal_ffb_moments_on_two_calls(form):
    import torch

    from fastfortranboltzmann_amd import Lattice, Settings

    precision, scale = form
    print(f"[gpu] no-alloc moments {precision}", flush=True)
    out = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda:0")
    for nx, ny, mode, steps in ((700, 530, 4, 9), (70, 37, 2, 5)):  # 1450 then 11 partial sums
        s = Settings(N_X=nx, N_Y=ny, SIM_MODE=mode, OMEGA=1.3, U_LID=0.1, U_WALL=0.05, precision=precision,
                     storage_scale_log2=scale)
        with Lattice(s) as lat:
            lat.init().step(steps)
            lat.sync()
            ref = lat.moments()
            got = _sums_noalloc(lat, out)
        assert got == ref, (nx, ny, got, ref)
        assert abs(got[0] / (nx * ny) - 1.0) < 1e-3  # mass per cell near rho_0


def test_noalloc_moments_rejects_bad_arguments():
    from fastfortranboltzmann_amd import Settings
    from fastfortranboltzmann_amd._lib import E_INVALID, lib

    print("[gpu] no-alloc moments argument checks", flush=True)
    cfg = Settings(N_X=64, N_Y=64, precision="fp32").config()
    dummy = ctypes.c_void_p(16)  # never dereferenced: every call below fails its checks first
    assert lib().ffb_moments_noalloc(None, dummy, 64, 4096, None, dummy) == E_INVALID
    assert lib().ffb_moments_noalloc(ctypes.byref(cfg), None, 64, 4096, None, dummy) == E_INVALID
    assert lib().ffb_moments_noalloc(ctypes.byref(cfg), dummy, 63, 4096, None, dummy) == E_INVALID
    big = Settings(N_X=1 << 25, N_Y=1, precision="fp32").config()
    assert lib().ffb_moments_noalloc(ctypes.byref(big), dummy, 1 << 25, 1 << 25, None, dummy) == E_INVALID
    assert b"2^24" in lib().ffb_last_error()
