"""Pins of the oracle's 16-bit storage emulation (the precision form): the
binary16 conversions against numpy's float16 (an independent IEEE
implementation, round to nearest even), the bfloat16 rounding against its
definition, and the FP16 run against the FP32 run it perturbs."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle.oracle import OracleLattice, Params


def test_half_to_float_is_exact_for_every_code():
    codes = np.arange(65536, dtype=np.uint16)
    ref = codes.view(np.float16).astype(np.float32)
    got = np.array([O.half_value(int(c)) for c in codes], dtype=np.float32)
    finite = np.isfinite(ref)
    np.testing.assert_array_equal(got[finite], ref[finite])
    assert np.all(np.isnan(got[~finite & np.isnan(ref)]))
    assert np.all(np.isinf(got[np.isinf(ref)]))


def test_float_to_half_rounds_like_numpy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(30000),
        (rng.standard_normal(20000) * 1e-5).astype(np.float32),   # subnormal range
        (rng.standard_normal(20000) * 1e-7).astype(np.float32),
        np.array([0.0, -0.0, 65504, 65519.99, 65520, 65536, 1e6, -1e6, 2.0 ** -24, 2.0 ** -25,
                  2.0 ** -25 * 1.0000001, 3 * 2.0 ** -25, 6.1035156e-05], dtype=np.float32),
    ])
    # exact halfway cases between neighbouring half values
    h = np.arange(0, 0x7c00, 97, dtype=np.uint16).view(np.float16).astype(np.float32)
    mids = ((h[:-1].astype(np.float64) + np.nextafter(h[:-1].astype(np.float16), np.float16(np.inf)).astype(np.float64)) / 2).astype(np.float32)
    xs = np.concatenate([xs, mids, -mids])
    with np.errstate(over="ignore"):
        ref = xs.astype(np.float16).view(np.uint16)
    got = np.array([O.half_bits(float(x)) for x in xs], dtype=np.uint16)
    np.testing.assert_array_equal(got, ref)


def test_float_to_bf16_is_nearest_even():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.standard_normal(20000).astype(np.float32),
                         (rng.standard_normal(2000) * 1e30).astype(np.float32)])
    for x in xs[:4000]:
        b = O.bf16_bits(float(x))
        v = O.bf16_value(b)
        lo = O.bf16_value(b - 1) if b & 0x7fff else v
        hi = O.bf16_value(b + 1)
        d, dl, dh = abs(v - x), abs(lo - x), abs(hi - x)
        assert d <= dl and d <= dh
        if d == dl or d == dh:        # a tie: the chosen code is even
            assert b % 2 == 0


@pytest.mark.parametrize("storage,scale,tol", [("fp16dev", 16, 2e-3), ("bf16dev", 0, 3e-2)])
def test_16bit_run_tracks_fp32(storage, scale, tol):
    p = Params(omega=1.3, u_lid=0.1)
    a = OracleLattice(4, 24, 20, p, np.float32).init().run(200)
    b = OracleLattice(4, 24, 20, p, storage=storage, scale_log2=scale).init().run(200)
    err = np.linalg.norm(b.u_x - a.u_x) / np.linalg.norm(a.u_x)
    assert 0 < err < tol
