"""Pins of the oracle's deviation forms (the precision form): the binary16
conversions against numpy's float16 (an independent IEEE implementation,
round to nearest even), the bfloat16 rounding against its definition, the
stochastic rounding against its definition (two neighbours, unbiased over its
random bits) and its effect on a slowly decaying wave, the
deviation-space algebra (halo rules and collision on D = (f - w) 2^s) against
the FP64 restatement of the reference in every mode -- unrounded FP32
deviations must land within FP32 rounding of it, so a dropped term, a wrong
sign or a swapped direction fails -- and the 16-bit runs against FP64."""
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


MODE_PARAMS = {
    1: Params(omega=1.3, u_max=0.05, n_sin=1.0),
    2: Params(omega=0.9, u_wall=0.08),
    3: Params(omega=1.2, rho_in=1.002, rho_out=0.998),
    4: Params(omega=1.7, u_lid=0.1),
    5: Params(omega=1.1, force=1e-4),
}


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5])
def test_fp32_deviation_form_matches_fp64_within_rounding(mode):
    """Unrounded FP32 deviations after 300 steps: max |f - f64| below 4e-7
    (the plain FP32 path reaches 4e-7 .. 1.5e-6 here) and the velocity within
    3e-5 relative L2; measured 2e-8 .. 3.4e-7."""
    p = MODE_PARAMS[mode]
    ref = OracleLattice(mode, 40, 30, p, np.float64).init().run(300)
    dev = OracleLattice(mode, 40, 30, p, storage="fp32dev").init().run(300)
    err = np.abs(dev.interior.astype(np.float64) - ref.interior).max()
    assert err < 4e-7, err
    uerr = np.linalg.norm(dev.u_x.astype(np.float64) - ref.u_x) / np.linalg.norm(ref.u_x)
    assert uerr < 3e-5, uerr
    rerr = np.abs(dev.rho.astype(np.float64) - ref.rho).max()
    assert rerr < 2e-6, rerr


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("storage,scale,tol", [("fp16dev", 16, 5e-5), ("bf16dev", 0, 2e-3)])
def test_16bit_deviation_forms_track_fp64(mode, storage, scale, tol):
    """max |f - f64| after 300 steps within a few storage roundings of the
    deviations (FP16 7e-6 .. 2.4e-5, BF16 2e-4 .. 9.5e-4 measured)."""
    p = MODE_PARAMS[mode]
    ref = OracleLattice(mode, 40, 30, p, np.float64).init().run(300)
    q = OracleLattice(mode, 40, 30, p, storage=storage, scale_log2=scale).init().run(300)
    err = np.abs(q.interior.astype(np.float64) - ref.interior).max()
    assert 0 < err < tol, err


def test_deviation_encode_decode_round_trip():
    """encode/decode of the initial state: |f - decode(encode(f))| within half
    a storage ulp of the deviation (times 2^-s), exact for FP32 at rest."""
    p = Params(omega=1.0, u_max=0.08, n_sin=1.0)
    f32 = OracleLattice(1, 16, 12, p, np.float32).init()
    for storage, scale, rel in (("fp16dev", 12, 2.0 ** -11), ("bf16dev", 0, 2.0 ** -8)):
        q = OracleLattice(1, 16, 12, p, storage=storage, scale_log2=scale).init()
        w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, np.float32)[:, None, None]
        dev = np.abs(f32.interior - w)
        err = np.abs(q.interior - f32.interior)
        assert np.all(err <= dev * rel + 1e-7)
    rest = OracleLattice(4, 8, 8, Params(), storage="fp32dev").init()
    assert np.all(rest.d[:, 1:-1, 1:-1] == 0)


def _neighbours16(x, kind):
    """The two representable FP16 / BF16 values around x (float32 in)."""
    if kind == "fp16":
        h = np.float16(x)
        lo = h if float(h) <= x else np.nextafter(h, np.float16(-np.inf))
        hi = np.nextafter(lo, np.float16(np.inf)) if float(lo) < x else lo
        return float(lo), float(hi)
    b = O.bf16_bits(x)
    v = O.bf16_value(b)
    if v == x:
        return v, v
    if (v < x) == (x > 0):
        lo, hi = sorted((v, O.bf16_value(b + 1)))
    else:
        lo, hi = sorted((v, O.bf16_value(b - 1)))
    return lo, hi


@pytest.mark.parametrize("kind", ["fp16", "bf16"])
def test_stochastic_rounding_is_unbiased_over_its_random_bits(kind):
    """For every random value r = 0..127 the result is one of the two
    neighbours of x, and the mean over all r lies within 1/256 of the
    neighbour spacing of x (7 random bits): an unbiased rounding, unlike
    nearest-even, which always returns the same neighbour."""
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-3000, 3000, 300), rng.uniform(-1, 1, 300) * 1e-3,
                         rng.uniform(-1, 1, 200) * 5e-6, rng.uniform(-1, 1, 100) * 5e-8]).astype(np.float32)
    for x in xs:
        x = float(x)
        lo, hi = _neighbours16(x, kind)
        outs = np.array([O.round_sr(kind, x, r) for r in range(128)])
        assert np.all((outs == lo) | (outs == hi)), (x, lo, hi, set(outs))
        if hi > lo:
            assert abs(outs.mean() - x) <= (hi - lo) / 256 + 1e-12 * abs(x), (x, outs.mean(), lo, hi)


def test_stochastic_rounding_keeps_slow_decay_that_nearest_even_freezes():
    """Shear wave 32x256, omega 1, u_max 0.05: an update of ~1e-4 of the
    amplitude per step is a fifth of an FP16 step.  Nearest-even returns the
    old value every step (the wave never decays); stochastic rounding keeps
    the decay, nu within 1e-3 of (1/omega - 1/2)/3 (2.6e-4 measured; the 32
    columns average the rounding noise of the fit, which for one column
    reaches 1.6e-3)."""
    from fastfortranboltzmann_amd.analysis import decay_amplitude, fit_viscosity

    p = Params(omega=1.0, u_max=0.05, n_sin=1.0)
    res = {}
    for rounding in ("nearest", "stochastic"):
        o = OracleLattice(1, 32, 256, p, storage="fp16dev", scale_log2=17, rounding=rounding).init()
        amps, times = [], []
        for t in range(200, 2001, 200):
            o.run(200)
            amps.append(decay_amplitude(o.u_x.astype(np.float64), 1.0))
            times.append(t)
        res[rounding] = (fit_viscosity(np.array(times), np.array(amps), 256, 1.0), amps[-1] / 0.05)
    nu = 1.0 / 6.0
    assert abs(res["nearest"][0]) < 1e-9 and res["nearest"][1] > 0.999
    assert abs(res["stochastic"][0] - nu) / nu < 1e-3
    assert abs(res["stochastic"][1] - np.exp(-nu * (2 * np.pi / 256) ** 2 * 2000)) < 2e-3
