"""C ABI entry points around the step, on the GPU: population write/read
round trips in every storage form, random states in the deviation forms (every
boundary rule carries a distinct value), FP64 moment sums, and parameter
changes between steps (ffb_set_parameters), each against the oracle."""
import dataclasses

import numpy as np
import pytest

from oracle.oracle import OracleLattice, Params

pytestmark = pytest.mark.gpu

# the 1-3-12 form at s = 2: the random states below reach |f - w| ~ 0.3 (rest
# direction), inside its range 2^(2-s) there
FORMS = [("fp64", 0), ("fp32", 0), ("fp32dev", 0), ("fp16dev", 12), ("bf16dev", 0), ("e3m12dev", 2)]
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def _oracle(mode, nx, ny, p, precision, scale):
    if precision in ("fp64", "fp32"):
        return OracleLattice(mode, nx, ny, p, np.float64 if precision == "fp64" else np.float32)
    return OracleLattice(mode, nx, ny, p, storage=precision, scale_log2=scale)


def _lattice(mode, nx, ny, p, precision, scale):
    from fastfortranboltzmann_amd import Lattice, Settings

    return Lattice(Settings(N_X=nx, N_Y=ny, SIM_MODE=mode, OMEGA=p.omega, U_MAX=p.u_max, N_SIN=p.n_sin,
                            U_WALL=p.u_wall, U_LID=p.u_lid, RHO_IN=p.rho_in, RHO_OUT=p.rho_out,
                            BODY_FORCE=p.force, precision=precision, storage_scale_log2=scale))


def _random_state(nx, ny, seed, dtype):
    rng = np.random.default_rng(seed)
    return (W[:, None, None] * (1 + 0.2 * rng.standard_normal((9, ny + 2, nx + 2)))).astype(dtype)


@pytest.mark.parametrize("precision,scale", FORMS)
def test_population_round_trip(precision, scale):
    nx, ny = 37, 11
    print(f"[gpu] population round trip {precision}", flush=True)
    p = Params(omega=1.1)
    dt = np.float64 if precision == "fp64" else np.float32
    f0 = _random_state(nx, ny, 7, dt)
    o = _oracle(4, nx, ny, p, precision, scale).init().set_populations(f0)
    with _lattice(4, nx, ny, p, precision, scale) as lat:
        lat.init()
        lat.set_populations(f0)
        np.testing.assert_array_equal(lat.interior(), o.interior)
        if precision in ("fp64", "fp32"):
            np.testing.assert_array_equal(lat.interior(), f0[:, 1:-1, 1:-1])


@pytest.mark.parametrize("precision,scale", FORMS[2:])
@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5])
def test_deviation_forms_from_random_state(mode, precision, scale):
    nx, ny = 37, 29
    print(f"[gpu] random state mode {mode} {precision}", flush=True)
    p = Params(omega=0.83, u_wall=0.05, u_lid=0.12, rho_in=1.01, rho_out=0.99,
               force=3e-4 if mode == 5 else 0.0)
    f0 = _random_state(nx, ny, 1234 + mode, np.float32)
    o = _oracle(mode, nx, ny, p, precision, scale).init().set_populations(f0).run(11)
    with _lattice(mode, nx, ny, p, precision, scale) as lat:
        lat.init()
        lat.set_populations(f0)
        lat.step(11, write_macro_last=True)
        np.testing.assert_array_equal(lat.interior(), o.interior)
        rho, ux, uy = lat.macros()
        np.testing.assert_array_equal(ux, o.u_x)
        np.testing.assert_array_equal(uy, o.u_y)
        if mode == 3:
            left, right = lat.pf_strips()
            np.testing.assert_array_equal(left, o.recv_macro_left)
            np.testing.assert_array_equal(right, o.recv_macro_right)


@pytest.mark.parametrize("precision,scale", FORMS)
def test_moments_are_fp64_sums_of_the_state(precision, scale):
    """mass and momentum from ffb_moments against float64 sums of the state
    read back (the read-back rounds a deviation form's w + D 2^-s to FP32,
    hence the tolerance)."""
    nx, ny = 64, 48
    print(f"[gpu] moments {precision}", flush=True)
    p = Params(omega=1.2, u_lid=0.1)
    with _lattice(4, nx, ny, p, precision, scale) as lat:
        lat.init().step(40)
        m, mx, my = lat.moments()
        f = lat.interior().astype(np.float64)
    cx = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float)[:, None, None]
    cy = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], float)[:, None, None]
    tol = 0 if precision in ("fp64", "fp32") else 1e-7
    assert m == pytest.approx(f.sum(), rel=1e-13 + tol)
    assert mx == pytest.approx((cx * f).sum(), abs=1e-10 + tol * nx * ny)
    assert my == pytest.approx((cy * f).sum(), abs=1e-10 + tol * nx * ny)


@pytest.mark.parametrize("precision,scale", FORMS)
@pytest.mark.parametrize("mode", [2, 4, 5])
def test_parameters_changed_between_steps(mode, precision, scale):
    """ffb_set_parameters between two runs of steps: the same as the oracle
    continuing with the new omega, wall speed and force."""
    from fastfortranboltzmann_amd import _lib

    nx, ny = 33, 21
    print(f"[gpu] set_parameters mode {mode} {precision}", flush=True)
    p = Params(omega=1.3, u_wall=0.06, u_lid=0.06, force=2e-4 if mode == 5 else 0.0)
    o = _oracle(mode, nx, ny, dataclasses.replace(p), precision, scale).init().run(7)
    o.params = dataclasses.replace(p, omega=0.9, u_wall=0.09, u_lid=0.09, force=5e-4 if mode == 5 else 0.0)
    o.run(6)
    with _lattice(mode, nx, ny, p, precision, scale) as lat:
        lat.init().step(7)
        _lib.check(_lib.lib().ffb_set_parameters(lat.handle, 0.9, o.params.force, 0.09, 0.09))
        lat.step(6, write_macro_last=True)
        np.testing.assert_array_equal(lat.interior(), o.interior)


@pytest.mark.parametrize("precision,scale", FORMS)
def test_deviation_extent_matches_the_state(precision, scale):
    """ffb_deviation_extent: the largest |f - w| per direction class (rest,
    axis, diagonal) equals the one computed from the oracle's state after the
    same steps (bitwise the GPU's), and the limit is the storage form's."""
    nx, ny, steps = 41, 23, 30
    print(f"[gpu] deviation extent {precision}", flush=True)
    p = Params(omega=1.3, u_lid=0.08)
    o = _oracle(4, nx, ny, p, precision, scale).init()
    dt = np.float64 if precision == "fp64" else np.float32
    f0 = _random_state(nx, ny, 11, dt)
    o.set_populations(f0).run(steps)
    with _lattice(4, nx, ny, p, precision, scale) as lat:
        lat.init()
        lat.set_populations(f0)
        lat.step(steps)
        dev, lim = lat.deviation_extent()
    if o.storage is None:
        w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=dt).astype(np.float64)
        a = np.abs(o.interior.astype(np.float64) - w[:, None, None])
    else:
        a = np.abs(o.d[:, 1:-1, 1:-1].astype(np.float64)) * 2.0 ** -o.work_log2
    want = np.array([a[0].max(), a[1:5].max(), a[5:9].max()])
    np.testing.assert_array_equal(dev, want)
    top = {"fp64": np.finfo(np.float64).max, "fp32": np.finfo(np.float32).max,
           "fp32dev": np.finfo(np.float32).max * 2.0 ** -scale, "fp16dev": 65504.0 * 2.0 ** -scale,
           "bf16dev": 3.3895313892515355e38 * 2.0 ** -scale}
    if precision == "e3m12dev":
        np.testing.assert_array_equal(lim, (2 - 2.0 ** -12) * 2.0 ** (-scale - np.array([-1, 0, 1])))
    else:
        np.testing.assert_array_equal(lim, [top[precision]] * 3)
    assert np.all(dev < lim)


def test_deviation_extent_shows_clamping_of_the_1_3_12_form():
    """Sliding lid u = 0.1 at omega 1.0085: the rest deviation at the lid's
    corners reaches 0.062.  With s = 5 it stays under the rest direction's
    limit 2^(2-s) (no clamping); with s = 7 (limit 2^-5) it is clamped, and
    the extent equals the limit."""
    res = {}
    for s in (5, 7):
        print(f"[gpu] deviation extent clamp s={s}", flush=True)
        p = Params(omega=1.008553, u_lid=0.1)
        with _lattice(4, 64, 64, p, "e3m12dev", s) as lat:
            lat.init().step(400)
            res[s] = lat.deviation_extent()
    dev, lim = res[5]
    assert 0.055 < dev[0] < lim[0] and np.all(dev < lim)
    dev, lim = res[7]
    assert dev[0] == lim[0]


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16dev", "e3m12dev"])
def test_density_check_reports_the_reference_stop(precision):
    """FFB_DENSITY_CHECKS (simulation.f90:229-233): a state with one cell whose
    pulled density is negative makes ffb_sync return FFB_E_DENSITY with the
    reference's message; the flag stays until ffb_init; without the check the
    same run syncs normally; a NaN density is reported as such."""
    import numpy as np

    from fastfortranboltzmann_amd import Lattice, Settings
    from fastfortranboltzmann_amd._lib import E_DENSITY, FfbError

    print(f"[gpu] density check {precision}", flush=True)
    nx, ny = 40, 30
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    dt = np.float64 if precision == "fp64" else np.float32
    f = np.broadcast_to(w[:, None, None], (9, ny + 2, nx + 2)).astype(dt).copy()
    f[:, 11:14, 16:19] = -2.0 * w[:, None, None]  # cell (16, 11) pulls all nine from this block: rho = -2
    for checks in (False, True):
        s = Settings(N_X=nx, N_Y=ny, SIM_MODE=4, OMEGA=1.2, U_LID=0.05, precision=precision,
                     storage_scale_log2=1 if precision.endswith("dev") else 0, DENSITY_CHECKS=checks)
        with Lattice(s) as lat:
            lat.init()
            lat.set_populations(f)
            lat.step(3)
            if not checks:
                lat.sync()
                continue
            with pytest.raises(FfbError) as err:
                lat.sync()
            assert err.value.code == E_DENSITY
            assert "error: density is zero in collision/streaming step (rho_val <= 0)" in str(err.value)
            with pytest.raises(FfbError):
                lat.sync()  # sticky until init
            lat.init().step(4)
            lat.sync()  # a healthy state runs clean
            if precision == "e3m12dev":
                continue  # the 1-3-12 layout has no NaN: a NaN written in is stored as a finite value
            g = np.broadcast_to(w[:, None, None], f.shape).astype(dt).copy()
            g[:, 5, 5] = np.nan
            lat.set_populations(g)
            lat.step(1)
            with pytest.raises(FfbError) as err:
                lat.sync()
            assert err.value.code == E_DENSITY and "NaN" in str(err.value)
