"""Parity of the HIP path (through the C ABI) with the reference.

  * against the golden vectors the reference's own kernels wrote
    (tests/golden/, bitwise, every mode and both precisions);
  * against the CPU oracle (itself bitwise-pinned to those vectors) on the
    same seeded inputs, at sizes the oracle finishes in seconds, bitwise;
  * edge cases: 1- and 2-cell-wide domains, odd non-square grids, random
    states that exercise every boundary rule with non-trivial values.
"""
import json
import os

import numpy as np
import pytest

from oracle.oracle import OracleLattice, Params

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))["cases"]


def _gpu_lattice(mode, nx, ny, params: Params, precision):
    from fastfortranboltzmann_amd import Lattice, Settings

    s = Settings(N_X=nx, N_Y=ny, SIM_MODE=mode, OMEGA=params.omega, RHO_0=params.rho_0,
                 U_MAX=params.u_max, N_SIN=params.n_sin, U_WALL=params.u_wall, U_LID=params.u_lid,
                 RHO_IN=params.rho_in, RHO_OUT=params.rho_out, precision=precision)
    return Lattice(s)


def _params(case):
    p = case["params"]
    return Params(omega=p["omega"], rho_0=p["rho_0"], u_max=p["u_max"], n_sin=p["n_sin"],
                  u_wall=p["u_wall"], u_lid=p["u_lid"], rho_in=p["rho_in"], rho_out=p["rho_out"])


@pytest.mark.parametrize("case", MANIFEST, ids=[c["name"] for c in MANIFEST])
def test_gpu_matches_reference_golden_bitwise(case):
    print(f"[gpu] golden {case['name']}", flush=True)
    g = np.load(os.path.join(GOLDEN, case["file"]))
    with _gpu_lattice(case["sim_mode"], case["n_x"], case["n_y"], _params(case),
                      case["precision"]) as lat:
        lat.init().step(case["n_steps"], write_macro_last=True)
        np.testing.assert_array_equal(lat.interior(), g["f"])
        rho, ux, uy = lat.macros()
        np.testing.assert_array_equal(rho, g["rho"])
        np.testing.assert_array_equal(ux, g["u_x"])
        np.testing.assert_array_equal(uy, g["u_y"])
        if case["sim_mode"] == 3:
            left, right = lat.pf_strips()
            np.testing.assert_array_equal(left, g["recv_macro"][0].reshape(3, -1))
            np.testing.assert_array_equal(right, g["recv_macro"][1].reshape(3, -1))


# includes row pitches that do not divide 4 KiB (300, 1000, 1030), so rows and
# planes start at every alignment within a page
SHAPES = [(1, 1), (1, 5), (5, 1), (2, 2), (2, 7), (7, 2), (3, 3), (8, 3), (9, 4), (64, 3),
          (65, 17), (300, 7), (257, 130), (1000, 9), (1030, 5)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("mode", [1, 2, 3, 4])
@pytest.mark.parametrize("shape", SHAPES, ids=[f"{a}x{b}" for a, b in SHAPES])
@pytest.mark.parametrize("vec", [4, 2, 1], ids=["vec4", "vec2", "scalar"])
def test_gpu_matches_oracle_small_shapes(shape, mode, precision, vec, monkeypatch):
    monkeypatch.setenv("FFB_VEC", str(vec))
    nx, ny = shape
    print(f"[gpu] oracle shapes {nx}x{ny} mode {mode} {precision}", flush=True)
    p = Params(omega=1.37, u_max=0.07, n_sin=1.0, u_wall=0.08, u_lid=0.09, rho_in=1.002,
               rho_out=0.998)
    dt = np.float64 if precision == "fp64" else np.float32
    steps = 23
    o = OracleLattice(mode, nx, ny, p, dt).init().run(steps)
    with _gpu_lattice(mode, nx, ny, p, precision) as lat:
        lat.init().step(steps, write_macro_last=True)
        np.testing.assert_array_equal(lat.interior(), o.interior)
        rho, ux, uy = lat.macros()
        np.testing.assert_array_equal(rho, o.rho)
        np.testing.assert_array_equal(ux, o.u_x)
        np.testing.assert_array_equal(uy, o.u_y)


@pytest.mark.parametrize("precision,storage_scale", [("fp64", 0), ("fp32", 0), ("fp32dev", 0), ("fp16dev", 16),
                                                     ("bf16dev", 0), ("e3m12dev", 5)])
@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("shape", [(1, 1), (2, 7), (65, 17), (300, 7), (1030, 5)], ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("vec", [1, 4], ids=["scalar", "vec4"])
def test_plane_layout_matches_oracle(shape, mode, precision, storage_scale, vec, monkeypatch):
    """Layout 0 (nine planes; the default is the row-blocked layout 1) gives
    the same results: bitwise equal to the oracle."""
    from fastfortranboltzmann_amd import Lattice, Settings

    monkeypatch.setenv("FFB_LAYOUT", "0")
    monkeypatch.setenv("FFB_VEC", str(vec))
    nx, ny = shape
    print(f"[gpu] plane layout {nx}x{ny} mode {mode} {precision}", flush=True)
    p = Params(omega=1.37, u_max=0.07, n_sin=1.0, u_wall=0.08, u_lid=0.09, rho_in=1.002,
               rho_out=0.998, force=2e-4 if mode == 5 else 0.0)
    steps = 13
    if precision in ("fp64", "fp32"):
        o = OracleLattice(mode, nx, ny, p, np.float64 if precision == "fp64" else np.float32)
    else:
        o = OracleLattice(mode, nx, ny, p, storage=precision, scale_log2=storage_scale)
    o.init().run(steps)
    s = Settings(N_X=nx, N_Y=ny, SIM_MODE=mode, OMEGA=p.omega, U_MAX=p.u_max, N_SIN=p.n_sin,
                 U_WALL=p.u_wall, U_LID=p.u_lid, RHO_IN=p.rho_in, RHO_OUT=p.rho_out, BODY_FORCE=p.force,
                 precision=precision, storage_scale_log2=storage_scale)
    with Lattice(s) as lat:
        lat.init().step(steps, write_macro_last=True)
        np.testing.assert_array_equal(lat.interior(), o.interior)
        rho, ux, uy = lat.macros()
        np.testing.assert_array_equal(rho, o.rho)
        np.testing.assert_array_equal(ux, o.u_x)
        np.testing.assert_array_equal(uy, o.u_y)


@pytest.mark.parametrize("precision,storage_scale", [("fp32", 0), ("fp64", 0), ("fp16dev", 16), ("e3m12dev", 5)])
@pytest.mark.parametrize("group", [1, 3, 16])
@pytest.mark.parametrize("shape", [(65, 17), (300, 7), (257, 130), (1030, 5)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_xcd_tile_groups_match_oracle(shape, group, precision, storage_scale, monkeypatch):
    """Any XCD group size G (G consecutive tiles per XCD, the last partial
    chunk in plain order) visits every tile once: results bitwise equal to
    the oracle, with 64-thread tiles so that small grids have many chunks."""
    from fastfortranboltzmann_amd import Lattice, Settings

    monkeypatch.setenv("FFB_XCD_GROUP", str(group))
    monkeypatch.setenv("FFB_TILE_W", "64")
    nx, ny = shape
    print(f"[gpu] xcd group {group} {nx}x{ny} {precision}", flush=True)
    p = Params(omega=1.37, u_lid=0.09)
    if precision in ("fp64", "fp32"):
        o = OracleLattice(4, nx, ny, p, np.float64 if precision == "fp64" else np.float32)
    else:
        o = OracleLattice(4, nx, ny, p, storage=precision, scale_log2=storage_scale)
    o.init().run(9)
    s = Settings(N_X=nx, N_Y=ny, SIM_MODE=4, OMEGA=p.omega, U_LID=p.u_lid, precision=precision,
                 storage_scale_log2=storage_scale)
    with Lattice(s) as lat:
        lat.init().step(9, write_macro_last=True)
        np.testing.assert_array_equal(lat.interior(), o.interior)
        np.testing.assert_array_equal(lat.macros()[1], o.u_x)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_gpu_matches_oracle_from_random_state(mode, precision):
    """A random positive state makes every halo rule carry a distinct value."""
    nx, ny = 37, 29
    print(f"[gpu] random state mode {mode} {precision}", flush=True)
    rng = np.random.default_rng(1234 + mode)
    dt = np.float64 if precision == "fp64" else np.float32
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    f0 = (w[:, None, None] * (1 + 0.2 * rng.standard_normal((9, ny + 2, nx + 2)))).astype(dt)
    p = Params(omega=0.83, u_wall=0.05, u_lid=0.12, rho_in=1.01, rho_out=0.99)
    o = OracleLattice(mode, nx, ny, p, dt).init()
    o.f[:] = f0
    o.run(11)
    with _gpu_lattice(mode, nx, ny, p, precision) as lat:
        lat.init()
        lat.set_populations(f0)
        lat.step(11, write_macro_last=True)
        np.testing.assert_array_equal(lat.interior(), o.interior)


def test_gpu_shear_wave_256_fp64_2000_steps_and_viscosity():
    """Shear-wave decay, 256x256 FP64, omega 1, 2000 steps: bitwise equal to the
    oracle, and the amplitude decays as exp(-nu k^2 t) with nu = (1/omega - 1/2)/3."""
    from fastfortranboltzmann_amd.analysis import decay_amplitude, fit_viscosity

    print("[gpu] shear wave 256^2 fp64 2000 steps", flush=True)
    p = Params(omega=1.0, u_max=0.05, n_sin=1.0)
    n, steps, chunk = 256, 2000, 200
    o = OracleLattice(1, n, n, p, np.float64).init().run(steps)
    amps, times = [], []
    with _gpu_lattice(1, n, n, p, "fp64") as lat:
        lat.init()
        for t in range(0, steps + 1, chunk):
            if t:
                lat.step(chunk, write_macro_last=True)
                _, ux, _ = lat.macros()
                amps.append(decay_amplitude(ux, 1.0))
                times.append(t)
        np.testing.assert_array_equal(lat.interior(), o.interior)
    nu_fit = fit_viscosity(np.array(times), np.array(amps), n, 1.0)
    nu = (1.0 / 1.0 - 0.5) / 3.0
    assert abs(nu_fit - nu) / nu < 2e-3, (nu_fit, nu)
    assert abs(amps[-1] / 0.05 - np.exp(-nu * (2 * np.pi / n) ** 2 * steps)) < 2e-3


def test_gpu_mass_conservation_and_moments():
    """Bounce-back (resting and moving walls) and periodic wraps conserve mass."""
    print("[gpu] moments", flush=True)
    for mode in (1, 2, 4):
        with _gpu_lattice(mode, 96, 80, Params(omega=1.6), "fp64") as lat:
            lat.init()
            m0, _, _ = lat.moments()
            lat.step(500)
            m1, _, _ = lat.moments()
            assert abs(m1 - m0) / m0 < 1e-12, (mode, m0, m1)


@pytest.mark.slow
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gpu_cavity_4096_step_1000_matches_oracle(precision):
    """The cavity workload at full size: 4096^2, Re 1000 (omega from
    calculate_reynolds, settings.f90:166-172), u_lid 0.1, 1000 steps, through
    the default multi-step launches (FP64 four steps, FP32 five).  Bar (FP64):
    max|df| <= 1e-12 against the reference restated by the oracle; the
    kernels are bitwise equal, so the test asserts 0 in both precisions."""
    import os

    from oracle.oracle import set_threads
    from fastfortranboltzmann_amd import omega_for_reynolds

    n, steps = 4096, 1000
    dt = np.float64 if precision == "fp64" else np.float32
    p = Params(omega=omega_for_reynolds(1000.0, 0.1, n), u_lid=0.1)
    print(f"[gpu] cavity {n}^2 {precision} {steps} steps vs oracle", flush=True)
    with _gpu_lattice(4, n, n, p, precision) as lat:
        lat.init().step(steps, write_macro_last=True)
        g = lat.interior()
        _, gux, _ = lat.macros()
    set_threads(min(16, os.cpu_count() or 1))
    o = OracleLattice(4, n, n, p, dt).init().run(steps)
    dmax = float(np.abs(g - o.interior).max())
    print(f"[gpu] cavity {n}^2 max|df| = {dmax:.3e}, max|u_x| = {float(np.abs(gux).max()):.4f}", flush=True)
    assert dmax <= 1e-12
    assert dmax == 0.0
    np.testing.assert_array_equal(gux, o.u_x)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("steps,decades", [(1, 3.0), (5, 0.3)], ids=["1step-6decades", "5steps-0.6decades"])
@pytest.mark.parametrize("mode", [1, 4, 5])
def test_quotients_over_wide_density_range(mode, steps, decades, precision):
    """FP32 and FP64 velocity quotients u = j / rho over a wide range of densities and
    momenta: a random state whose cells carry scales 10^U(-decades, decades),
    so that a cell's pulled populations differ by up to 2 * decades decades
    (rho from about 1e-3 to 1e3 and |u| up to about 1 for one step).  The
    kernels divide with a shared refined reciprocal (the compiler's division
    without its range scaling); the oracle divides with IEEE division.  One
    step (every pulled value positive, so rho > 0) in a single-step launch,
    and five steps from a milder state (FP32: one five-step launch; FP64: a
    triple and a pair); bitwise equal in the shear wave, the sliding lid and
    the Guo-forced channel."""
    from fastfortranboltzmann_amd import Lattice, Settings

    nx, ny = 150, 61
    print(f"[gpu] wide-range quotients mode {mode} steps {steps} decades {decades} {precision}", flush=True)
    dt = np.float64 if precision == "fp64" else np.float32
    rng = np.random.default_rng(4321 + mode)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    scale = 10.0 ** rng.uniform(-decades, decades, (1, ny + 2, nx + 2))
    f0 = (w[:, None, None] * scale * (1 + 0.3 * rng.uniform(-1, 1, (9, ny + 2, nx + 2)))).astype(dt)
    p = Params(omega=0.83, u_lid=0.12, force=3e-4)
    o = OracleLattice(mode, nx, ny, p, dt).init()
    o.f[:] = f0
    o.run(steps)
    s = Settings(N_X=nx, N_Y=ny, SIM_MODE=mode, OMEGA=p.omega, U_LID=p.u_lid, BODY_FORCE=p.force,
                 precision=precision)
    with Lattice(s) as lat:
        lat.init()
        lat.set_populations(f0)
        lat.step(steps, write_macro_last=True)
        assert np.isfinite(o.interior).all()
        np.testing.assert_array_equal(lat.interior(), o.interior)
        np.testing.assert_array_equal(lat.macros()[1], o.u_x)
        np.testing.assert_array_equal(lat.macros()[2], o.u_y)
