"""The scaled-deviation storage forms on the GPU.

  * bitwise against the oracle's restatement of the same form (state
    D = (f - w) * 2^s rounded to FP16/BF16/the 1-3-12 layout or unrounded
    FP32, deviation-space halo rules and collision in FP32, ffbo_q_step; the
    conversions are pinned to numpy's float16 and to the 1-3-12 layout's
    definition, and the deviation algebra to the FP64 oracle, in
    tests/test_oracle_storage16.py and tests/test_oracle_e3m12.py), every
    mode, odd shapes;
  * error of the velocity field against an FP64 run of the same cavity,
    with the bound written in the test.
"""
import numpy as np
import pytest

from oracle.oracle import OracleLattice, Params

pytestmark = pytest.mark.gpu

SHAPES = [(13, 9), (65, 17), (300, 7), (1030, 5)]


@pytest.mark.parametrize("storage,scale,rounding", [
    ("fp16dev", 16, "stochastic"), ("fp16dev", 11, "stochastic"), ("fp16dev", 16, "nearest"),
    ("bf16dev", 0, "stochastic"), ("bf16dev", 0, "nearest"), ("fp32dev", 0, "nearest"), ("fp32dev", 3, "nearest"),
    ("e3m12dev", 5, "stochastic"), ("e3m12dev", 3, "stochastic"), ("e3m12dev", 5, "nearest")])
@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("shape", SHAPES, ids=[f"{a}x{b}" for a, b in SHAPES])
@pytest.mark.parametrize("vec", [1, 2, 8])
def test_16bit_storage_matches_oracle_emulation(shape, mode, storage, scale, rounding, vec, monkeypatch):
    from fastfortranboltzmann_amd import Lattice, Settings

    monkeypatch.setenv("FFB_VEC", str(vec))
    nx, ny = shape
    print(f"[gpu] storage {storage} s={scale} {rounding} {nx}x{ny} mode {mode}", flush=True)
    p = Params(omega=1.37, u_max=0.07, n_sin=1.0, u_wall=0.08, u_lid=0.09, rho_in=1.002, rho_out=0.998,
               force=2e-4 if mode == 5 else 0.0)
    steps = 17
    o = OracleLattice(mode, nx, ny, p, storage=storage, scale_log2=scale, rounding=rounding).init()
    o.run(9).run(steps - 9)  # two calls: the rounding step index continues
    s = Settings(N_X=nx, N_Y=ny, SIM_MODE=mode, OMEGA=p.omega, U_MAX=p.u_max, N_SIN=p.n_sin,
                 U_WALL=p.u_wall, U_LID=p.u_lid, RHO_IN=p.rho_in, RHO_OUT=p.rho_out, BODY_FORCE=p.force,
                 precision=storage, storage_scale_log2=scale, storage_rounding=rounding)
    with Lattice(s) as lat:
        lat.init().step(steps, write_macro_last=True)
        np.testing.assert_array_equal(lat.interior(), o.interior)
        rho, ux, uy = lat.macros()
        np.testing.assert_array_equal(rho, o.rho)
        np.testing.assert_array_equal(ux, o.u_x)


def test_16bit_storage_error_against_fp64_cavity():
    """Cavity 512^2, Re 1000, 6000 steps: error of each storage form against
    FP64 (relative L2 of the velocity field, density error) and FP64-summed
    mass drift, plus the shear-wave viscosity fitted in each form.  Measured
    on MI355X: FP32 9.6e-5, FP16 deviations (s = 16, stochastic) 4.1e-4,
    1-3-12 deviations (s = 5, stochastic) 1.3e-4, BF16 3.1e-3.  The bounds
    are about twice the measured values (regression guards, not targets)."""
    from fastfortranboltzmann_amd.precision import precision_study

    print("[gpu] precision study 512^2", flush=True)
    res = precision_study(512, 1000.0, 6000, forms=("fp16dev", "bf16dev", "fp32", "fp32dev", "e3m12dev"),
                          log=lambda m: None)
    f = res["forms"]
    print({k: {kk: v[kk] for kk in ("rel_l2_u", "max_abs_rho", "mass_drift") if kk in v} for k, v in f.items()},
          flush=True)
    print({k: v["shear_wave"]["nu_rel_err"] for k, v in f.items() if "shear_wave" in v}, flush=True)
    assert f["fp32"]["rel_l2_u"] < 2e-4
    assert f["fp32dev"]["rel_l2_u"] < f["fp32"]["rel_l2_u"] / 3
    assert f["fp16dev"]["rel_l2_u"] < 1e-3
    assert f["e3m12dev"]["rel_l2_u"] < 3e-4
    assert f["e3m12dev"]["rel_l2_u"] < f["fp16dev"]["rel_l2_u"] / 2
    assert f["bf16dev"]["rel_l2_u"] < 2e-2
    assert abs(f["fp64"]["mass_drift"]) < 1e-11
    assert abs(f["fp32"]["mass_drift"]) < 2e-4
    assert abs(f["fp32dev"]["mass_drift"]) < 1e-8
    assert abs(f["fp16dev"]["mass_drift"]) < 1e-5
    assert abs(f["e3m12dev"]["mass_drift"]) < 1e-5
    for form, bound in (("fp32", 1e-4), ("fp32dev", 1e-4), ("fp16dev", 2e-3), ("bf16dev", 5e-3),
                        ("e3m12dev", 2e-3)):
        assert f[form]["shear_wave"]["nu_rel_err"] < bound, (form, f[form]["shear_wave"])


def test_error_target_met_on_1024_cavity():
    """The precision target, relative L2 velocity error <= 1e-4 against FP64,
    on the cavity 1024^2, Re 1000, 2000 steps (the bench's error case): met by
    the 1-3-12 deviations (16 bits, stochastic rounding; measured 8.16e-5) and
    the FP32 deviations (measured 1.81e-6).  FP16 deviations (2.85e-4) do not
    meet it and are not asserted here."""
    from fastfortranboltzmann_amd.precision import precision_study

    print("[gpu] precision target 1024^2", flush=True)
    res = precision_study(1024, 1000.0, 2000, forms=("e3m12dev", "fp32dev"), log=lambda m: None)
    f = res["forms"]
    print({k: v.get("rel_l2_u") for k, v in f.items()}, flush=True)
    assert f["e3m12dev"]["rel_l2_u"] <= 1e-4
    assert f["fp32dev"]["rel_l2_u"] <= 1e-4
