"""The CPU oracle against the golden vectors written by the reference's own
kernels (tests/golden/make_golden.py: source/layouts/xyd/simulation.f90,
modes/*.f90, initialization.f90 compiled by flang).  Bitwise equality."""
import json
import os

import numpy as np
import pytest

from oracle.oracle import OracleLattice, Params

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))["cases"]


def params_of(case):
    p = case["params"]
    return Params(omega=p["omega"], rho_0=p["rho_0"], u_max=p["u_max"], n_sin=p["n_sin"],
                  u_wall=p["u_wall"], u_lid=p["u_lid"], rho_in=p["rho_in"], rho_out=p["rho_out"])


@pytest.mark.parametrize("case", MANIFEST, ids=[c["name"] for c in MANIFEST])
def test_oracle_matches_reference_bitwise(case):
    g = np.load(os.path.join(GOLDEN, case["file"]))
    dt = np.float64 if case["precision"] == "fp64" else np.float32
    o = OracleLattice(case["sim_mode"], case["n_x"], case["n_y"], params_of(case), dt)
    o.init().run(case["n_steps"], write_macro_last=True)
    assert g["f"].dtype == dt
    np.testing.assert_array_equal(o.interior, g["f"])
    np.testing.assert_array_equal(o.rho, g["rho"])
    np.testing.assert_array_equal(o.u_x, g["u_x"])
    np.testing.assert_array_equal(o.u_y, g["u_y"])
    if case["sim_mode"] == 3:
        np.testing.assert_array_equal(o.recv_macro_left, g["recv_macro"][0])
        np.testing.assert_array_equal(o.recv_macro_right, g["recv_macro"][1])


def test_manifest_covers_every_mode_and_precision():
    seen = {(c["sim_mode"], c["precision"]) for c in MANIFEST}
    assert seen >= {(m, p) for m in (1, 2, 3, 4) for p in ("fp64", "fp32")}
