"""Batched form on the GPU: every lattice of a mixed batch, advanced by the
persistent hipGraph-replayed launches, is bitwise equal to the oracle run of
that lattice alone (FP32, FP64 and the deviation forms at each lattice's own
scale, the 16-bit ones with stochastic rounding, whose step index the graph
replays advance on the device), across graph replays, odd/even step counts,
re-tuning (the graph is captured again) and the autotuner's
re-initialisation; with step triples (the default for the 4- and 8-byte forms)
and step pairs in one launch under several tunings, persistent grids smaller
than the tile count included, and with one launch per step under several
single-step tunings.  The 512-lattice mixed
batch is checked against an FP64 batch of the same lattices, lattice by
lattice."""
import numpy as np
import pytest

from oracle.oracle import OracleLattice, Params

pytestmark = pytest.mark.gpu

SW, CF, SL, BF = 1, 2, 4, 5


def small_batch():
    from fastfortranboltzmann_amd.batch import BatchLattice

    return [
        BatchLattice(13, 9, SW, 1.3, u_max=0.05, n_sin=1.0),
        BatchLattice(65, 17, CF, 0.7, u_wall=0.05),
        BatchLattice(130, 70, BF, 1.6, body_force=1e-5),
        BatchLattice(1, 5, SW, 1.1, u_max=0.02, n_sin=1.0),
        BatchLattice(300, 7, SL, 1.9, u_wall=0.1),
        BatchLattice(64, 64, BF, 0.6, body_force=3e-6),
        BatchLattice(97, 33, CF, 1.45, u_wall=0.08),
        BatchLattice(2, 2, SL, 1.0, u_wall=0.03),
    ]


def oracle_of(L, dtype, steps, scale=0):
    p = Params(omega=L.omega, rho_0=L.rho_0, u_max=L.u_max, n_sin=L.n_sin, u_wall=L.u_wall,
               u_lid=L.u_wall, force=L.body_force)
    if isinstance(dtype, str):  # a deviation form at the lattice's scale
        return OracleLattice(L.kind, L.n_x, L.n_y, p, storage=dtype, scale_log2=scale).init().run(steps)
    return OracleLattice(L.kind, L.n_x, L.n_y, p, dtype).init().run(steps)


# (mode, tuning): "5" / "4" five / four steps per launch (4-byte forms; three
# for the others), "3" step triples, "2" step pairs, with tile shape / XCD group /
# persistent grid (forced for every form), or "0" one launch per step
# with workgroup size / persistent grid
MULTI = [{"pair_shape": 2, "pair_group": 32, "pair_grid": 0},
         {"pair_shape": 1, "pair_group": 3, "pair_grid": 8},
         {"pair_shape": 2, "pair_group": 0, "pair_grid": 40}]
TUNINGS = [("5", t) for t in MULTI] + [("4", t) for t in MULTI] + [("3", t) for t in MULTI] + [("2", t) for t in MULTI] + [
    ("0", {"tile_w": 256, "grid_blocks": 0}),
    ("0", {"tile_w": 128, "grid_blocks": 8}),
    ("0", {"tile_w": 512, "grid_blocks": 3}),
    ("0", {"tile_w": 1024, "grid_blocks": 0}),
]


@pytest.mark.parametrize("mode,tuning", TUNINGS,
                         ids=[f"k{p}-" + "-".join(f"{k}{v}" for k, v in t.items()) for p, t in TUNINGS])
@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp32dev", "fp16dev", "bf16dev", "e3m12dev"])
def test_batch_matches_oracle_per_lattice(monkeypatch, precision, mode, tuning):
    """mode "5"/"4"/"3"/"2": five to two steps per launch (the tuning applies to the
    multi-step kernel; a remainder takes single steps); "0": one launch per
    step everywhere."""
    from fastfortranboltzmann_amd.batch import Batch

    print(f"[gpu] batch {precision} steps per launch {mode} tuning={tuning}", flush=True)
    if mode == "0":
        monkeypatch.setenv("FFB_BATCH_PAIRS", "0")
        monkeypatch.delenv("FFB_BATCH_STEPS_PER_PASS", raising=False)
    else:
        monkeypatch.setenv("FFB_BATCH_PAIRS", "1")
        monkeypatch.setenv("FFB_BATCH_STEPS_PER_PASS", mode)
    lats = small_batch()
    dt = {"fp64": np.float64, "fp32": np.float32}.get(precision, precision)
    with Batch(lats, precision) as b:
        b.set_tuning(**tuning)
        b.init()
        done = 0
        for chunk in (5, 6, 1, 11, 8):       # parities 1/0 at the chunk starts; graph cycles inside
            b.step(chunk, write_macro_last=True)
            done += chunk
            for i, L in enumerate(lats):
                o = oracle_of(L, dt, done, b.scales[i])
                np.testing.assert_array_equal(b.populations(i)[:, 1:-1, 1:-1], o.interior,
                                              err_msg=f"lattice {i} after {done} steps")
                rho, ux, uy = b.macros(i)
                np.testing.assert_array_equal(ux, o.u_x)
                np.testing.assert_array_equal(rho, o.rho)
            if done == 12:  # re-tune between graph replays: the graph is captured again
                b.set_tuning(pair_group=5, grid_blocks=16)


@pytest.mark.parametrize("precision", ["fp32", "e3m12dev"])
def test_batch_autotune_reinitialises_and_keeps_results(precision):
    """The autotuner's trials advance the state (and, in the 16-bit forms, the
    step index of the stochastic rounding); its re-initialisation restarts
    both, so the run afterwards equals the oracle's from step 0."""
    from fastfortranboltzmann_amd.batch import Batch

    print(f"[gpu] batch autotune {precision}", flush=True)
    lats = small_batch()
    with Batch(lats, precision) as b:
        best = b.autotune(4)
        assert best["tile_w"] in (64, 128, 256, 512, 1024) and best["ms_per_step"] > 0
        assert best["pair_shape"] in (1, 2) and best["pair_group"] >= 0 and best["pair_grid"] >= 0
        assert b.info()["tuning"] == {k: v for k, v in best.items() if k != "ms_per_step"}
        b.step(9, write_macro_last=True)
        dt = np.float32 if precision == "fp32" else precision
        for i, L in enumerate(lats):
            np.testing.assert_array_equal(b.populations(i)[:, 1:-1, 1:-1],
                                          oracle_of(L, dt, 9, b.scales[i]).interior)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_single_lattice_body_force_channel_matches_oracle(precision):
    from fastfortranboltzmann_amd import Lattice, Settings

    print(f"[gpu] body-force channel {precision}", flush=True)
    dt = np.float64 if precision == "fp64" else np.float32
    p = Params(omega=1.3, force=2e-5)
    o = OracleLattice(5, 77, 31, p, dt).init().run(40)
    with Lattice(Settings(N_X=77, N_Y=31, SIM_MODE=5, OMEGA=1.3, BODY_FORCE=2e-5,
                          precision=precision)) as lat:
        lat.init().step(40, write_macro_last=True)
        np.testing.assert_array_equal(lat.interior(), o.interior)
        np.testing.assert_array_equal(lat.macros()[1], o.u_x)


@pytest.mark.parametrize("precision,bound", [("fp32dev", 1e-4), ("fp32", 1.5e-1)])
def test_mixed_batch_512_against_fp64(precision, bound):
    """The benchmark's 512-lattice mixed batch (seed 0), autotuned, 300 steps:
    every lattice's relative L2 error of u against an FP64 batch of the same
    lattices is below `bound` (FP32 deviations: the precision form's 1e-4
    target; plain FP32: the body-force channels, whose force of 1e-6 moves
    f ~ w by a few FP32 ulps per step, carry errors of several per cent)."""
    from fastfortranboltzmann_amd.batch import batch_study, mixed_batch

    print(f"[gpu] mixed batch 512 {precision} vs fp64", flush=True)
    res = batch_study(mixed_batch(0, 512), 298, autotune=True, trial_steps=4, forms=(precision,),
                      log=lambda m: print(m, flush=True))
    err = res["forms"][precision]["rel_l2_u_vs_fp64"]
    print(f"[gpu] mixed batch 512 {precision}: {err}", flush=True)
    assert err["max"] < bound, err
