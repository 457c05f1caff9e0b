"""Step pairs and triples in one pass (temporal blocking: step2_kernel,
march2_kernel, stepk_kernel) against one pass per step (step_kernel):
bitwise equal populations and macro fields in every
mode without a strip pass and every storage form, from random states that
make every boundary rule carry a distinct value, on shapes smaller and larger
than one tile (62 x 14 and 62 x 30 for 4-byte and narrower elements, both
region shapes; FP64 62 x 6), with odd step counts (the last step of an odd
count runs alone) and macro output on the last step.  Triples (a pair or a
single step for a remainder) are the default for every form, so the oracle
parity tests cover them too;
FFB_TWO_STEP, FFB_STEPS_PER_PASS, FFB_PAIR_KERNEL (tiles or column march),
FFB_PAIR_SHAPE, FFB_TRIPLE_SHAPE, FFB_MARCH_SHAPE and FFB_MARCH_ROWS select
the path when a lattice is created."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = {1: "shear wave", 2: "Couette", 4: "sliding lid", 5: "body-force channel"}
FORMS = [("fp64", 0), ("fp32", 0), ("fp32dev", 0), ("fp16dev", 10), ("bf16dev", 2), ("e3m12dev", 3)]
SHAPES = [(1, 1), (2, 3), (5, 8), (63, 17), (130, 31), (125, 15), (70, 67), (300, 101)]


def _run(monkeypatch, pairs, mode, nx, ny, precision, scale, f0, steps, env=None):
    from fastfortranboltzmann_amd import Lattice, Settings

    monkeypatch.setenv("FFB_TWO_STEP", "1" if pairs else "0")
    for k, v in (env or {}).items():
        monkeypatch.setenv(k, str(v))
    s = Settings(N_X=nx, N_Y=ny, SIM_MODE=mode, OMEGA=0.83, U_MAX=0.02, N_SIN=1.0, U_WALL=0.05, U_LID=0.07,
                 BODY_FORCE=1e-5, precision=precision, storage_scale_log2=scale,
                 storage_rounding="stochastic")
    with Lattice(s) as lat:
        lat.init()
        lat.set_populations(f0)
        lat.step(steps, write_macro_last=True)
        return lat.populations(), lat.macros(), lat.steps


def _random_state(seed, precision, nx, ny):
    rng = np.random.default_rng(seed)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    dt = np.float64 if precision == "fp64" else np.float32
    return (w[:, None, None] * (1 + 0.05 * rng.standard_normal((9, ny + 2, nx + 2)))).astype(dt)


PAIRS = {"FFB_STEPS_PER_PASS": 2}
KINDS = {"shape8x2": {**PAIRS, "FFB_PAIR_KERNEL": 0, "FFB_PAIR_SHAPE": 2},
         "shape16x2": {**PAIRS, "FFB_PAIR_KERNEL": 0, "FFB_PAIR_SHAPE": 1},
         # XCD blocks of 2 x 3 tiles (partial chunks at the grid's edges)
         "xcd2x3": {**PAIRS, "FFB_PAIR_KERNEL": 0, "FFB_PAIR_SHAPE": 2, "FFB_PAIR_XCD_GROUP": 2, "FFB_PAIR_XCD_ROWS": 3},
         # column march: bands of 8 x 1, 8 x 2, 16 x 1 and 4 x 2 rows, chunks of
         # 5, 16, 40 and 3 rows (several chunks per strip, partial last chunks,
         # ring rows reused many times, chunks shorter than a band)
         "march8x1": {**PAIRS, "FFB_PAIR_KERNEL": 1, "FFB_MARCH_SHAPE": 1, "FFB_MARCH_ROWS": 5},
         "march8x2": {**PAIRS, "FFB_PAIR_KERNEL": 1, "FFB_MARCH_SHAPE": 2, "FFB_MARCH_ROWS": 16},
         "march16x1": {**PAIRS, "FFB_PAIR_KERNEL": 1, "FFB_MARCH_SHAPE": 3, "FFB_MARCH_ROWS": 40},
         "march4x2": {**PAIRS, "FFB_PAIR_KERNEL": 1, "FFB_MARCH_SHAPE": 4, "FFB_MARCH_ROWS": 3},
         # step triples (stepk_kernel): regions 64 x 16 and 64 x 32, outputs 60 x 12 and 60 x 28
         "triple8x2": {"FFB_STEPS_PER_PASS": 3, "FFB_TRIPLE_SHAPE": 2},
         "triple16x2": {"FFB_STEPS_PER_PASS": 3, "FFB_TRIPLE_SHAPE": 1},
         # four steps per launch (FP32, FP64, FP32 and FP16 deviations; the other forms
         # run triples under the same setting): regions 64 x 24, 64 x 32, 64 x 16
         "quad12x2": {"FFB_STEPS_PER_PASS": 4, "FFB_QUAD_SHAPE": 2},
         "quad16x2": {"FFB_STEPS_PER_PASS": 4, "FFB_QUAD_SHAPE": 1},
         "quad8x2": {"FFB_STEPS_PER_PASS": 4, "FFB_QUAD_SHAPE": 3},
         # 8 waves x 4 rows (64 x 32 regions) by stepk_kernel for two, three and
         # four steps (4-byte and narrower elements; FP64 takes its default shape)
         "pair8x4": {"FFB_STEPS_PER_PASS": 2, "FFB_PAIR_KERNEL": 0, "FFB_PAIR_SHAPE": 3},
         "triple8x4": {"FFB_STEPS_PER_PASS": 3, "FFB_TRIPLE_SHAPE": 3},
         "quad8x4": {"FFB_STEPS_PER_PASS": 4, "FFB_QUAD_SHAPE": 4},
         # five steps per launch (FP32 and FP32 deviations; the other forms run
         # four): regions 64 x 32 (8 waves x 4 rows) and 64 x 24 (12 waves x 2 rows)
         "penta8x4": {"FFB_STEPS_PER_PASS": 5, "FFB_PENTA_SHAPE": 1},
         "penta12x2": {"FFB_STEPS_PER_PASS": 5, "FFB_PENTA_SHAPE": 2},
         # persistent grids that load the next tile during the LDS steps
         # (stepk_pipe_kernel): the default grid, and 1 or 3 workgroups, so
         # that each workgroup walks several tiles
         "triple_pipe": {"FFB_STEPS_PER_PASS": 3, "FFB_TRIPLE_SHAPE": 2, "FFB_MULTI_PIPE": 1},
         "triple_pipe_g3": {"FFB_STEPS_PER_PASS": 3, "FFB_TRIPLE_SHAPE": 2, "FFB_MULTI_PIPE": 1, "FFB_MULTI_GRID": 3},
         "quad_pipe": {"FFB_STEPS_PER_PASS": 4, "FFB_QUAD_SHAPE": 2, "FFB_MULTI_PIPE": 1},
         "penta_pipe_g3": {"FFB_STEPS_PER_PASS": 5, "FFB_PENTA_SHAPE": 1, "FFB_MULTI_PIPE": 1, "FFB_MULTI_GRID": 3},
         "quad_pipe_g1": {"FFB_STEPS_PER_PASS": 4, "FFB_QUAD_SHAPE": 2, "FFB_MULTI_PIPE": 1, "FFB_MULTI_GRID": 1},
         "quad_pipe_g3": {"FFB_STEPS_PER_PASS": 4, "FFB_QUAD_SHAPE": 2, "FFB_MULTI_PIPE": 1, "FFB_MULTI_GRID": 3}}


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("shape", SHAPES, ids=[f"{a}x{b}" for a, b in SHAPES])
@pytest.mark.parametrize("form", FORMS, ids=[f[0] for f in FORMS])
@pytest.mark.parametrize("mode", list(MODES), ids=list(MODES.values()))
def test_pairs_equal_single_steps_bitwise(monkeypatch, mode, form, shape, kind):
    precision, scale = form
    nx, ny = shape
    print(f"[gpu] two-step pairs ({kind}) mode {mode} {precision} {nx}x{ny}", flush=True)
    f0 = _random_state(100 * mode + nx + 7 * ny, precision, nx, ny)
    # triples: 3 (one triple), 8 (two triples and a pair), 7 (two triples and a single step);
    # four steps: 4, 11 (two and a triple), 9 (two and a single step)
    # five steps: 5, 14 (two and a four), 12 (two and a pair)
    steps_of = {"triple": (3, 8, 7), "quad": (4, 11, 9), "penta": (5, 14, 12)}
    for steps in steps_of.get(next((k for k in steps_of if kind.startswith(k)), None), (2, 7)):
        fa, ma, na = _run(monkeypatch, True, mode, nx, ny, precision, scale, f0, steps, KINDS[kind])
        fb, mb, nb = _run(monkeypatch, False, mode, nx, ny, precision, scale, f0, steps)
        assert na == nb == steps
        np.testing.assert_array_equal(fa, fb)
        for a, b in zip(ma, mb):
            np.testing.assert_array_equal(a, b)
        assert not np.array_equal(fa, f0.astype(fa.dtype))  # the state moved


def test_pairs_in_chunks_equal_one_call(monkeypatch):
    """A run split into calls of odd length (pairs, then a single step per
    call) lands on the same state as one call."""
    from fastfortranboltzmann_amd import Lattice, Settings

    print("[gpu] two-step pairs chunked 200x90 fp32 cavity", flush=True)
    monkeypatch.setenv("FFB_TWO_STEP", "1")
    s = Settings(N_X=200, N_Y=90, SIM_MODE=4, OMEGA=1.2, U_LID=0.1, precision="fp32")
    with Lattice(s) as a, Lattice(s) as b:
        a.init().step(30, write_macro_last=True)
        b.init()
        for n in (3, 5, 1, 7, 14):
            b.step(n, write_macro_last=True)
        np.testing.assert_array_equal(a.populations(), b.populations())
        for x, y in zip(a.macros(), b.macros()):
            np.testing.assert_array_equal(x, y)
