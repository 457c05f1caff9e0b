"""Differentiation form: the hand-written adjoint against central finite
differences of the FP64 forward solver (h = 1e-4 relative), checkpoint-
interval invariance, FP32 against FP64 gradients, the torch.autograd.Function,
and the reverse-sweep cost."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # mode, params (true / perturbed), parameters to differentiate
    (5, dict(OMEGA=1.3, BODY_FORCE=4e-5), dict(OMEGA=1.25, BODY_FORCE=5e-5), ("omega", "body_force")),
    (2, dict(OMEGA=0.9, U_WALL=0.05), dict(OMEGA=0.95, U_WALL=0.06), ("omega", "u_wall")),
    (4, dict(OMEGA=1.6, U_LID=0.08, U_WALL=0.08), dict(OMEGA=1.5, U_LID=0.07, U_WALL=0.07), ("omega", "u_wall")),
    (1, dict(OMEGA=1.1, U_MAX=0.04, N_SIN=1.0), dict(OMEGA=1.2, U_MAX=0.04, N_SIN=1.0), ("omega",)),
]


def _settings(mode, kw, precision, nx=48, ny=32):
    from fastfortranboltzmann_amd import Settings

    return Settings(N_X=nx, N_Y=ny, SIM_MODE=mode, precision=precision, **kw)


def _target(mode, kw, steps, precision="fp64"):
    from fastfortranboltzmann_amd.adjoint import Adjoint

    with Adjoint(_settings(mode, kw, precision), steps, 50) as a:
        _, tx, ty = a.forward()
    return tx.astype(np.float64), ty.astype(np.float64)


@pytest.mark.parametrize("case", CASES, ids=["body_force", "couette", "sliding_lid", "shear_wave"])
def test_adjoint_matches_central_differences_fp64(case):
    from fastfortranboltzmann_amd.adjoint import Adjoint, finite_difference, loss_and_gradient

    mode, true_kw, start_kw, which = case
    steps = 300
    print(f"[gpu] adjoint FD mode {mode}", flush=True)
    tx, ty = _target(mode, true_kw, steps)
    with Adjoint(_settings(mode, start_kw, "fp64"), steps, 64) as a:
        loss, g = loss_and_gradient(a, tx, ty)
        assert loss > 0
        grads = dict(zip(("omega", "body_force", "u_wall"), g))
        for name in which:
            fd = finite_difference(a, tx, ty, name, 1e-4)
            rel = abs(grads[name] - fd) / abs(fd)
            print(f"  {name}: adjoint {grads[name]:.10e} FD {fd:.10e} rel {rel:.2e}", flush=True)
            assert rel < 1e-6, (name, grads[name], fd)


def test_adjoint_is_independent_of_checkpoint_interval():
    from fastfortranboltzmann_amd.adjoint import Adjoint, loss_and_gradient

    print("[gpu] adjoint checkpoint intervals", flush=True)
    mode, true_kw, start_kw, _ = CASES[0]
    steps = 120
    tx, ty = _target(mode, true_kw, steps)
    ref = None
    for k in (1, 7, 50, 120, 500):
        with Adjoint(_settings(mode, start_kw, "fp64"), steps, k) as a:
            _, g = loss_and_gradient(a, tx, ty)
        if ref is None:
            ref = g
        assert g == ref, (k, g, ref)


def test_adjoint_fp32_tracks_fp64():
    from fastfortranboltzmann_amd.adjoint import Adjoint, loss_and_gradient

    print("[gpu] adjoint fp32 vs fp64", flush=True)
    mode, true_kw, start_kw, _ = CASES[0]
    steps = 300
    tx, ty = _target(mode, true_kw, steps)
    out = {}
    for prec in ("fp64", "fp32", "fp32dev"):
        with Adjoint(_settings(mode, start_kw, prec), steps, 100) as a:
            out[prec] = loss_and_gradient(a, tx, ty)[1]
    for prec in ("fp32", "fp32dev"):
        rel = [abs(out[prec][k] - out["fp64"][k]) / abs(out["fp64"][k]) for k in (0, 1)]
        print(f"  {prec}: relative gradient error vs fp64 {rel}", flush=True)
        assert max(rel) < 1e-3


def test_torch_function_gradients_equal_adjoint():
    torch = pytest.importorskip("torch")
    from fastfortranboltzmann_amd.adjoint import Adjoint, LatticeVelocity, loss_and_gradient

    print("[gpu] torch autograd function", flush=True)
    mode, true_kw, start_kw, _ = CASES[0]
    steps = 100
    tx, ty = _target(mode, true_kw, steps)
    with Adjoint(_settings(mode, start_kw, "fp64"), steps, 25) as a:
        _, g = loss_and_gradient(a, tx, ty)
        om = torch.tensor(start_kw["OMEGA"], dtype=torch.float64, requires_grad=True)
        fo = torch.tensor(start_kw["BODY_FORCE"], dtype=torch.float64, requires_grad=True)
        uw = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
        ux, uy = LatticeVelocity.apply(om, fo, uw, a)
        loss = torch.sum((ux - torch.from_numpy(tx)) ** 2 + (uy - torch.from_numpy(ty)) ** 2)
        loss.backward()
    assert float(om.grad) == pytest.approx(g[0], rel=1e-12)
    assert float(fo.grad) == pytest.approx(g[1], rel=1e-12)


def test_torch_function_on_device_equals_host_path():
    """Parameters on the GPU: fields and seeds stay in device memory and the
    gradients equal those of the host path bit for bit (same seed values)."""
    torch = pytest.importorskip("torch")
    from fastfortranboltzmann_amd.adjoint import Adjoint, LatticeVelocity

    print("[gpu] torch autograd function on device", flush=True)
    mode, true_kw, start_kw, _ = CASES[0]
    steps = 100
    tx, ty = _target(mode, true_kw, steps)
    grads = {}
    for dev in ("cpu", "cuda"):
        with Adjoint(_settings(mode, start_kw, "fp64"), steps, 25) as a:
            om = torch.tensor(start_kw["OMEGA"], dtype=torch.float64, device=dev, requires_grad=True)
            fo = torch.tensor(start_kw["BODY_FORCE"], dtype=torch.float64, device=dev, requires_grad=True)
            uw = torch.tensor(0.0, dtype=torch.float64, device=dev, requires_grad=True)
            ux, uy = LatticeVelocity.apply(om, fo, uw, a)
            assert ux.device.type == dev
            t_x = torch.from_numpy(tx).to(dev)
            t_y = torch.from_numpy(ty).to(dev)
            loss = torch.sum((ux - t_x) ** 2 + (uy - t_y) ** 2)
            loss.backward()
            grads[dev] = (float(om.grad), float(fo.grad), float(loss))
    assert grads["cuda"][:2] == grads["cpu"][:2], grads
    assert grads["cuda"][2] == pytest.approx(grads["cpu"][2], rel=1e-12)


def test_adjoint_cost_against_forward():
    """Reverse sweep (recomputation + adjoint steps) against the forward pass
    with checkpoints, 512^2 channel, FP32, 2000 steps, checkpoint every 100."""
    from fastfortranboltzmann_amd.adjoint import Adjoint

    print("[gpu] adjoint cost 512^2", flush=True)
    with Adjoint(_settings(5, dict(OMEGA=1.2, BODY_FORCE=1e-6), "fp32", 512, 512), 2000, 100) as a:
        _, ux, uy = a.forward()
        a.backward(np.ones_like(ux), np.zeros_like(uy))
        _, ux, uy = a.forward()
        a.backward(np.ones_like(ux), np.zeros_like(uy))
        f_ms, b_ms = a.timing()
    print(f"  forward {f_ms:.2f} ms, backward {b_ms:.2f} ms, ratio {b_ms / f_ms:.2f}", flush=True)
    assert b_ms / f_ms <= 3.0  # the differentiation form's cost target


def test_inverse_channel_512_gradients():
    """The inverse channel problem at its full size: 512^2 body-force channel,
    5000 steps, a checkpoint every 100 steps, loss ||u - u_target||^2 with the
    target from the forward solver at seeded true (omega, g) (seed 0: omega
    1.43696, g 3.464e-7) and the gradient taken at the perturbed start.  The
    FP64 adjoint is within 1e-6 of central finite differences of the FP64
    forward (h = 1e-4 relative); the FP32-deviation adjoint within 5e-3 of the
    FP64 gradient (measured 9e-4 / 2e-4); plain FP32, whose forward misses the
    force's few-ulp changes of f ~ w, within 0.2 (measured 7e-2 / 5e-2)."""
    from fastfortranboltzmann_amd.adjoint import gradient_check

    print("[gpu] inverse channel 512^2 gradient check", flush=True)
    rng = np.random.Generator(np.random.PCG64(0))
    t_om, t_g = float(rng.uniform(0.8, 1.8)), float(10 ** rng.uniform(-7, -5))
    gc = gradient_check(512, 5000, 100, t_om * 0.85 + 0.1, t_g * 1.6, t_om, t_g, rel_h=1e-4)
    print({k: v for k, v in gc.items() if k.startswith("rel_err")}, flush=True)
    for k in ("d_omega", "d_force"):
        assert gc[f"rel_err_fp64_vs_fd_{k}"] < 1e-6
        assert gc[f"rel_err_fp32dev_vs_fp64_{k}"] < 5e-3
        assert gc[f"rel_err_fp32_vs_fp64_{k}"] < 0.2


def test_inverse_channel_512_adam_recovers_omega_and_force():
    """The 50 Adam iterations of the inverse channel problem (512^2, 5000 steps,
    checkpoint every 100, FP32-deviation forward; seed 0, start 0.85 omega + 0.1
    and 1.6 g) recover omega and g within 5 % (measured 1.3 % and 0.7 %; the
    loss is nearly flat along omega, whose first Adam steps move it away) and
    lower the loss by three orders of magnitude."""
    from fastfortranboltzmann_amd.adjoint import inverse_channel

    print("[gpu] inverse channel 512^2 adam", flush=True)
    r = inverse_channel(512, 5000, 100, seed=0, iters=50, precision="fp32dev", log=lambda m: print(m, flush=True))
    print({k: v for k, v in r.items() if k != "history"}, flush=True)
    assert r["rel_err_omega"] < 0.05 and r["rel_err_g"] < 0.05
    assert r["loss_final"] < 1e-3 * r["loss_start"]
