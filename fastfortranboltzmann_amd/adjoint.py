"""Differentiation form: gradients of a loss of the final velocity field with
respect to omega, the body force and the wall/lid speed, from the hand-written
adjoint in libffb_hip.so (ffb_adjoint_* in include/ffb.h; checkpoint every K
steps, recomputation inside each segment).  No reference counterpart.

``LatticeVelocity`` is a ``torch.autograd.Function``: its forward runs the
solver for n_steps through the C ABI and returns (u_x, u_y); its backward
hands dL/du to the adjoint and returns dL/domega, dL/dF, dL/du_wall.  Torch is
only the autograd and tensor plumbing here; every step and every adjoint
step runs in the HIP kernels.  With parameters on a GPU device the velocity
fields and their gradients stay in device memory: the library copies the
macro fields into torch's tensors and reads the seed from them directly.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib
from .lattice import Lattice, Settings


class Adjoint:
    """A lattice plus its adjoint workspace for a fixed number of steps."""

    def __init__(self, settings: Settings, n_steps: int, checkpoint_interval: int = 100):
        self.lattice = Lattice(settings)
        self.settings = settings
        self.n_steps = int(n_steps)
        self.nx, self.ny = int(settings.N_X), int(settings.N_Y)
        self.dtype = self.lattice.dtype
        h = ctypes.c_void_p()
        check(lib().ffb_adjoint_create(self.lattice.handle, self.n_steps, int(checkpoint_interval),
                                       ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ffb_adjoint_destroy(self._h)
        self._h = None
        if getattr(self, "lattice", None) is not None:
            self.lattice.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_parameters(self, omega=None, body_force=None, u_wall=None):
        s = self.settings
        if omega is not None:
            s.OMEGA = float(omega)
        if body_force is not None:
            s.BODY_FORCE = float(body_force)
        if u_wall is not None:
            s.U_WALL = float(u_wall)
            s.U_LID = float(u_wall)
        check(lib().ffb_set_parameters(self.lattice.handle, s.OMEGA, s.BODY_FORCE, s.U_WALL, s.U_LID))

    def forward(self):
        """Runs the n_steps steps; returns (rho, u_x, u_y) of the last step (numpy)."""
        rho = np.empty((self.ny, self.nx), self.dtype)
        ux, uy = np.empty_like(rho), np.empty_like(rho)
        check(lib().ffb_adjoint_forward(self._h, ctypes.c_void_p(rho.ctypes.data),
                                        ctypes.c_void_p(ux.ctypes.data), ctypes.c_void_p(uy.ctypes.data)))
        return rho, ux, uy

    def forward_device(self, device):
        """As forward, into torch tensors on a GPU device (no host copy); returns
        (u_x, u_y) in the lattice's element type."""
        dt = torch.float64 if self.dtype == np.float64 else torch.float32
        ux = torch.empty((self.ny, self.nx), dtype=dt, device=device)
        uy = torch.empty_like(ux)
        torch.cuda.current_stream(device).synchronize()  # the library works on its own stream
        check(lib().ffb_adjoint_forward(self._h, None, ctypes.c_void_p(ux.data_ptr()),
                                        ctypes.c_void_p(uy.data_ptr())))
        return ux, uy

    def backward_device(self, grad_u_x, grad_u_y):
        """As backward, for seeds held in torch tensors on a GPU device."""
        dt = torch.float64 if self.dtype == np.float64 else torch.float32
        gx = grad_u_x.detach().to(dt).contiguous()
        gy = grad_u_y.detach().to(dt).contiguous()
        torch.cuda.current_stream(gx.device).synchronize()  # seeds written by torch's stream
        d = [ctypes.c_double() for _ in range(3)]
        check(lib().ffb_adjoint_backward(self._h, ctypes.c_void_p(gx.data_ptr()), ctypes.c_void_p(gy.data_ptr()),
                                         *[ctypes.byref(x) for x in d]))
        return d[0].value, d[1].value, d[2].value

    def backward(self, grad_u_x: np.ndarray, grad_u_y: np.ndarray):
        """dL/domega, dL/d(body force), dL/d(wall speed) for the seed dL/du."""
        gx = np.ascontiguousarray(grad_u_x, dtype=self.dtype)
        gy = np.ascontiguousarray(grad_u_y, dtype=self.dtype)
        d = [ctypes.c_double() for _ in range(3)]
        check(lib().ffb_adjoint_backward(self._h, ctypes.c_void_p(gx.ctypes.data),
                                         ctypes.c_void_p(gy.ctypes.data), *[ctypes.byref(x) for x in d]))
        return d[0].value, d[1].value, d[2].value

    def timing(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        check(lib().ffb_adjoint_timing(self._h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value


def velocity_loss(ux, uy, tx, ty) -> float:
    """L = sum over cells of |u - u_target|^2 (FP64 sum)."""
    return float(np.sum((ux.astype(np.float64) - tx) ** 2 + (uy.astype(np.float64) - ty) ** 2))


def loss_and_gradient(adj: Adjoint, tx, ty):
    """Loss ||u - u_target||^2 and its gradient (d omega, d F, d u_wall)."""
    _, ux, uy = adj.forward()
    loss = velocity_loss(ux, uy, tx, ty)
    g = adj.backward(2.0 * (ux.astype(np.float64) - tx), 2.0 * (uy.astype(np.float64) - ty))
    return loss, g


def finite_difference(adj: Adjoint, tx, ty, which: str, rel_h: float = 1e-4) -> float:
    """Central difference of the loss in one parameter (h = rel_h * |value|)."""
    s = adj.settings
    base = {"omega": s.OMEGA, "body_force": s.BODY_FORCE, "u_wall": s.U_WALL}[which]
    h = rel_h * abs(base)
    vals = []
    for sign in (1, -1):
        adj.set_parameters(**{which: base + sign * h})
        _, ux, uy = adj.forward()
        vals.append(velocity_loss(ux, uy, tx, ty))
    adj.set_parameters(**{which: base})
    return (vals[0] - vals[1]) / (2 * h)


try:  # torch is plumbing only; the Function exists where torch is importable
    import torch

    class LatticeVelocity(torch.autograd.Function):
        """(omega, body_force, u_wall) -> (u_x, u_y) after adj.n_steps steps."""

        @staticmethod
        def forward(ctx, omega, body_force, u_wall, adj: Adjoint):
            adj.set_parameters(float(omega), float(body_force), float(u_wall))
            ctx.adj = adj
            dev = omega.device
            if dev.type == "cuda":  # device-resident fields, no host round trip
                ux, uy = adj.forward_device(dev)
                return ux.to(torch.float64), uy.to(torch.float64)
            _, ux, uy = adj.forward()
            return (torch.from_numpy(ux.astype(np.float64)).to(dev),
                    torch.from_numpy(uy.astype(np.float64)).to(dev))

        @staticmethod
        def backward(ctx, grad_ux, grad_uy):
            adj = ctx.adj
            if grad_ux.device.type == "cuda":
                d_om, d_f, d_w = adj.backward_device(grad_ux, grad_uy)
            else:
                d_om, d_f, d_w = adj.backward(grad_ux.detach().cpu().numpy(), grad_uy.detach().cpu().numpy())
            dev = grad_ux.device
            t = lambda v: torch.tensor(v, dtype=torch.float64, device=dev)  # noqa: E731
            return t(d_om), t(d_f), t(d_w), None
except ImportError:  # pragma: no cover
    torch = None
    LatticeVelocity = None


def channel_settings(n: int, omega: float, force: float, precision: str) -> Settings:
    return Settings(N_X=n, N_Y=n, SIM_MODE=5, OMEGA=omega, BODY_FORCE=force, precision=precision)


def inverse_channel(n: int = 512, steps: int = 5000, interval: int = 100, seed: int = 0,
                    iters: int = 50, lr: float = 0.08, precision: str = "fp32", log=print,
                    lr_decay: float = 0.85, decay_start: int = 30) -> dict:
    """Inverse body-force channel: the target u field is the forward solution at
    seeded true (omega, g) with omega in [0.8, 1.8], g in [1e-7, 1e-5]; Adam on
    (omega, g / 1e-6) from a perturbed start recovers them.  Learning rate lr
    for the first decay_start iterations, then times lr_decay per iteration."""
    rng = np.random.Generator(np.random.PCG64(seed))
    om_true = float(rng.uniform(0.8, 1.8))
    g_true = float(10 ** rng.uniform(-7, -5))
    with Adjoint(channel_settings(n, om_true, g_true, precision), steps, interval) as adj:
        _, tx, ty = adj.forward()
        tx, ty = tx.astype(np.float64), ty.astype(np.float64)
        om0, g0 = om_true * 0.85 + 0.1, g_true * 1.6
        theta = torch.tensor([om0, g0 / 1e-6], dtype=torch.float64, requires_grad=True)
        opt = torch.optim.Adam([theta], lr=lr)
        sched = torch.optim.lr_scheduler.LambdaLR(
            opt, lambda k: 1.0 if k < decay_start else lr_decay ** (k - decay_start + 1))
        hist = []
        for it in range(iters):
            opt.zero_grad()
            ux, uy = LatticeVelocity.apply(theta[0], theta[1] * 1e-6, torch.tensor(0.0, dtype=torch.float64), adj)
            loss = torch.sum((ux - torch.from_numpy(tx)) ** 2 + (uy - torch.from_numpy(ty)) ** 2)
            loss.backward()
            opt.step()
            sched.step()
            with torch.no_grad():
                theta[0].clamp_(0.55, 1.95)
                theta[1].clamp_(1e-3, 100.0)
            hist.append((float(loss), float(theta[0]), float(theta[1]) * 1e-6))
            if it % 10 == 0 or it == iters - 1:
                log(f"adam {it}: loss {hist[-1][0]:.3e} omega {hist[-1][1]:.6f} g {hist[-1][2]:.4e}")
    return {"precision": precision, "omega_true": om_true, "g_true": g_true, "omega_start": om0, "g_start": g0,
            "omega_final": hist[-1][1], "g_final": hist[-1][2], "loss_start": hist[0][0],
            "loss_final": hist[-1][0], "rel_err_omega": abs(hist[-1][1] - om_true) / om_true,
            "rel_err_g": abs(hist[-1][2] - g_true) / g_true, "history": hist}


def gradient_check(n: int, steps: int, interval: int, omega: float, force: float,
                   t_omega: float, t_force: float, rel_h: float = 1e-4) -> dict:
    """Adjoint gradients of ||u - u_target||^2 in FP64, FP32 and FP32
    deviations against central differences of the FP64 forward solver
    (h = rel_h relative) and against each other."""
    out = {}
    with Adjoint(channel_settings(n, t_omega, t_force, "fp64"), steps, interval) as a:
        _, tx, ty = a.forward()
    for prec in ("fp64", "fp32", "fp32dev"):
        with Adjoint(channel_settings(n, omega, force, prec), steps, interval) as a:
            loss, g = loss_and_gradient(a, tx, ty)
            out[prec] = {"loss": loss, "d_omega": g[0], "d_force": g[1]}
            if prec == "fp64":
                out["fd"] = {"d_omega": finite_difference(a, tx, ty, "omega", rel_h),
                             "d_force": finite_difference(a, tx, ty, "body_force", rel_h)}
    for k in ("d_omega", "d_force"):
        fd = out["fd"][k]
        out[f"rel_err_fp64_vs_fd_{k}"] = abs(out["fp64"][k] - fd) / abs(fd)
        for prec in ("fp32", "fp32dev"):
            out[f"rel_err_{prec}_vs_fp64_{k}"] = abs(out[prec][k] - out["fp64"][k]) / abs(out["fp64"][k])
    return out


def main(argv=None) -> int:
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description="inverse body-force channel: gradient check and Adam recovery")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--interval", type=int, default=100)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--lr", type=float, default=0.08)
    ap.add_argument("--lr-decay", type=float, default=0.85)
    ap.add_argument("--decay-start", type=int, default=30)
    ap.add_argument("--precisions", default="fp32,fp32dev", help="forward storage of the Adam runs")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    rng = np.random.Generator(np.random.PCG64(args.seed))
    t_om, t_g = float(rng.uniform(0.8, 1.8)), float(10 ** rng.uniform(-7, -5))
    log(f"gradient check at ({t_om * 0.85 + 0.1:.4f}, {t_g * 1.6:.3e}) against target ({t_om:.4f}, {t_g:.3e})")
    gc = gradient_check(args.n, args.steps, args.interval, t_om * 0.85 + 0.1, t_g * 1.6, t_om, t_g)
    log(json.dumps(gc))
    inv = {}
    for prec in args.precisions.split(","):
        inv[prec] = inverse_channel(args.n, args.steps, args.interval, args.seed, args.iters, args.lr, prec, log,
                                    args.lr_decay, args.decay_start)
        inv[prec].pop("history")
        log(json.dumps(inv[prec]))
    with Adjoint(channel_settings(args.n, t_om, t_g, "fp32"), args.steps, args.interval) as a:
        for _ in range(2):
            _, ux, uy = a.forward()
            a.backward(np.ones_like(ux), np.zeros_like(uy))
        f_ms, b_ms = a.timing()
    print(json.dumps({"gradient_check": gc, "inverse": inv,
                      "cost": {"forward_ms": f_ms, "backward_ms": b_ms, "ratio": b_ms / f_ms}}))
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
