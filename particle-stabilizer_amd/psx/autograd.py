"""PyTorch shim over the differentiable window of the HIP backend.

`window(sim, x0, v0, spring, restitution, gravity_z, nsteps)` runs nsteps steps of the hot path
(psx_window_forward) and returns the final positions and velocities as tensors on the GPU; its
backward (psx_window_backward) runs the hand-written adjoint kernels, recomputing each segment
between checkpoints, and returns dL/dx0, dL/dv0, dL/dspring, dL/drestitution, dL/dgravity_z.
Contact sets of every step are held fixed, as the gradient of the piecewise-smooth step requires.

Streams: the library runs on its own stream; torch's current stream is synchronised before each
library call and the library's stream after it, so tensors passed by device pointer are complete
on both sides.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from .sim import Params, Simulation


def _torch_done(dev) -> None:
    torch.cuda.current_stream(dev).synchronize()


class _Window(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, spring, restitution, gravity_z, sim, nsteps, every):
        dt = torch.float32 if sim.dtype == np.float32 else torch.float64
        dev = x0.device
        p = dataclasses.replace(sim.params, spring=float(spring), restitution=float(restitution),
                                gravity=(sim.params.gravity[0], sim.params.gravity[1],
                                         float(gravity_z)))
        sim.set_params(p)
        x0c = x0.detach().to(dt).contiguous()
        v0c = v0.detach().to(dt).contiguous()
        xN = torch.empty_like(x0c)
        vN = torch.empty_like(v0c)
        _torch_done(dev)
        sim.set_state_device(x0c.data_ptr(), v0c.data_ptr())
        wid = sim.window_forward(nsteps, every)
        sim.get_state_device(xN.data_ptr(), vN.data_ptr())
        sim.synchronize()
        ctx.sim = sim
        # the handle keeps the checkpoints of its latest window only: a backward after another
        # window (or a coefficient change) on the same handle is refused, not silently wrong
        ctx.window_id = wid
        ctx.dtypes = (x0.dtype, v0.dtype)
        return xN, vN

    @staticmethod
    def backward(ctx, gx, gv):
        sim = ctx.sim
        dt = torch.float32 if sim.dtype == np.float32 else torch.float64
        gx = (torch.zeros((sim.n, 3), device="cuda", dtype=dt) if gx is None
              else gx.to(dt).contiguous())
        gv = (torch.zeros((sim.n, 3), device=gx.device, dtype=dt) if gv is None
              else gv.to(dt).contiguous())
        gx0 = torch.empty_like(gx)
        gv0 = torch.empty_like(gv)
        _torch_done(gx.device)
        gp = sim.window_backward(ctx.window_id, gx.data_ptr(), gv.data_ptr(), gx0.data_ptr(),
                                 gv0.data_ptr())[0]
        sim.synchronize()
        t = lambda v: torch.tensor(v, dtype=torch.float64, device=gx.device)  # noqa: E731
        return (gx0.to(ctx.dtypes[0]), gv0.to(ctx.dtypes[1]), t(gp[0]), t(gp[1]), t(gp[2]),
                None, None, None)


def window(sim: Simulation, x0: torch.Tensor, v0: torch.Tensor, spring, restitution, gravity_z,
           nsteps: int, checkpoint_every: int = 100):
    """Differentiable window of nsteps steps; returns (positions, velocities) after the window."""
    as_t = lambda v: v if torch.is_tensor(v) else torch.tensor(float(v), dtype=torch.float64)  # noqa: E731
    return _Window.apply(x0, v0, as_t(spring), as_t(restitution), as_t(gravity_z), sim,
                         int(nsteps), int(checkpoint_every))


def soft_histogram(z: torch.Tensor, lo: float, hi: float, bins: int = 64,
                   sigma: float | None = None) -> torch.Tensor:
    """Gaussian-kernel histogram of heights, normalised to sum to one."""
    centers = torch.linspace(lo, hi, bins, device=z.device, dtype=z.dtype)
    sigma = sigma or (hi - lo) / bins
    w = torch.exp(-0.5 * ((z[:, None] - centers[None, :]) / sigma) ** 2)
    h = w.sum(0)
    return h / h.sum()


def packing_loss(x: torch.Tensor, target_hist: torch.Tensor, target_pos: torch.Tensor | None,
                 lo: float, hi: float, pos_weight: float = 1.0) -> torch.Tensor:
    """L2 of the soft height-histogram mismatch plus the mean-square final-position error."""
    h = soft_histogram(x[:, 2], lo, hi, target_hist.shape[0])
    loss = ((h - target_hist) ** 2).sum()
    if target_pos is not None:
        loss = loss + pos_weight * ((x - target_pos) ** 2).sum(1).mean()
    return loss


@dataclasses.dataclass
class CalibrationResult:
    spring: float
    restitution: float
    gravity_z: float
    v0: np.ndarray
    losses: list
    grads: list
    best_loss: float


def calibrate(sim: Simulation, x0: np.ndarray, v0: np.ndarray, target_hist: torch.Tensor,
              target_pos: torch.Tensor | None, nsteps: int, spring: float, restitution: float,
              gravity_z: float, iters: int = 10, lr: float = 0.05, lr_v0: float | None = None,
              lo: float = 0.0, hi: float = 30.0, checkpoint_every: int = 100) -> CalibrationResult:
    """Fits spring, restitution, gravity_z and the initial velocities against a target packing
    (packing_loss), the window's adjoint supplying every gradient: Adam on (log spring,
    restitution, gravity_z) with step lr and on v0 with step lr_v0 (default lr / 10).

    Contact dynamics make the loss piecewise smooth and, over long windows, rough: a step that
    raises the loss is undone and both steps are halved (the gradient at the best point is a
    descent direction for a small enough step), so the best loss never rises. Every evaluated loss
    is returned in `losses`, the parameters of the best one in the result."""
    dev = torch.device("cuda")
    dt = torch.float32 if sim.dtype == np.float32 else torch.float64
    x0t = torch.tensor(x0, dtype=dt, device=dev)
    v0t = torch.tensor(v0, dtype=dt, device=dev, requires_grad=True)
    logk = torch.tensor(np.log(spring), dtype=torch.float64, device=dev, requires_grad=True)
    es = torch.tensor(restitution, dtype=torch.float64, device=dev, requires_grad=True)
    gz = torch.tensor(gravity_z, dtype=torch.float64, device=dev, requires_grad=True)
    params = [logk, es, gz, v0t]
    steps = [lr, lr, lr, lr / 10.0 if lr_v0 is None else lr_v0]

    def make_opt():
        return torch.optim.Adam([{"params": [q], "lr": a} for q, a in zip(params, steps)])

    opt = make_opt()
    losses, grads = [], []
    best, best_loss = None, float("inf")
    for _ in range(iters):
        opt.zero_grad()
        xN, _ = window(sim, x0t, v0t, torch.exp(logk), es, gz, nsteps, checkpoint_every)
        loss = packing_loss(xN.to(torch.float64), target_hist, target_pos, lo, hi)
        loss.backward()
        lv = float(loss)
        losses.append(lv)
        grads.append((float(logk.grad), float(es.grad), float(gz.grad)))
        if lv < best_loss:
            best_loss = lv
            best = [(q.detach().clone(), q.grad.detach().clone()) for q in params]
        else:  # undo: back to the best point with its gradient, half the steps
            steps = [a * 0.5 for a in steps]
            opt = make_opt()
            with torch.no_grad():
                for q, (val, g) in zip(params, best):
                    q.copy_(val)
                    q.grad = g.clone()
        opt.step()
        with torch.no_grad():
            es.clamp_(0.0, 1.0)
    with torch.no_grad():
        for q, (val, _) in zip(params, best):
            q.copy_(val)
    return CalibrationResult(float(torch.exp(logk)), float(es), float(gz),
                             v0t.detach().double().cpu().numpy(), losses, grads, best_loss)
