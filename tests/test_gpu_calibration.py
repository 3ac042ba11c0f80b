"""The differentiable window on the calibration workload and its guard rails:

- window ids: a backward after a newer window, or after a coefficient change, is refused instead
  of using the newer checkpoints (at the C-ABI wrapper and through torch.autograd);
- the full calibration window (16,384 spheres in the 80-triangle box mesh, 2,000 steps): gradients
  independent of the checkpoint spacing and of graph replay, bit for bit, with contacts in the
  window (spring and restitution gradients nonzero);
- calibration against a target packing built from spring 8e3 and restitution 0.6 (64-bin soft
  height histogram plus final-position error): the loss falls under the adjoint's gradients.
"""
import numpy as np
import pytest
import torch

from psx import PsxError, scenes
from psx.autograd import calibrate, soft_histogram, window

pytestmark = pytest.mark.gpu

PSX_ERR_STATE = 4


def _cotangents(n, seed=2):
    g = torch.Generator(device="cpu").manual_seed(seed)
    gx = torch.randn((n, 3), generator=g, dtype=torch.float32).cuda()
    gv = torch.randn((n, 3), generator=g, dtype=torch.float32).cuda()
    return gx, gv


def test_stale_window_is_refused(psx_mod):
    sc = scenes.calibration_scene(n_xy=8, n_z=4)
    n = sc.radius.shape[0]
    gx, gv = _cotangents(n)
    ox, ov = torch.empty_like(gx), torch.empty_like(gv)
    torch.cuda.synchronize()
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        w1 = s.window_forward(50, 10)
        w2 = s.window_forward(30, 10)
        assert w2 != w1
        with pytest.raises(PsxError) as e:
            s.window_backward(w1, gx.data_ptr(), gv.data_ptr(), ox.data_ptr(), ov.data_ptr())
        assert e.value.code == PSX_ERR_STATE and "stale" in str(e.value)
        g = s.window_backward(w2, gx.data_ptr(), gv.data_ptr(), ox.data_ptr(), ov.data_ptr())
        assert np.isfinite(g).all()
        # a backward leaves the state alone: the handle still holds window 2's final state
        p, _ = s.state()
        s.set_params(sc.params)  # the coefficients the window ran with are no longer held
        with pytest.raises(PsxError) as e:
            s.window_backward(w2, gx.data_ptr(), gv.data_ptr(), ox.data_ptr(), ov.data_ptr())
        assert e.value.code == PSX_ERR_STATE
        p2, _ = s.state()
        assert np.array_equal(p, p2)


def test_stale_window_through_autograd(psx_mod):
    sc = scenes.calibration_scene(n_xy=8, n_z=4)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        x0 = torch.tensor(sc.pos, dtype=torch.float32, device="cuda")
        v0 = torch.tensor(sc.vel, dtype=torch.float32, device="cuda", requires_grad=True)
        x1, _ = window(s, x0, v0, 1e4, 0.5, -9.8, 40, 10)
        x2, _ = window(s, x0, v0, 1e4, 0.5, -9.8, 40, 10)
        with pytest.raises(PsxError):
            x1.sum().backward()
        x2.sum().backward()
        assert torch.isfinite(v0.grad).all()


def _full_window(psx_mod, sc, every, graph, nsteps=2000):
    n = sc.radius.shape[0]
    gx, gv = _cotangents(n)
    ox, ov = torch.empty_like(gx), torch.empty_like(gv)
    torch.cuda.synchronize()
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        s.set_tuning(use_graph=1 if graph else 0)
        wid = s.window_forward(nsteps, every)
        g = s.window_backward(wid, gx.data_ptr(), gv.data_ptr(), ox.data_ptr(), ov.data_ptr())
        s.synchronize()
        st = s.statistics()[0]
    return ox.cpu().numpy(), ov.cpu().numpy(), g[0], st


def test_full_window_checkpoint_spacing_and_graph_invariant(psx_mod):
    sc = scenes.calibration_scene()
    assert sc.radius.shape[0] == 16384 and sc.mesh.shape[0] == 80
    a = _full_window(psx_mod, sc, 100, True)
    b = _full_window(psx_mod, sc, 100, False)
    c = _full_window(psx_mod, sc, 250, True)
    for x, y, what in ((a, b, "graph vs eager"), (a, c, "checkpoints 100 vs 250")):
        assert np.array_equal(x[0], y[0]), f"dL/dx0 {what}"
        assert np.array_equal(x[1], y[1]), f"dL/dv0 {what}"
        assert np.array_equal(x[2], y[2]), f"dL/dparams {what}"
    assert np.isfinite(a[0]).all() and np.isfinite(a[1]).all() and np.isfinite(a[2]).all()
    # the window holds contacts (the lattice lands on the floor and the layers above run into
    # it): spring and restitution act, so their gradients are not zero
    assert a[2][0] != 0.0 and a[2][1] != 0.0


def test_calibration_loss_falls(psx_mod):
    """A calibration-shaped scene at reduced size (12 x 12 x 8 lattice in the 80-triangle box mesh,
    1,500-step window, checkpoints every 100): the target packing comes from spring 8e3 and
    restitution 0.6; starting from spring 1e4 and restitution 0.5, Adam steps on log-spring,
    restitution, gravity_z and the initial velocities (undone and halved when they raise the
    loss) lower the loss."""
    nsteps, lo, hi = 1500, 0.0, 12.0
    sc = scenes.calibration_scene(n_xy=12, n_z=8)
    target = scenes.calibration_scene(n_xy=12, n_z=8, spring=8e3, restitution=0.6)
    with psx_mod.Simulation(target.pos, target.vel, target.radius, target.mass, target.params,
                            target.mesh) as s:
        s.step(nsteps)
        tp, _ = s.state()
    tpos = torch.tensor(tp, dtype=torch.float64, device="cuda")
    thist = soft_histogram(tpos[:, 2], lo, hi, 64)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        res = calibrate(s, sc.pos, sc.vel, thist, tpos, nsteps, spring=1e4, restitution=0.5,
                        gravity_z=-9.8, iters=10, lr=0.01, lo=lo, hi=hi, checkpoint_every=100)
    print("calibration losses", res.losses, "gradients", res.grads)
    assert all(np.isfinite(res.losses)) and all(np.isfinite(g).all() for g in map(np.asarray, res.grads))
    assert res.losses[0] > 0.0
    assert res.best_loss == min(res.losses) < res.losses[0]
