"""Gradients of the differentiable window (hand-written adjoint kernels) against autograd of the
float64 PyTorch restatement of the step, and against float64 central finite differences, on windows
short enough that the contact sets stay fixed under the perturbations."""
import numpy as np
import pytest
import torch

import torch_step
from psx import meshes
from psx.autograd import window
from test_torch_step import adjoint_scene

pytestmark = pytest.mark.gpu
STEPS = 50


def _weights(n, seed=3):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 3)), rng.normal(size=(n, 3))


def _hip(psx_mod, sc, wx, wv, steps=STEPS, every=10, graph=False, **tuning):
    sim = psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh)
    if graph:
        sim.set_tuning(use_graph=1)
    if tuning:
        sim.set_tuning(**tuning)
    dt = torch.float64 if sc.params.dtype == "f64" else torch.float32
    x0 = torch.tensor(sc.pos, dtype=dt, device="cuda", requires_grad=True)
    v0 = torch.tensor(sc.vel, dtype=dt, device="cuda", requires_grad=True)
    k = torch.tensor(sc.params.spring, dtype=torch.float64, device="cuda", requires_grad=True)
    es = torch.tensor(sc.params.restitution, dtype=torch.float64, device="cuda", requires_grad=True)
    gz = torch.tensor(sc.params.gravity[2], dtype=torch.float64, device="cuda", requires_grad=True)
    xN, vN = window(sim, x0, v0, k, es, gz, steps, every)
    loss = (xN.double() * torch.tensor(wx, device="cuda")).sum() + \
        (vN.double() * torch.tensor(wv, device="cuda")).sum()
    loss.backward()
    out = {"x0": x0.grad.double().cpu().numpy(), "v0": v0.grad.double().cpu().numpy(),
           "k": float(k.grad), "es": float(es.grad), "gz": float(gz.grad),
           "loss": float(loss.detach())}
    # the backward leaves the handle's state (the window's final state) as it is
    p, v = sim.state()
    assert np.array_equal(p, xN.detach().double().cpu().numpy())
    assert np.array_equal(v, vN.detach().double().cpu().numpy())
    sim.close()
    return out


def _torch(sc, wx, wv, steps=STEPS):
    p = sc.params
    x0 = torch.tensor(sc.pos, dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(sc.vel, dtype=torch.float64, requires_grad=True)
    k = torch.tensor(p.spring, dtype=torch.float64, requires_grad=True)
    es = torch.tensor(p.restitution, dtype=torch.float64, requires_grad=True)
    gz = torch.tensor(p.gravity[2], dtype=torch.float64, requires_grad=True)
    if sc.mesh is not None:
        vv = sc.mesh.reshape(-1, 3)
        grid = torch_step.Grid(float(sc.radius.max()), vv.min(0), vv.max(0))
        nrm = meshes.normals(sc.mesh)
    else:
        grid = torch_step.Grid(float(sc.radius.max()), p.box_min, p.box_max)
        nrm = None
    x, v = x0, v0
    for _ in range(steps):
        x, v = torch_step.step(x, v, sc.radius, sc.mass, k, es, gz, dt=p.dt, gxy=p.gravity[:2],
                               walls=p.walls, box_min=p.box_min, box_max=p.box_max, grid=grid,
                               mesh=sc.mesh, mesh_normals=nrm)
    loss = (x * torch.tensor(wx)).sum() + (v * torch.tensor(wv)).sum()
    loss.backward()
    return {"x0": x0.grad.numpy(), "v0": v0.grad.numpy(), "k": float(k.grad),
            "es": float(es.grad), "gz": float(gz.grad), "loss": float(loss)}


def _close(a, b, rtol):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() <= rtol * scale


@pytest.mark.parametrize("mesh", [False, True])
def test_adjoint_matches_torch_autograd(psx_mod, mesh):
    sc = adjoint_scene(mesh)
    wx, wv = _weights(sc.radius.shape[0])
    h = _hip(psx_mod, sc, wx, wv)
    t = _torch(sc, wx, wv)
    assert abs(h["loss"] - t["loss"]) <= 1e-8 * abs(t["loss"])
    for key in ("x0", "v0", "k", "es", "gz"):
        assert _close(h[key], t[key], 1e-6), (key, h[key], t[key])


def test_adjoint_checkpoint_spacing_and_graph_invariant(psx_mod):
    """Checkpoint spacing and graph replay change neither the recomputed states nor the sums."""
    sc = adjoint_scene(True)
    wx, wv = _weights(sc.radius.shape[0])
    a = _hip(psx_mod, sc, wx, wv, every=7)
    b = _hip(psx_mod, sc, wx, wv, every=50)
    c = _hip(psx_mod, sc, wx, wv, every=10, graph=True)
    for key in ("x0", "v0", "k", "es", "gz"):
        assert np.array_equal(np.asarray(a[key]), np.asarray(b[key])), key
        assert np.array_equal(np.asarray(a[key]), np.asarray(c[key])), key


@pytest.mark.parametrize("mesh", [False, True])
def test_adjoint_finite_differences(psx_mod, mesh):
    """Central differences of the float64 forward window (the GPU forward, contact sets fixed by
    the small perturbation) against the adjoint: relative error <= 1e-3."""
    sc = adjoint_scene(mesh)
    n = sc.radius.shape[0]
    wx, wv = _weights(n)
    g = _hip(psx_mod, sc, wx, wv)

    def loss_of(spring=None, es=None, gz=None, v0=None):
        p = sc.params
        prm = psx_mod.Params(gravity=(p.gravity[0], p.gravity[1], p.gravity[2] if gz is None else gz),
                             walls=p.walls, spring=p.spring if spring is None else spring,
                             restitution=p.restitution if es is None else es, dt=p.dt,
                             box_min=p.box_min, box_max=p.box_max, dtype="f64")
        with psx_mod.Simulation(sc.pos, sc.vel if v0 is None else v0, sc.radius, sc.mass, prm,
                                sc.mesh) as s:
            s.step(STEPS)
            s.synchronize()
            x, v = s.state()
        return float((x * wx).sum() + (v * wv).sum())

    def fd_fixed_contacts(f, base, eps):
        """Central difference at eps, accepted only when it agrees with the one at eps/2 (a contact
        switch inside +-eps makes the two disagree); otherwise eps is reduced."""
        for _ in range(4):
            d1 = (f(base + eps) - f(base - eps)) / (2 * eps)
            d2 = (f(base + eps / 2) - f(base - eps / 2)) / eps
            if abs(d1 - d2) <= 1e-4 * max(abs(d1), 1e-6):
                return d1
            eps /= 8
        raise AssertionError("no perturbation small enough to keep the contact sets fixed")

    p = sc.params
    # step sizes: small enough that no contact switches inside +-eps (restitution moves every
    # contact's normal velocity, so it needs the smallest), large enough that the chaotic rounding
    # of the 50-step trajectory stays below the tolerance
    checks = [("k", lambda t: loss_of(spring=t), p.spring, 1e-7 * p.spring),
              ("es", lambda t: loss_of(es=t), p.restitution, 1e-7),
              ("gz", lambda t: loss_of(gz=t), p.gravity[2], 1e-4)]
    for key, f, b, eps in checks:
        fd = fd_fixed_contacts(f, b, eps)
        assert abs(fd - g[key]) <= 1e-3 * max(abs(fd), 1e-6), (key, fd, g[key])
    rng = np.random.default_rng(4)
    for i in rng.choice(n, 3, replace=False):
        for c in range(3):
            def f(t, i=i, c=c):
                vv = sc.vel.copy()
                vv[i, c] = t
                return loss_of(v0=vv)
            fd = fd_fixed_contacts(f, sc.vel[i, c], 1e-5)
            assert abs(fd - g["v0"][i, c]) <= 1e-3 * max(abs(fd), 1e-3), (i, c, fd, g["v0"][i, c])


def test_adjoint_on_grid_above_direct_scan_limit(psx_mod):
    """The window on a box of 104^3 = 1.12M cells (more than the direct scan's 2^20): the forward
    checkpoints and the backward's recomputed steps take the look-back scan, whose hash passes carry
    no tile counts. Gradients against autograd of the float64 restatement, and eager against
    graph-replayed recomputation bit for bit."""
    sc = adjoint_scene(False)
    sc.params.box_max = (120.0, 120.0, 120.0)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        assert s.num_cells > (1 << 20)
    wx, wv = _weights(sc.radius.shape[0])
    h = _hip(psx_mod, sc, wx, wv)
    t = _torch(sc, wx, wv)
    assert abs(h["loss"] - t["loss"]) <= 1e-8 * abs(t["loss"])
    for key in ("x0", "v0", "k", "es", "gz"):
        assert _close(h[key], t[key], 1e-6), (key, h[key], t[key])
    g = _hip(psx_mod, sc, wx, wv, every=7, graph=True)
    for key in ("x0", "v0", "k", "es", "gz"):
        assert np.array_equal(np.asarray(h[key]), np.asarray(g[key])), key
