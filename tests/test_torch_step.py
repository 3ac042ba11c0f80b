"""The float64 PyTorch restatement used to check the adjoint reproduces the oracle's forward step
(box and mesh), so its autograd differentiates the same dynamics."""
import numpy as np
import pytest
import torch

import torch_step
from helpers import dense_box, oracle_run
from psx import meshes


def _run_torch(sc, steps):
    x = torch.tensor(sc.pos, dtype=torch.float64)
    v = torch.tensor(sc.vel, dtype=torch.float64)
    p = sc.params
    if sc.mesh is not None:
        vv = sc.mesh.reshape(-1, 3)
        grid = torch_step.Grid(float(sc.radius.max()), vv.min(0), vv.max(0))
        nrm = meshes.normals(sc.mesh)
    else:
        grid = torch_step.Grid(float(sc.radius.max()), p.box_min, p.box_max)
        nrm = None
    k = torch.tensor(p.spring, dtype=torch.float64)
    es = torch.tensor(p.restitution, dtype=torch.float64)
    gz = torch.tensor(p.gravity[2], dtype=torch.float64)
    for _ in range(steps):
        x, v = torch_step.step(x, v, sc.radius, sc.mass, k, es, gz, dt=p.dt, gxy=p.gravity[:2],
                               walls=p.walls, box_min=p.box_min, box_max=p.box_max, grid=grid,
                               mesh=sc.mesh, mesh_normals=nrm)
    return x.numpy(), v.numpy()


def adjoint_scene(mesh: bool, n_side=8, seed=0, dtype="f64"):
    sc = dense_box(n_side=n_side, spacing=0.98, seed=seed, dtype=dtype, spring=300.0, es=0.5,
                   vscale=0.5, box=(10.0, 10.0, 14.0), walls=(0.5,) * 6)
    sc.params.dt = 1e-3
    if mesh:
        sc.mesh = meshes.box_mesh((0.0, 0.0, 0.0), (10.0, 10.0, 14.0))
    return sc


@pytest.mark.parametrize("mesh", [False, True])
def test_torch_step_matches_oracle(oracle, mesh):
    sc = adjoint_scene(mesh)
    tp, tv = _run_torch(sc, 30)
    rp, rv = oracle_run(oracle, sc, 30)
    assert np.abs(tp - rp).max() < 1e-9 and np.abs(tv - rv).max() < 1e-7
