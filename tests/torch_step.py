"""A float64 PyTorch (CPU) restatement of one step, differentiable in the state and in spring,
sphere restitution and gravity_z, with every discrete decision (cells, wall and triangle hits, pair
contacts) taken from detached values. Test infrastructure: the checker for the adjoint kernels
(autograd of this step must equal the hand-written adjoint)."""
from __future__ import annotations

import numpy as np
import torch


def _axis_index(x: np.ndarray, mn: float, h: float, n: int) -> np.ndarray:
    q = (x - mn) / h
    ok = (q >= 1.0) & (q < 9223372036854775808.0)
    idx = np.where(ok, np.floor(np.where(ok, q, 0.0)), 0.0).astype(np.int64)
    return np.minimum(idx, n - 1)


class Grid:
    def __init__(self, rmax: float, mn, mx):
        self.h = rmax * 2.0 * float(np.float32(1.05))
        self.min = np.asarray(mn, np.float64)
        self.n = (np.floor((np.asarray(mx, np.float64) - self.min) / self.h) + 1).astype(np.int64)

    def cells(self, x: np.ndarray) -> np.ndarray:
        return np.stack([_axis_index(x[:, a], self.min[a], self.h, int(self.n[a]))
                         for a in range(3)], axis=1)


def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _segment(P, a, b):
    ab = b - a
    ap = P - a
    t = _dot(ap, ab) / _dot(ab, ab)
    t = np.where(t < 1.0, t, 1.0)
    t = np.where(0.0 < t, t, 0.0)
    return a + ab * t[..., None]


def mesh_hits(x: np.ndarray, r: np.ndarray, tri: np.ndarray, nrm: np.ndarray) -> list:
    """Per particle, the ascending triangle indices it touches (PolygonContainer.cpp:244-278)."""
    A, B, C = tri[:, 0:3], tri[:, 3:6], tri[:, 6:9]
    lo = np.minimum(np.minimum(A, B), C)
    hi = np.maximum(np.maximum(A, B), C)
    P = x[:, None, :]
    rr = r[:, None]
    reject = ((P[..., 0] - rr > hi[None, :, 0]) | (P[..., 0] + rr < lo[None, :, 0]) |
              (P[..., 1] - rr > hi[None, :, 1]) | (P[..., 1] + rr < lo[None, :, 1]) |
              (P[..., 2] - rr > hi[None, :, 2]) | (P[..., 2] + rr < lo[None, :, 2]))
    v = P - A[None]
    ip = _dot(v, nrm[None])
    proj = P - nrm[None] * ip[..., None]
    v0, v1, v2 = (C - A)[None], (B - A)[None], proj - A[None]
    d00, d01, d02 = _dot(v0, v0), _dot(v0, v1), _dot(v0, v2)
    d11, d12 = _dot(v1, v1), _dot(v1, v2)
    inv = 1.0 / (d00 * d11 - d01 * d01)
    u = (d11 * d02 - d01 * d12) * inv
    w = (d00 * d12 - d01 * d02) * inv
    inside = (u >= 0) & (w >= 0) & (u + w <= 1)
    Ab, Bb, Cb = A[None].repeat(len(x), 0), B[None].repeat(len(x), 0), C[None].repeat(len(x), 0)
    q = _segment(P, Ab, Bb)
    best = np.sqrt(_dot(q - P, q - P))
    for s0, s1 in ((Bb, Cb), (Cb, Ab)):
        qq = _segment(P, s0, s1)
        dd = np.sqrt(_dot(qq - P, qq - P))
        better = dd < best
        q = np.where(better[..., None], qq, q)
        best = np.where(better, dd, best)
    q = np.where(inside[..., None], proj, q)
    dist = np.sqrt(_dot(q - P, q - P))
    hit = (~reject) & (dist < rr)
    return [np.nonzero(hit[i])[0] for i in range(len(x))]


def step(x, v, r, m, spring, restitution, gz, *, dt, gxy, walls, box_min, box_max, grid: Grid,
         mesh=None, mesh_normals=None):
    """One step on torch float64 tensors x, v (n, 3); returns (x, v) after it."""
    g = torch.stack([torch.as_tensor(gxy[0], dtype=x.dtype), torch.as_tensor(gxy[1], dtype=x.dtype),
                     gz])
    x1 = x + v * dt
    v1 = v + g * dt
    rt = torch.as_tensor(r, dtype=x.dtype)
    rn = np.asarray(r, np.float64)
    bucket = grid.cells(x1.detach().numpy())
    if mesh is None:
        lo = x1.detach().numpy() - rn[:, None]
        hi = x1.detach().numpy() + rn[:, None]
        bmin, bmax = np.asarray(box_min, np.float64), np.asarray(box_max, np.float64)
        hmin = torch.as_tensor(lo < bmin)
        hmax = torch.as_tensor(hi > bmax)
        emin = torch.as_tensor([walls[2], walls[4], walls[0]], dtype=x.dtype)
        emax = torch.as_tensor([walls[3], walls[5], walls[1]], dtype=x.dtype)
        x2 = torch.where(hmin, torch.as_tensor(bmin) + rt[:, None], x1)
        x2 = torch.where(hmax, torch.as_tensor(bmax) - rt[:, None], x2)
        v2 = v1 * torch.where(hmin, -emin, torch.ones_like(emin)) * \
            torch.where(hmax, -emax, torch.ones_like(emax))
    else:
        tri = np.asarray(mesh, np.float64).reshape(-1, 9)
        hits = mesh_hits(x1.detach().numpy(), rn, tri, mesh_normals)
        cnt = np.array([len(h) for h in hits])
        hmax_ = int(cnt.max()) if len(cnt) else 0
        nt = torch.as_tensor(mesh_normals, dtype=x.dtype)
        at = torch.as_tensor(tri[:, 0:3], dtype=x.dtype)
        e = walls[0]
        y = x1
        acc = torch.zeros_like(v1)
        for k in range(hmax_):
            idx = np.array([h[k] if len(h) > k else 0 for h in hits])
            mask = torch.as_tensor(cnt > k)[:, None]
            n = nt[idx]
            a = at[idx]
            refl = v1 - (2.0 * _tdot(v1, n))[:, None] * n
            acc = torch.where(mask, acc + e * refl, acc)
            mod = rt - _tdot(-n, y - a)
            y = torch.where(mask, y - mod[:, None] * n, y)
        has = torch.as_tensor(cnt > 0)[:, None]
        x2 = torch.where(has, y, x1)
        v2 = torch.where(has, acc / torch.as_tensor(np.maximum(cnt, 1), dtype=x.dtype)[:, None], v1)
    # pairs: i <- j when j's bucket cell is within one cell of i's home cell and they touch
    home = grid.cells(x2.detach().numpy())
    xd = x2.detach().numpy()
    dd = xd[None, :, :] - xd[:, None, :]
    dist = np.sqrt((dd[..., 0] ** 2 + dd[..., 1] ** 2) + dd[..., 2] ** 2)
    near = (np.abs(bucket[None, :, :] - home[:, None, :]) <= 1).all(-1)
    M = near & (dist < rn[:, None] + rn[None, :])
    np.fill_diagonal(M, False)
    Mt = torch.as_tensor(M)
    d = x2[:, None, :] - x2[None, :, :]
    D = torch.where(Mt, _tdot(d, d), torch.ones_like(Mt, dtype=x.dtype))
    L = torch.sqrt(D)
    u = v2[:, None, :] - v2[None, :, :]
    mt = torch.as_tensor(m, dtype=x.dtype)
    a = 2.0 * mt[None, :] / (mt[:, None] + mt[None, :])
    R = rt[:, None] + rt[None, :]
    w = _tdot(u, d)
    f = -(a * (w / D) * restitution)[..., None] * d + (spring * (R - L) / L)[..., None] * d
    f = torch.where(Mt[..., None], f, torch.zeros_like(f))
    return x2, v2 + f.sum(1)


def _tdot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]
