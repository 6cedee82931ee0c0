"""Synthetic initial conditions of the benchmark and test workloads (numpy, host).

* ``turbulence(side)``: jittered periodic lattice (sigma = 0.1 spacing) in the unit box,
  velocities a Gaussian random field with P(k) ~ k^(-11/3) for 1 <= |k| <= 3
  (direct sum over the integer modes), scaled to rms Mach 0.3 at u = 1, gamma = 5/3;
  h for ~100 neighbors.
* ``uniform_random(n)``: i.i.d. uniform positions in the periodic unit cube,
  v ~ N(0, 0.1^2), u = 1; h = 0.5 (ng / (4/3 pi n))^(1/3) * 2 for ``ng`` neighbors
  in the unit box.
* ``evrard(n)``: open-boundary sphere of radius 1 with rho ~ 1/r (r = sqrt(U)),
  u = 0.05, v = 0, h from the local density for ~100 neighbors.
* ``jittered_lattice(side, sigma)``: periodic lattice with Gaussian jitter, fixed h.

All return (fields dict of float64 arrays x, y, z, h, m, vx, vy, vz, u; box tuple
(lo, hi, periodic)).
"""
from __future__ import annotations

import numpy as np

GAMMA = 5.0 / 3.0


def h_for_neighbors(n, ng=100, volume=1.0):
    """Smoothing length whose 2h sphere holds ~ng particles at mean density n/volume."""
    return 0.5 * (3.0 * ng * volume / (4.0 * np.pi * n)) ** (1.0 / 3.0)


def jittered_lattice(side, sigma=0.1, seed=20261202, ng=100):
    rng = np.random.default_rng(seed)
    n = side ** 3
    c = (np.arange(side) + 0.5) / side
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    pos = []
    for A in (X, Y, Z):
        a = A.ravel() + rng.normal(0.0, sigma / side, n)
        pos.append(a - np.floor(a))
    h = np.full(n, h_for_neighbors(n, ng))
    m = np.full(n, 1.0 / n)
    zero = np.zeros(n)
    return dict(x=pos[0], y=pos[1], z=pos[2], h=h, m=m, vx=zero, vy=zero.copy(), vz=zero.copy(),
                u=np.ones(n)), (0.0, 1.0, True)


def turbulent_velocity(x, y, z, seed=7, kmin=1.0, kmax=3.0, slope=-11.0 / 3.0, chunk=1 << 20):
    """Gaussian random velocity field sum_k Re(c_k exp(2 pi i k.x)), |c_k|^2 ~ P(|k|)."""
    rng = np.random.default_rng(seed)
    kk = int(np.ceil(kmax))
    modes = []
    for a in range(-kk, kk + 1):
        for b in range(-kk, kk + 1):
            for c in range(-kk, kk + 1):
                k = np.sqrt(a * a + b * b + c * c)
                # half space: keep one of (k, -k)
                if kmin <= k <= kmax and (a, b, c) > (0, 0, 0):
                    modes.append((a, b, c, k))
    K = np.array([m[:3] for m in modes], np.float64)
    amp = np.array([m[3] for m in modes]) ** (slope / 2.0)
    cr = rng.normal(size=(len(modes), 3)) * amp[:, None]
    ci = rng.normal(size=(len(modes), 3)) * amp[:, None]
    n = len(x)
    v = np.zeros((3, n))
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        phase = 2.0 * np.pi * (np.outer(K[:, 0], x[s:e]) + np.outer(K[:, 1], y[s:e]) + np.outer(K[:, 2], z[s:e]))
        cp, sp = np.cos(phase), np.sin(phase)
        for d in range(3):
            v[d, s:e] = cr[:, d] @ cp - ci[:, d] @ sp
    return v


def turbulent_velocity_torch(x, y, z, seed=7, kmin=1.0, kmax=3.0, slope=-11.0 / 3.0, chunk=1 << 21):
    """turbulent_velocity on torch tensors (same modes and coefficients), on their device."""
    import torch

    rng = np.random.default_rng(seed)
    kk = int(np.ceil(kmax))
    modes = []
    for a in range(-kk, kk + 1):
        for b in range(-kk, kk + 1):
            for c in range(-kk, kk + 1):
                k = np.sqrt(a * a + b * b + c * c)
                if kmin <= k <= kmax and (a, b, c) > (0, 0, 0):
                    modes.append((a, b, c, k))
    K = torch.tensor([m[:3] for m in modes], dtype=torch.float64, device=x.device)
    amp = np.array([m[3] for m in modes]) ** (slope / 2.0)
    cr = torch.from_numpy(rng.normal(size=(len(modes), 3)) * amp[:, None]).to(x.device)
    ci = torch.from_numpy(rng.normal(size=(len(modes), 3)) * amp[:, None]).to(x.device)
    n = x.numel()
    v = torch.zeros((3, n), dtype=torch.float64, device=x.device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        phase = 2.0 * np.pi * (K[:, 0:1] * x[s:e] + K[:, 1:2] * y[s:e] + K[:, 2:3] * z[s:e])
        cp, sp = torch.cos(phase), torch.sin(phase)
        v[:, s:e] = cr.T @ cp - ci.T @ sp
    return v


def turbulence(side=256, mach=0.3, seed=20261202, device=None):
    """Host arrays, or with ``device`` a dict of float64 torch tensors on that device."""
    f, box = jittered_lattice(side, 0.1, seed)
    if device is not None:
        import torch

        t = {k: torch.from_numpy(v).to(device) for k, v in f.items()}
        v = turbulent_velocity_torch(t["x"], t["y"], t["z"], seed + 1)
        cs = np.sqrt(GAMMA * (GAMMA - 1.0) * 1.0)
        v *= mach * cs / torch.sqrt(torch.mean(torch.sum(v * v, dim=0)))
        t["vx"], t["vy"], t["vz"] = v[0].contiguous(), v[1].contiguous(), v[2].contiguous()
        return t, box
    v = turbulent_velocity(f["x"], f["y"], f["z"], seed + 1)
    cs = np.sqrt(GAMMA * (GAMMA - 1.0) * 1.0)
    v *= mach * cs / np.sqrt(np.mean(np.sum(v * v, axis=0)))
    f["vx"], f["vy"], f["vz"] = v[0].copy(), v[1].copy(), v[2].copy()
    return f, box


def uniform_random(n, seed=20261202, ng=None):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    v = rng.normal(0.0, 0.1, (3, n))
    if ng is None:
        h = 0.5 * (100.0 / (4.0 / 3.0 * np.pi * n)) ** (1.0 / 3.0) * 2.0
    else:
        h = h_for_neighbors(n, ng)
    return dict(x=x, y=y, z=z, h=np.full(n, h), m=np.full(n, 1.0 / n), vx=v[0].copy(), vy=v[1].copy(),
                vz=v[2].copy(), u=np.ones(n)), (0.0, 1.0, True)


def evrard(n, seed=20261202):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.random(n))
    ct = 2.0 * rng.random(n) - 1.0
    phi = 2.0 * np.pi * rng.random(n)
    st = np.sqrt(1.0 - ct * ct)
    x, y, z = r * st * np.cos(phi), r * st * np.sin(phi), r * ct
    h = 0.5 * (150.0 * r / n) ** (1.0 / 3.0)
    zero = np.zeros(n)
    lo = min(x.min(), y.min(), z.min())
    hi = max(x.max(), y.max(), z.max())
    return dict(x=x, y=y, z=z, h=h, m=np.full(n, 1.0 / n), vx=zero, vy=zero.copy(), vz=zero.copy(),
                u=np.full(n, 0.05)), ((x.min(), y.min(), z.min()), (x.max(), y.max(), z.max()), False)
