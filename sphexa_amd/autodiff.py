"""Differentiable SPH density and momentum-energy (torch.autograd.Function wrappers).

The reference's loops are forward-only; these operators add hand-written adjoints
(sphx_density_adjoint, sphx_momentum_energy_adjoint) behind torch.autograd so that
losses built from SPH fields can be differentiated with respect to particle positions.

Both operators take particles in any order.  Each call sorts them along the Hilbert
curve, builds the cornerstone tree, the spatial target groups and the neighbor lists
at FIXED smoothing lengths (no h iteration: ng0 = 0, and a particle with more than
ngmax neighbors is an error rather than a reason to change h), evaluates the loop,
and returns the outputs in the caller's order.  Backward recomputes the lists from the
saved inputs and runs the adjoint kernel.

* ``DensityOp``: rho(x, y, z, h, m) (hydro_std/density.hpp:41-60); gradients for x, y,
  z and h (m is not differentiated).
* ``MomentumEnergyOp``: (ax, ay, az, du)(x, y, z; v, h, m, rho, p, c, c11..c33)
  (hydro_std/momentum_energy_kern.hpp:12-134); gradients for x, y, z only, with every
  other input held fixed (they must not require grad).
"""
from __future__ import annotations

import torch

from . import hydro
from ._lib import FP64, MIXED

_ME_FIELDS = ("vx", "vy", "vz", "h", "m", "rho", "p", "c", "c11", "c12", "c13", "c22", "c23", "c33")


class _Pipeline:
    """Sort, tree, groups and fixed-h neighbor lists for one set of positions."""

    def __init__(self, ctx: hydro.Context, box, precision=FP64, ngmax=256, bucket=64, sinc_index=6.0):
        self.ctx, self.box, self.precision, self.ngmax, self.bucket = ctx, box, precision, ngmax, bucket
        self.sinc_index = sinc_index
        self.d = None
        self.order = None
        self.groups = None

    @property
    def dtype(self):
        return torch.float64 if self.precision == FP64 else torch.float32

    def _data(self, n):
        if self.d is None or self.d.n != n:
            self.d = hydro.ParticlesData(n, self.precision, 0, self.ngmax, sincIndex=self.sinc_index)
            self.d.allocate(self.ctx.device)
            self.keys = torch.empty(n, dtype=torch.int64, device=self.ctx.device)
            self.order = torch.empty(n, dtype=torch.int32, device=self.ctx.device)
        return self.d

    def prepare(self, x, y, z, fields: dict):
        """Sort by Hilbert key, gather x, y, z and `fields` into self.d, build the tree,
        groups and lists.  Returns self.d."""
        ctx, box = self.ctx, self.box
        n = x.numel()
        for name, t in (("y", y), ("z", z)) + tuple(fields.items()):
            if t.numel() != n:
                raise ValueError(f"{name}: {t.numel()} elements, x has {n}")
        # FP64 contiguous copies of the positions, used for the keys and the gather alike
        x, y, z = (t.detach().to(torch.float64).contiguous() for t in (x, y, z))
        d = self._data(n)
        self.keys.zero_()
        hydro.compute_sfc_keys(ctx, x, y, z, box, self.keys)
        ctx.check(ctx.lib.sphx_sequence(ctx.h, hydro._ptr(self.order), n, 0))
        hydro.sort_by_key(ctx, self.keys, self.order)
        for name, src in (("x", x), ("y", y), ("z", z)):
            hydro.gather(ctx, self.order, src, getattr(d, name))
        for name, src in fields.items():
            hydro.gather(ctx, self.order, src.detach().to(self.dtype).contiguous(), getattr(d, name))
        leaves, counts, layout = hydro.compute_octree(ctx, self.keys, self.bucket)
        o = hydro.build_octree(ctx, leaves)
        o.layout, o.counts = layout, counts
        hydro.node_fp_centers(ctx, o, box)
        self.groups = hydro.compute_groups(ctx, d.x, d.y, d.z, d.h, o, box)
        h0 = d.h.clone()
        hydro.find_neighbors_sfc(ctx, d, box, o, self.groups)
        if not torch.equal(h0, d.h):
            over = int((d.nc.to(torch.int64) - 1).max().item())
            raise ValueError(f"more than ngmax = {self.ngmax} neighbors at fixed h (up to {over} after the "
                             f"smoothing-length update); raise ngmax")
        self.octree = o
        return d

    def to_sorted(self, t):
        return t.detach().to(self.dtype).contiguous()[self.order.long()]

    def unsort(self, t_sorted, dtype=None):
        out = torch.empty_like(t_sorted, dtype=dtype or t_sorted.dtype)
        out.index_copy_(0, self.order.long(), t_sorted.to(out.dtype))
        return out


class _DensityFn(torch.autograd.Function):
    @staticmethod
    def forward(fctx, x, y, z, h, m, op):
        pipe = op.pipe
        d = pipe.prepare(x, y, z, {"h": h, "m": m})
        hydro.compute_density(pipe.ctx, d, pipe.box, pipe.groups)
        fctx.op = op
        fctx.save_for_backward(x, y, z, h, m)
        return pipe.unsort(d.rho, dtype=h.dtype)

    @staticmethod
    def backward(fctx, g_rho):
        op = fctx.op
        pipe = op.pipe
        x, y, z, h, m = fctx.saved_tensors
        d = pipe.prepare(x, y, z, {"h": h, "m": m})
        gx, gy, gz, gh = hydro.density_adjoint(pipe.ctx, d, pipe.box, pipe.to_sorted(g_rho), pipe.groups)
        return (pipe.unsort(gx, x.dtype), pipe.unsort(gy, y.dtype), pipe.unsort(gz, z.dtype), pipe.unsort(gh, h.dtype),
                None, None)


class DensityOp:
    """rho = DensityOp(ctx, box)(x, y, z, h, m), differentiable in x, y, z and h."""

    def __init__(self, ctx: hydro.Context, box, precision=FP64, ngmax=256, bucket=64, sinc_index=6.0):
        self.pipe = _Pipeline(ctx, box, precision, ngmax, bucket, sinc_index)

    @property
    def K(self):
        return self.pipe.d.K if self.pipe.d is not None else None

    def __call__(self, x, y, z, h, m):
        return _DensityFn.apply(x, y, z, h, m, self)


class _MomentumEnergyFn(torch.autograd.Function):
    @staticmethod
    def forward(fctx, x, y, z, op, *fields):
        pipe = op.pipe
        f = dict(zip(_ME_FIELDS, fields))
        d = pipe.prepare(x, y, z, f)
        hydro.compute_momentum_energy_std(pipe.ctx, d, pipe.box, pipe.groups)
        fctx.op = op
        fctx.save_for_backward(x, y, z, *fields)
        dt = f["h"].dtype
        return (pipe.unsort(d.ax, dt), pipe.unsort(d.ay, dt), pipe.unsort(d.az, dt), pipe.unsort(d.du, dt))

    @staticmethod
    def backward(fctx, g_ax, g_ay, g_az, g_du):
        op = fctx.op
        pipe = op.pipe
        x, y, z, *fields = fctx.saved_tensors
        d = pipe.prepare(x, y, z, dict(zip(_ME_FIELDS, fields)))
        srt = lambda g: None if g is None else pipe.to_sorted(g)
        gx, gy, gz = hydro.momentum_energy_adjoint(pipe.ctx, d, pipe.box, srt(g_ax), srt(g_ay), srt(g_az), srt(g_du),
                                                   pipe.groups)
        return (pipe.unsort(gx, x.dtype), pipe.unsort(gy, y.dtype), pipe.unsort(gz, z.dtype), None) + \
            (None,) * len(fields)


class MomentumEnergyOp:
    """(ax, ay, az, du) = MomentumEnergyOp(ctx, box)(x, y, z, vx=..., vy=..., ..., c33=...),
    differentiable in x, y, z; the other fields are constants of the operator."""

    def __init__(self, ctx: hydro.Context, box, precision=FP64, ngmax=256, bucket=64, sinc_index=6.0):
        self.pipe = _Pipeline(ctx, box, precision, ngmax, bucket, sinc_index)

    def __call__(self, x, y, z, **fields):
        missing = [k for k in _ME_FIELDS if k not in fields]
        if missing:
            raise TypeError(f"missing fields {missing}")
        if any(fields[k].requires_grad for k in _ME_FIELDS):
            raise ValueError("only x, y, z are differentiated; pass the other fields detached")
        return _MomentumEnergyFn.apply(x, y, z, self, *[fields[k] for k in _ME_FIELDS])


def relax_glass(ctx: hydro.Context, box, x, y, z, h, m, iterations=200, step=None, precision=FP64, ngmax=256,
                callback=None):
    """Glass relaxation by gradient descent on L = sum_i (rho_i - mean rho)^2 with the density
    adjoint: x <- x - step * dL/dx (positions wrapped into a periodic box).  The default
    step moves the particle with the largest gradient by 0.05 of its smoothing length in
    the first iteration.  Returns (x, y, z, history of L)."""
    op = DensityOp(ctx, box, precision, ngmax)
    x, y, z = (t.detach().clone() for t in (x, y, z))
    lo = torch.tensor([box.xmin, box.ymin, box.zmin], dtype=torch.float64)
    hi = torch.tensor([box.xmax, box.ymax, box.zmax], dtype=torch.float64)
    periodic = (box.bx == 1, box.by == 1, box.bz == 1)
    hist = []
    for it in range(iterations):
        xs = [t.requires_grad_(True) for t in (x, y, z)]
        rho = op(*xs, h, m)
        loss = ((rho.double() - rho.double().mean()) ** 2).sum()
        gx, gy, gz = torch.autograd.grad(loss, xs)
        hist.append(float(loss.item()))
        if step is None:
            gmax = torch.sqrt(gx * gx + gy * gy + gz * gz).max().item()
            step = 0.05 * float(h.double().mean().item()) / max(gmax, 1e-300)
        with torch.no_grad():
            out = []
            for k, (t, g) in enumerate(((x, gx), (y, gy), (z, gz))):
                t = t.detach() - step * g
                if periodic[k]:
                    L = float(hi[k] - lo[k])
                    t = lo[k].item() + torch.remainder(t - lo[k].item(), L)
                out.append(t)
            x, y, z = out
        if callback is not None:
            callback(it, hist[-1])
    return x, y, z, hist
