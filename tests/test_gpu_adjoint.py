"""GPU adjoints of density and momentum-energy against the FP64 torch reference
(tests/adjoint_ref.py, itself pinned to the oracle by value and by central differences).

Inputs are the GPU forward's own sorted fields and exported lists, so the comparison
isolates the adjoint kernels.  FP64 tolerance 1e-11 of the gradient's max (atomic
summation order); MIXED 2e-3 (FP32 pair arithmetic and the polynomial kernel's exact
derivative against the table's segment slopes)."""
import numpy as np
import pytest
import torch

from tests import adjoint_ref as R
from tests.helpers import uniform_box

pytestmark = pytest.mark.gpu

ME_F = ("vx", "vy", "vz", "h", "m", "rho", "p", "c", "c11", "c12", "c13", "c22", "c23", "c33")


@pytest.fixture(scope="module")
def ctx():
    from sphexa_amd import hydro

    c = hydro.Context(0)
    yield c
    c.close()


def _forward(ctx, n, seed, precision, periodic=True):
    from sphexa_amd import hydro

    inp, box = uniform_box(n, seed=seed, periodic=periodic)
    hb = hydro.make_box(box.xmin, box.xmax, box.bx)
    step = hydro.HydroStep(ctx, n, hb, precision)
    dt = step.d.dtype
    d = step.compute_forces({k: torch.tensor(v, dtype=torch.float64 if k in ("x", "y", "z") else dt, device=ctx.device)
                             for k, v in inp.items()})
    torch.cuda.synchronize()
    nb = hydro.export_neighbors(ctx, d).cpu().numpy().view(np.uint32)
    nc = d.nc.cpu().numpy().view(np.uint32)
    i, j = R.pairs_from_lists(nb, nc, d.ngmax)
    wh = torch.from_numpy(hydro.kernel_tables(6.0)[0])
    return step, d, hb, box, i, j, wh


def _err(got, ref):
    return float((got.cpu().double() - ref).abs().max() / ref.abs().max())


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("periodic", [True, False])
def test_density_adjoint(ctx, mode, periodic):
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64, MIXED

    print("case density_adjoint", mode, periodic, flush=True)
    step, d, hb, box, i, j, wh = _forward(ctx, 12000, 81, FP64 if mode == "fp64" else MIXED, periodic)
    g = torch.from_numpy(np.random.default_rng(82).normal(size=d.n))
    gx, gy, gz, gh = hydro.density_adjoint(ctx, d, hb, g.to(ctx.device), step.groups)
    xyz = [d.x.cpu().double().requires_grad_(True), d.y.cpu().double().requires_grad_(True),
           d.z.cpu().double().requires_grad_(True)]
    h = d.h.cpu().double().requires_grad_(True)
    rho = R.density(*xyz, h, d.m.cpu().double(), i, j, wh, d.K, box)
    ref = torch.autograd.grad((rho * g).sum(), xyz + [h])
    tol = 1e-11 if mode == "fp64" else 2e-3
    errs = [_err(a, b) for a, b in zip((gx, gy, gz, gh), ref)]
    print("errors", errs, flush=True)
    assert max(errs) < tol


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_momentum_energy_adjoint(ctx, mode):
    from sphexa_amd._lib import FP64, MIXED

    print("case me_adjoint", mode, flush=True)
    step, d, hb, box, i, j, wh = _forward(ctx, 12000, 83, FP64 if mode == "fp64" else MIXED)
    W = torch.from_numpy(np.random.default_rng(84).normal(size=(4, d.n)))
    gx, gy, gz = step.momentum_energy_vjp(*[W[k].to(ctx.device) for k in range(4)])
    f = {k: getattr(d, k).cpu().double() for k in ME_F}
    xyz = [d.x.cpu().double().requires_grad_(True), d.y.cpu().double().requires_grad_(True),
           d.z.cpu().double().requires_grad_(True)]
    out = R.momentum_energy(*xyz, f, i, j, wh, d.K, box)
    ref = torch.autograd.grad(sum((W[k] * out[k]).sum() for k in range(4)), xyz)
    tol = 1e-11 if mode == "fp64" else 2e-3
    errs = [_err(a, b) for a, b in zip((gx, gy, gz), ref)]
    print("errors", errs, flush=True)
    assert max(errs) < tol
    # one output gradient alone (the others NULL) is the matching part of the sum
    gx1, _, _ = step.momentum_energy_vjp(None, None, None, W[3].to(ctx.device), recompute_lists=False)
    ref1 = torch.autograd.grad((W[3] * R.momentum_energy(*xyz, f, i, j, wh, d.K, box)[3]).sum(), xyz)
    assert _err(gx1, ref1[0]) < tol


def test_density_op_autograd_unsorted_inputs(ctx):
    """DensityOp: caller-order inputs, autograd gradients equal the sorted adjoint mapped back."""
    from sphexa_amd import autodiff, hydro
    from sphexa_amd._lib import FP64

    print("case density_op", flush=True)
    inp, box = uniform_box(8000, seed=85)
    hb = hydro.make_box(box.xmin, box.xmax, box.bx)
    op = autodiff.DensityOp(ctx, hb, FP64, ngmax=256)
    t = {k: torch.tensor(inp[k], device=ctx.device) for k in ("x", "y", "z", "h", "m")}
    xyz = [t[k].clone().requires_grad_(True) for k in "xyz"]
    h = t["h"].clone().requires_grad_(True)
    rho = op(*xyz, h, t["m"])
    loss = ((rho - rho.mean()) ** 2).sum()
    grads = torch.autograd.grad(loss, xyz + [h])
    # reference in the caller's order: pairs from the oracle-free exported lists of the op
    d, order = op.pipe.d, op.pipe.order.long().cpu()
    nb = hydro.export_neighbors(ctx, d).cpu().numpy().view(np.uint32)
    i, j = R.pairs_from_lists(nb, d.nc.cpu().numpy().view(np.uint32), d.ngmax)
    wh = torch.from_numpy(hydro.kernel_tables(6.0)[0])
    xs = [torch.tensor(inp[k])[order].requires_grad_(True) for k in "xyz"]
    hs = torch.tensor(inp["h"])[order].requires_grad_(True)
    rr = R.density(*xs, hs, torch.tensor(inp["m"])[order], i, j, wh, d.K, box)
    assert float((rr - rho.detach().cpu()[order]).abs().max() / rr.abs().max()) < 1e-13
    ref = torch.autograd.grad(((rr - rr.mean()) ** 2).sum(), xs + [hs])
    for g, r in zip(grads, ref):
        assert _err(g.cpu()[order], r) < 1e-11


def test_glass_relaxation_reduces_density_variance(ctx):
    from sphexa_amd import autodiff, hydro
    from sphexa_amd._lib import FP64
    from sphexa_amd.workloads import jittered_lattice

    print("case relax_glass", flush=True)
    side = 16
    f = jittered_lattice(side, 0.3, seed=86)[0]
    x, y, z = f["x"], f["y"], f["z"]
    n = side ** 3
    hb = hydro.make_box(0.0, 1.0, 1)
    dev = ctx.device
    h = torch.full((n,), 0.5 * (300 / (4 * np.pi * n)) ** (1 / 3), dtype=torch.float64, device=dev)
    m = torch.full((n,), 1.0 / n, dtype=torch.float64, device=dev)
    X = [torch.tensor(a, device=dev) for a in (x, y, z)]
    _, _, _, hist = autodiff.relax_glass(ctx, hb, *X, h, m, iterations=20, precision=FP64)
    print("loss", hist[0], hist[-1], flush=True)
    assert hist[-1] < 0.5 * hist[0]


def _lattice20(ctx, seed=87):
    """The differentiable-density case at 20^3: periodic unit lattice, Gaussian jitter
    sigma = 0.3 spacing, m = 1/N, fixed h = 0.5 (300 / (4 pi N))^(1/3)."""
    from sphexa_amd.workloads import jittered_lattice

    side = 20
    n = side ** 3
    f = jittered_lattice(side, 0.3, seed=seed)[0]
    h = np.full(n, 0.5 * (300 / (4 * np.pi * n)) ** (1 / 3))
    m = np.full(n, 1.0 / n)
    return f["x"], f["y"], f["z"], h, m


def test_density_adjoint_matches_finite_differences(ctx):
    """GPU adjoint of L = sum (rho - mean rho)^2 against central differences of the oracle's
    density (fixed lists) on 64 random particles: <= 1e-6 of the gradient scale."""
    import oracle
    from sphexa_amd import autodiff, hydro
    from sphexa_amd._lib import FP64

    print("case density_adjoint_fd", flush=True)
    x, y, z, h, m = _lattice20(ctx)
    hb = hydro.make_box(0.0, 1.0, 1)
    op = autodiff.DensityOp(ctx, hb, FP64, ngmax=256)
    xyz = [torch.tensor(a, device=ctx.device).requires_grad_(True) for a in (x, y, z)]
    rho = op(*xyz, torch.tensor(h, device=ctx.device), torch.tensor(m, device=ctx.device))
    grads = [g.cpu().numpy() for g in torch.autograd.grad(((rho - rho.mean()) ** 2).sum(), xyz)]
    scale = max(np.abs(g).max() for g in grads)
    # oracle density over the op's own lists, in the op's sorted order
    d, order = op.pipe.d, op.pipe.order.long().cpu().numpy()
    nb = hydro.export_neighbors(ctx, d).cpu().numpy().view(np.uint32)
    nc = d.nc.cpu().numpy().view(np.uint32)
    cpu = oracle.Cpu("oracle")
    wh, _, K = cpu.kernel_tables()
    box = oracle.Box.cube(0.0, 1.0, oracle.BC_PERIODIC)
    s = {k: np.ascontiguousarray(v[order]) for k, v in dict(x=x, y=y, z=z, h=h, m=m).items()}

    def loss(f):
        r = cpu.density(nb, nc, d.ngmax, f["x"], f["y"], f["z"], s["h"], s["m"], wh, K, box)
        return float(np.sum((r - r.mean()) ** 2))

    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    worst = 0.0
    delta = 2e-9
    for p in np.random.default_rng(88).choice(len(x), 64, replace=False):
        for a, key in enumerate("xyz"):
            plus, minus = dict(s), dict(s)
            plus[key], minus[key] = s[key].copy(), s[key].copy()
            plus[key][inv[p]] += delta
            minus[key][inv[p]] -= delta
            fd = (loss(plus) - loss(minus)) / (2 * delta)
            worst = max(worst, abs(fd - grads[a][p]) / scale)
    print("worst", worst, flush=True)
    assert worst < 1e-6


def test_momentum_energy_adjoint_matches_finite_differences(ctx):
    """GPU FP64 adjoint of L = sum_i (w1 ax + w2 ay + w3 az + w4 du) against central differences
    of the oracle's momentum-energy loop over the GPU's own lists and fields (v, h, m, rho, p,
    c and the IAD tensors held fixed) on 24 random particles: <= 1e-6 of the gradient scale."""
    import oracle
    from sphexa_amd._lib import FP64

    print("case me_adjoint_fd", flush=True)
    step, d, hb, box, i, j, wh = _forward(ctx, 6000, 91, FP64)
    n = d.n
    Wt = np.random.default_rng(92).normal(size=(4, n))
    grads = [g.cpu().numpy() for g in step.momentum_energy_vjp(*[torch.from_numpy(Wt[k]).to(ctx.device)
                                                                 for k in range(4)])]
    scale = max(np.abs(g).max() for g in grads)
    nb = hydro_export(ctx, d)
    nc = d.nc.cpu().numpy().view(np.uint32)
    f = {k: getattr(d, k).cpu().double().numpy() for k in ("x", "y", "z") + ME_F}
    cs = [f[k] for k in ("c11", "c12", "c13", "c22", "c23", "c33")]
    cpu = oracle.Cpu("oracle")
    owh, _, K = cpu.kernel_tables()

    def loss(g):
        out = cpu.momentum_energy(nb, nc, d.ngmax, g["x"], g["y"], g["z"], f["vx"], f["vy"], f["vz"], f["h"], f["m"],
                                  f["rho"], f["p"], f["c"], cs, owh, K, 0.2, box)
        return float(sum(np.dot(Wt[k], out[k]) for k in range(4)))

    worst, delta = 0.0, 2e-9
    for p in np.random.default_rng(93).choice(n, 24, replace=False):
        for a, key in enumerate("xyz"):
            plus, minus = dict(f), dict(f)
            plus[key], minus[key] = f[key].copy(), f[key].copy()
            plus[key][p] += delta
            minus[key][p] -= delta
            fd = (loss(plus) - loss(minus)) / (2 * delta)
            worst = max(worst, abs(fd - grads[a][p]) / scale)
    print("worst", worst, flush=True)
    assert worst < 1e-6


def hydro_export(ctx, d):
    from sphexa_amd import hydro

    return hydro.export_neighbors(ctx, d).cpu().numpy().view(np.uint32)


def test_density_adjoint_vs_bruteforce_autograd(ctx):
    """GPU adjoint against FP64 torch autograd of an all-pairs density (no neighbor lists:
    every pair, minimum image, the same table-linear W, which is 0 beyond 2h) at 20^3."""
    from sphexa_amd import autodiff, hydro
    from sphexa_amd._lib import FP64

    print("case density_adjoint_bruteforce", flush=True)
    x, y, z, h, m = _lattice20(ctx, seed=89)
    n = len(x)
    hb = hydro.make_box(0.0, 1.0, 1)
    op = autodiff.DensityOp(ctx, hb, FP64, ngmax=256)
    dev = ctx.device
    xyz = [torch.tensor(a, device=dev).requires_grad_(True) for a in (x, y, z)]
    ht = torch.tensor(h, device=dev).requires_grad_(True)
    mt = torch.tensor(m, device=dev)
    w = torch.from_numpy(np.random.default_rng(90).normal(size=n)).to(dev)
    rho = op(*xyz, ht, mt)
    got = torch.autograd.grad((rho * w).sum(), xyz + [ht])
    # brute force, rows in chunks
    wh = torch.from_numpy(hydro.kernel_tables(6.0)[0]).to(dev)
    K = op.K
    bx = [torch.tensor(a, device=dev).requires_grad_(True) for a in (x, y, z)]
    bh = torch.tensor(h, device=dev).requires_grad_(True)
    parts = []
    for r0 in range(0, n, 1000):
        rows = slice(r0, min(n, r0 + 1000))
        d2 = 0
        for c in bx:
            dd = c[rows, None] - c[None, :]
            dd = dd - torch.round(dd)  # minimum image in the unit box
            d2 = d2 + dd * dd
        idx = torch.arange(r0, min(n, r0 + 1000), device=dev)
        self_mask = torch.zeros_like(d2, dtype=torch.bool)
        self_mask[torch.arange(len(idx), device=dev), idx] = True
        d2 = torch.where(self_mask, torch.ones_like(d2), d2)  # keep sqrt differentiable on the diagonal
        v = torch.sqrt(d2) / bh[rows, None]
        Wv = R.table_w(wh, v.reshape(-1)).reshape(v.shape)
        Wv = torch.where(self_mask, torch.zeros_like(Wv), Wv)
        rho0 = mt[rows] + (Wv * mt[None, :]).sum(1)
        parts.append(K / bh[rows] ** 3 * rho0)
    rb = torch.cat(parts)
    assert float((rb - rho.detach()).abs().max() / rb.abs().max()) < 1e-13
    ref = torch.autograd.grad((rb * w).sum(), bx + [bh])
    errs = [_err(a, b.cpu().double()) for a, b in zip(got, ref)]
    print("errors", errs, flush=True)
    assert max(errs) < 1e-11


def test_density_adjoint_coincident_particles(ctx):
    """Two particles at the same position are neighbors (the search excludes only i itself):
    their pair contributes W(0) to the density and nothing to the gradient, so the adjoint
    stays finite and matches central differences of the oracle's density."""
    import oracle
    from sphexa_amd import autodiff, hydro
    from sphexa_amd._lib import FP64

    print("case density_adjoint_coincident", flush=True)
    inp, box = uniform_box(3000, seed=94)
    for k in ("x", "y", "z"):
        inp[k][5] = inp[k][17]
    hb = hydro.make_box(box.xmin, box.xmax, box.bx)
    op = autodiff.DensityOp(ctx, hb, FP64, ngmax=256)
    xyz = [torch.tensor(inp[k], device=ctx.device).requires_grad_(True) for k in "xyz"]
    rho = op(*xyz, torch.tensor(inp["h"], device=ctx.device), torch.tensor(inp["m"], device=ctx.device))
    grads = [g.cpu().numpy() for g in torch.autograd.grad(((rho - rho.mean()) ** 2).sum(), xyz)]
    assert all(np.all(np.isfinite(g)) for g in grads)
    scale = max(np.abs(g).max() for g in grads)
    d, order = op.pipe.d, op.pipe.order.long().cpu().numpy()
    nb = hydro.export_neighbors(ctx, d).cpu().numpy().view(np.uint32)
    nc = d.nc.cpu().numpy().view(np.uint32)
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    assert inv[17] in nb[inv[5], :nc[inv[5]] - 1]  # the two are each other's neighbors
    cpu = oracle.Cpu("oracle")
    wh, _, K = cpu.kernel_tables()
    s = {k: np.ascontiguousarray(inp[k][order]) for k in ("x", "y", "z", "h", "m")}

    def loss(f):
        r = cpu.density(nb, nc, d.ngmax, f["x"], f["y"], f["z"], s["h"], s["m"], wh, K, box)
        return float(np.sum((r - r.mean()) ** 2))

    worst, delta = 0.0, 2e-9
    for p in [5, 17] + list(np.random.default_rng(95).choice(3000, 6, replace=False)):
        for a, key in enumerate("xyz"):
            plus, minus = dict(s), dict(s)
            plus[key], minus[key] = s[key].copy(), s[key].copy()
            plus[key][inv[p]] += delta
            minus[key][inv[p]] -= delta
            fd = (loss(plus) - loss(minus)) / (2 * delta)
            worst = max(worst, abs(fd - grads[a][p]) / scale)
    print("worst", worst, flush=True)
    assert worst < 1e-6


def test_momentum_energy_op_autograd_unsorted_inputs(ctx):
    """MomentumEnergyOp on caller-order inputs: the forward equals the hydro step's outputs
    mapped back through the sort order, torch.autograd.grad equals the step's momentum-energy
    VJP mapped back (to 1e-11, atomic summation order), and both match central differences of
    the oracle's momentum-energy loop on the op's own lists (<= 1e-6 of the gradient scale).

    The table-linear kernel has a kink at every table node (and the viscosity one at w = 0),
    where the derivative is undefined; a central difference whose step straddles one mixes two
    slopes.  Each sample is therefore differenced with two steps, 2e-9 and 5e-10 (outputs
    differenced element-wise, so round-off stays far below 1e-8): where they disagree by more
    than 1e-8 of the scale a kink lies within the step and the sample is skipped; at least 24
    of the 36 samples must be clean."""
    import oracle
    from sphexa_amd import autodiff, hydro
    from sphexa_amd._lib import FP64

    print("case momentum_energy_op", flush=True)
    n = 5000
    inp, box = uniform_box(n, seed=96)
    hb = hydro.make_box(box.xmin, box.xmax, box.bx)
    step = hydro.HydroStep(ctx, n, hb, FP64, ngmax=256)
    d = step.compute_forces({k: torch.tensor(v, device=ctx.device) for k, v in inp.items()})
    torch.cuda.synchronize()
    order = step.order.long()
    # caller-order copies of every input of the loop (h after the step's iteration)
    fields = {}
    for k in ME_F:
        t = torch.empty_like(getattr(d, k))
        t.index_copy_(0, order, getattr(d, k))
        fields[k] = t
    xyz = [torch.tensor(inp[k], device=ctx.device).requires_grad_(True) for k in "xyz"]
    op = autodiff.MomentumEnergyOp(ctx, hb, FP64, ngmax=256)
    out = op(*xyz, **fields)
    for k, o in zip(("ax", "ay", "az", "du"), out):
        want = getattr(d, k)
        got = o.detach()[order]
        assert float((got - want).abs().max() / want.abs().max()) < 1e-12, k
    Wt = torch.from_numpy(np.random.default_rng(97).normal(size=(4, n))).to(ctx.device)
    loss = sum((Wt[k] * out[k]).sum() for k in range(4))
    grads = torch.autograd.grad(loss, xyz)
    vjp = step.momentum_energy_vjp(*[Wt[k][order] for k in range(4)])
    for g, v in zip(grads, vjp):
        assert _err(g[order], v.cpu().double()) < 1e-11
    # central differences of the oracle over the op's lists and sorted fields
    pd, porder = op.pipe.d, op.pipe.order.long().cpu().numpy()
    nb = hydro_export(ctx, pd)
    nc = pd.nc.cpu().numpy().view(np.uint32)
    f = {k: getattr(pd, k).cpu().double().numpy() for k in ("x", "y", "z") + ME_F}
    cs = [f[k] for k in ("c11", "c12", "c13", "c22", "c23", "c33")]
    W = Wt.cpu().numpy()[:, porder]
    cpu = oracle.Cpu("oracle")
    owh, _, K = cpu.kernel_tables()

    def outputs(g):
        return cpu.momentum_energy(nb, nc, pd.ngmax, g["x"], g["y"], g["z"], f["vx"], f["vy"], f["vz"], f["h"],
                                   f["m"], f["rho"], f["p"], f["c"], cs, owh, K, 0.2, box)[:4]

    gnp = [g.cpu().numpy() for g in grads]
    scale = max(np.abs(g).max() for g in gnp)
    inv = np.empty_like(porder)
    inv[porder] = np.arange(n)

    def central(key, q, delta):
        # outputs differenced element-wise before the weighted sum: the particles the move does
        # not reach contribute exactly zero, so no large total loss cancels
        plus, minus = dict(f), dict(f)
        plus[key], minus[key] = f[key].copy(), f[key].copy()
        plus[key][q] += delta
        minus[key][q] -= delta
        op_, om = outputs(plus), outputs(minus)
        return sum(float(np.dot(W[k], op_[k] - om[k])) for k in range(4)) / (2 * delta)

    worst, clean = 0.0, 0
    for p in np.random.default_rng(98).choice(n, 12, replace=False):
        for a, key in enumerate("xyz"):
            fd, fd_small = central(key, inv[p], 2e-9), central(key, inv[p], 5e-10)
            if abs(fd - fd_small) > 1e-8 * scale:
                continue  # a kink within the step
            clean += 1
            worst = max(worst, abs(fd - gnp[a][p]) / scale)
    print("worst", worst, "clean samples", clean, flush=True)
    assert clean >= 24
    assert worst < 1e-6


def test_density_adjoint_1m_lattice_finite_differences_and_relaxation(ctx):
    """The differentiable-density configuration at its full size: 1M particles on a 100^3
    periodic lattice, Gaussian jitter 0.3 spacing, fixed h = 0.5 (300 / (4 pi N))^(1/3).

    * The GPU adjoint of L = sum (rho - mean rho)^2 against central differences of the oracle's
      density on 64 random particles, <= 1e-6 of the gradient scale.  Moving particle p changes
      only rho_i for i in {p} and its neighbors (h is uniform, so the lists are symmetric), and
      dL/dx_p = sum_i 2 (rho_i - mean) d rho_i/dx_p exactly (the mean's derivative multiplies
      sum_i (rho_i - mean) = 0), so each difference needs only those densities, computed by the
      oracle's one-particle loop over the GPU's lists.  Steps 2e-9 and 5e-10; a sample where
      they disagree by more than 1e-8 of the scale straddles a kink of the table-linear kernel
      and is skipped (at least 150 of the 192 samples must be clean).
    * 20 iterations of the glass relaxation lower the loss steadily, and by at least 30 %."""
    import oracle
    from sphexa_amd import autodiff, hydro
    from sphexa_amd._lib import FP64
    from sphexa_amd.workloads import jittered_lattice

    print("case density adjoint 1M lattice", flush=True)
    side = 100
    n = side ** 3
    f = jittered_lattice(side, 0.3, seed=99)[0]
    h0 = 0.5 * (300 / (4 * np.pi * n)) ** (1 / 3)
    dev = ctx.device
    hb = hydro.make_box(0.0, 1.0, 1)
    op = autodiff.DensityOp(ctx, hb, FP64, ngmax=256)
    xyz = [torch.tensor(f[k], device=dev).requires_grad_(True) for k in ("x", "y", "z")]
    ht = torch.full((n,), h0, dtype=torch.float64, device=dev)
    mt = torch.full((n,), 1.0 / n, dtype=torch.float64, device=dev)
    rho = op(*xyz, ht, mt)
    grads = [g.cpu().numpy() for g in torch.autograd.grad(((rho - rho.mean()) ** 2).sum(), xyz)]
    scale = max(np.abs(g).max() for g in grads)
    d, order = op.pipe.d, op.pipe.order.long().cpu().numpy()
    inv = np.empty_like(order)
    inv[order] = np.arange(n)
    rho_s = d.rho.cpu().numpy()
    mean = float(rho_s.mean())
    sf = {k: getattr(d, k).cpu().numpy() for k in ("x", "y", "z", "h", "m")}
    nc = d.nc.cpu().numpy().view(np.uint32)
    full = hydro.export_neighbors(ctx, d)
    cpu = oracle.Cpu("oracle")
    wh, _, K = cpu.kernel_tables()
    obox = oracle.Box.cube(0.0, 1.0, oracle.BC_PERIODIC)

    rows = {}

    def fetch(ids):
        """Neighbor rows of the particles ids (sorted order), one device gather."""
        ids = np.array(sorted(set(int(i) for i in ids) - set(rows)), np.int64)
        if len(ids):
            got = full[torch.as_tensor(ids, device=full.device)].cpu().numpy().view(np.uint32)
            for i, r in zip(ids, got):
                rows[int(i)] = r[: int(nc[i]) - 1].astype(np.int64)

    def row(i):
        return rows[int(i)]

    sample = np.random.default_rng(100).choice(n, 64, replace=False)
    fetch(inv[sample])
    fetch(np.concatenate([row(inv[p]) for p in sample]))

    def density_of(i, moved, key, delta):
        """rho_i from the oracle's one-particle loop over i's list, particle `moved` shifted."""
        ids = np.concatenate([[i], row(i)])
        loc = {k: np.ascontiguousarray(sf[k][ids]) for k in sf}
        loc[key][ids == moved] += delta
        lnb = np.arange(1, len(ids), dtype=np.uint32)
        xm = cpu.xmass_one(0, lnb, loc["x"], loc["y"], loc["z"], loc["h"], loc["m"], wh, K, obox)
        return loc["m"][0] / xm

    worst, clean = 0.0, 0
    for p in sample:
        s = inv[p]
        affected = np.concatenate([[s], row(s)])
        for a, key in enumerate(("x", "y", "z")):
            fds = []
            for delta in (2e-9, 5e-10):
                g = 0.0
                for i in affected:
                    drho = (density_of(i, s, key, delta) - density_of(i, s, key, -delta)) / (2 * delta)
                    g += 2.0 * (rho_s[i] - mean) * drho
                fds.append(g)
            if abs(fds[0] - fds[1]) > 1e-8 * scale:
                continue
            clean += 1
            worst = max(worst, abs(fds[0] - grads[a][p]) / scale)
    print("worst", worst, "clean samples", clean, flush=True)
    del full
    torch.cuda.empty_cache()
    assert clean >= 150
    assert worst < 1e-6

    x0 = [torch.tensor(f[k], device=dev) for k in ("x", "y", "z")]
    _, _, _, hist = autodiff.relax_glass(ctx, hb, *x0, ht, mt, iterations=20, precision=FP64)
    print("relaxation loss", hist[0], hist[-1], flush=True)
    assert all(np.isfinite(hist))
    assert all(b < a for a, b in zip(hist, hist[1:]))
    assert hist[-1] < 0.7 * hist[0]
