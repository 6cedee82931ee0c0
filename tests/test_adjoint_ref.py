"""CPU checks of the torch FP64 reference used for the GPU adjoints (tests/adjoint_ref.py):
its forward values equal the oracle's density and momentum-energy, and its autograd
gradients equal central differences of the ORACLE's loops (fixed neighbor lists).  A
wrong sign, a dropped term or a transposed index in the reference fails one of them."""
import numpy as np
import pytest
import torch

import oracle
from tests import adjoint_ref as R
from tests.helpers import uniform_box

CPU = oracle.Cpu("oracle")


def _setup(n=3000, seed=71, periodic=True, ng=60):
    inp, box = uniform_box(n, seed=seed, ng=ng, periodic=periodic)
    keys = CPU.sfc_keys(inp["x"], inp["y"], inp["z"], box)
    keys, order = CPU.sort_by_key(keys, np.arange(n, dtype=np.uint32))
    s = {k: np.ascontiguousarray(v[order]) for k, v in inp.items()}
    t = CPU.tree(keys, box)
    ngmax = 200
    nb, nc = CPU.find_neighbors(s["x"], s["y"], s["z"], s["h"], box, t, ngmax)
    assert nc.max() <= ngmax
    nc = nc + 1  # with self, as findNeighborsSph stores it
    wh, _, K = CPU.kernel_tables()
    return s, box, nb, nc, ngmax, wh, K


def _oracle_fields(s, box, nb, nc, ngmax, wh, K):
    rho = CPU.density(nb, nc, ngmax, s["x"], s["y"], s["z"], s["h"], s["m"], wh, K, box)
    p, c = CPU.eos(s["u"], rho, 5.0 / 3.0)
    cs = CPU.iad(nb, nc, ngmax, s["x"], s["y"], s["z"], s["h"], s["m"], rho, wh, K, box)
    return rho, p, c, cs


@pytest.mark.parametrize("periodic", [True, False])
def test_reference_density_matches_oracle(periodic):
    s, box, nb, nc, ngmax, wh, K = _setup(periodic=periodic)
    i, j = R.pairs_from_lists(nb, nc, ngmax)
    rho_ref = CPU.density(nb, nc, ngmax, s["x"], s["y"], s["z"], s["h"], s["m"], wh, K, box)
    rho = R.density(R.tt(s["x"]), R.tt(s["y"]), R.tt(s["z"]), R.tt(s["h"]), R.tt(s["m"]), i, j, R.tt(wh), K, box)
    assert np.max(np.abs(rho.numpy() - rho_ref) / rho_ref) < 1e-13


def test_reference_momentum_energy_matches_oracle():
    s, box, nb, nc, ngmax, wh, K = _setup()
    rho, p, c, cs = _oracle_fields(s, box, nb, nc, ngmax, wh, K)
    ax, ay, az, du, _ = CPU.momentum_energy(nb, nc, ngmax, s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["h"],
                                            s["m"], rho, p, c, cs, wh, K, 0.2, box)
    f = {k: R.tt(s[k]) for k in ("vx", "vy", "vz", "h", "m")}
    f.update(rho=R.tt(rho), p=R.tt(p), c=R.tt(c))
    f.update({k: R.tt(v) for k, v in zip(("c11", "c12", "c13", "c22", "c23", "c33"), cs)})
    i, j = R.pairs_from_lists(nb, nc, ngmax)
    got = R.momentum_energy(R.tt(s["x"]), R.tt(s["y"]), R.tt(s["z"]), f, i, j, R.tt(wh), K, box)
    for g, ref in zip(got, (ax, ay, az, du)):
        assert np.max(np.abs(g.numpy() - ref)) / np.max(np.abs(ref)) < 1e-12


def _fd_check(loss_oracle, loss_torch, s, probe, delta):
    xyz = [R.tt(s[k]).requires_grad_(True) for k in ("x", "y", "z")]
    grads = torch.autograd.grad(loss_torch(*xyz), xyz)
    scale = max(float(g.abs().max()) for g in grads)
    worst = 0.0
    for p in probe:
        for a, key in enumerate(("x", "y", "z")):
            plus, minus = dict(s), dict(s)
            plus[key] = s[key].copy()
            minus[key] = s[key].copy()
            plus[key][p] += delta
            minus[key][p] -= delta
            fd = (loss_oracle(plus) - loss_oracle(minus)) / (2 * delta)
            worst = max(worst, abs(fd - float(grads[a][p])) / scale)
    return worst


def test_density_gradient_vs_oracle_finite_differences():
    """L = sum (rho - mean rho)^2; FD step 2e-9 (inside one table segment for nearly all pairs)."""
    s, box, nb, nc, ngmax, wh, K = _setup()
    i, j = R.pairs_from_lists(nb, nc, ngmax)

    def lo(f):
        rho = CPU.density(nb, nc, ngmax, f["x"], f["y"], f["z"], f["h"], f["m"], wh, K, box)
        return float(np.sum((rho - rho.mean()) ** 2))

    def lt(x, y, z):
        rho = R.density(x, y, z, R.tt(s["h"]), R.tt(s["m"]), i, j, R.tt(wh), K, box)
        return ((rho - rho.mean()) ** 2).sum()

    probe = np.random.default_rng(3).choice(len(s["x"]), 24, replace=False)
    assert _fd_check(lo, lt, s, probe, 2e-9) < 1e-6  # measured 2.2e-7


def test_density_h_gradient_vs_oracle_finite_differences():
    s, box, nb, nc, ngmax, wh, K = _setup()
    i, j = R.pairs_from_lists(nb, nc, ngmax)
    h = R.tt(s["h"]).requires_grad_(True)
    rho = R.density(R.tt(s["x"]), R.tt(s["y"]), R.tt(s["z"]), h, R.tt(s["m"]), i, j, R.tt(wh), K, box)
    (gh,) = torch.autograd.grad(((rho - rho.mean()) ** 2).sum(), [h])
    scale = float(gh.abs().max())

    def lo(hh):
        r = CPU.density(nb, nc, ngmax, s["x"], s["y"], s["z"], hh, s["m"], wh, K, box)
        return float(np.sum((r - r.mean()) ** 2))

    worst = 0.0
    for p in np.random.default_rng(4).choice(len(h), 16, replace=False):
        hp, hm = s["h"].copy(), s["h"].copy()
        hp[p] += 1e-9
        hm[p] -= 1e-9
        fd = (lo(hp) - lo(hm)) / 2e-9
        worst = max(worst, abs(fd - float(gh[p])) / scale)
    # step 1e-9: rounding dominates below it (1e-10: 2.6e-6, 1e-11: 3.2e-5); measured 3.0e-7
    assert worst < 1e-6


def test_momentum_energy_gradient_vs_oracle_finite_differences():
    """L = sum_i (w1 ax + w2 ay + w3 az + w4 du) with random weights, rho, p, c and the IAD
    tensors held fixed."""
    s, box, nb, nc, ngmax, wh, K = _setup()
    rho, p, c, cs = _oracle_fields(s, box, nb, nc, ngmax, wh, K)
    rng = np.random.default_rng(5)
    W = rng.normal(size=(4, len(rho)))
    f = {k: R.tt(s[k]) for k in ("vx", "vy", "vz", "h", "m")}
    f.update(rho=R.tt(rho), p=R.tt(p), c=R.tt(c))
    f.update({k: R.tt(v) for k, v in zip(("c11", "c12", "c13", "c22", "c23", "c33"), cs)})
    i, j = R.pairs_from_lists(nb, nc, ngmax)

    def lo(g):
        out = CPU.momentum_energy(nb, nc, ngmax, g["x"], g["y"], g["z"], s["vx"], s["vy"], s["vz"], s["h"], s["m"],
                                  rho, p, c, cs, wh, K, 0.2, box)
        return float(sum(np.dot(W[k], out[k]) for k in range(4)))

    def lt(x, y, z):
        out = R.momentum_energy(x, y, z, f, i, j, R.tt(wh), K, box)
        return sum((R.tt(W[k]) * out[k]).sum() for k in range(4))

    probe = rng.choice(len(rho), 24, replace=False)
    assert _fd_check(lo, lt, s, probe, 2e-9) < 1e-6  # measured 1.4e-8
