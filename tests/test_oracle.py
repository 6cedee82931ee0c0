"""CPU checks that pin the oracle (oracle/sphx_oracle.cpp) independently of the GPU path.

Each test re-asserts a known answer from the reference's own test suite (file:line
cited), a closed form, or a brute-force computation; a plausible mistake in the
oracle (dropped term, wrong sign/index, transposed operand) fails at least one.
"""
import math
import os

import numpy as np
import pytest

import oracle
from oracle import BC_OPEN, BC_PERIODIC, Box, Cpu

CPU = Cpu("oracle")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROOT_RANGE = 1 << 63


# ----------------------------------------------------------------------------- Hilbert keys
def test_hilbert_first_order_curve():
    """domain/test/unit/sfc/hilbert.cpp:83-114: octant order of the first level."""
    hilbert_to_morton = [0, 1, 3, 2, 6, 7, 5, 4]
    half = (1 << 21) // 2
    for xi in (0, 1):
        for yi in (0, 1):
            for zi in (0, 1):
                morton = 4 * xi + 2 * yi + zi
                for off in (0, half - 1):
                    key = CPU.ihilbert(half * xi + off, half * yi + off, half * zi + off)
                    octant = (key >> 60) & 7
                    assert morton == hilbert_to_morton[octant]


def test_hilbert_continuity():
    """hilbert.cpp:116-145: consecutive keys across octant boundaries are unit neighbours."""
    for level in range(1, 21):
        for octant in range(8 if level > 1 else 7):
            last = (octant + 1) * (1 << (3 * (21 - level))) - 1
            a = CPU.decode_hilbert(last)
            b = CPU.decode_hilbert(last + 1)
            assert sum(abs(u - v) for u, v in zip(a, b)) == 1


def test_hilbert_inversion():
    """hilbert.cpp:147-180: decode(encode(p)) == p for random integer points."""
    rng = np.random.default_rng(0)
    for p in rng.integers(0, 1 << 21, (1000, 3)):
        assert CPU.decode_hilbert(CPU.ihilbert(*map(int, p))) == tuple(map(int, p))


def test_hilbert_ibox_known_answer():
    """hilbert.cpp:200-208: node 03 at level 1 is the box x<1/2, y>=1/2, z<1/2; checked
    through the node centre/size of that node in the unit box."""
    prefix = np.array([(1 << 3) | 3], np.uint64)  # placeholder-bit prefix of octal digit 3 at level 1
    c, s = CPU.node_centers(prefix, Box.cube(0.0, 1.0))
    assert np.allclose(c[0], [0.25, 0.75, 0.25]) and np.allclose(s[0], [0.25, 0.25, 0.25])


def test_sfc_keys_quantisation():
    """sfc.hpp:141-159: keys of the box corners and the clamp at the upper face."""
    box = Box.cube(-1.0, 3.0)
    x = np.array([-1.0, 3.0, 1.0])
    k = CPU.sfc_keys(x, x.copy(), x.copy(), box)
    assert k[0] == 0
    m = (1 << 21) - 1
    assert CPU.decode_hilbert(int(k[1])) == (m, m, m)
    assert CPU.decode_hilbert(int(k[2])) == (1 << 20,) * 3


def test_sfc_keys_remove_key_untouched():
    box = Box.cube(0.0, 1.0)
    x = np.array([0.1, 0.2])
    keys = np.array([1 << 63, 0], np.uint64)
    k = CPU.sfc_keys(x, x.copy(), x.copy(), box, keys)
    assert k[0] == 1 << 63 and k[1] != 0


# ----------------------------------------------------------------------------- sort
def test_sort_stable_bruteforce():
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 40, 5000).astype(np.uint64)
    vals = np.arange(5000, dtype=np.uint32)
    k, v = CPU.sort_by_key(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


# ----------------------------------------------------------------------------- cornerstone leaves
def _tree_divide_root_and_octant0():
    """OctreeMaker{}.divide().divide(0).makeTree() (tree/cs_util.hpp:137-200)."""
    l1, l2 = 1 << 60, 1 << 57
    return np.array([c * l2 for c in range(8)] + [c * l1 for c in range(1, 8)] + [ROOT_RANGE], np.uint64)


def test_rebalance_decision_known_answer():
    """domain/test/unit/tree/csarray.cpp:102-126."""
    tree = _tree_divide_root_and_octant0()
    counts = np.array([1, 1, 1, 0, 0, 0, 0, 0, 2, 3, 4, 5, 6, 7, 8], np.uint32)
    ops = np.zeros(15, np.int32)
    conv = CPU._rebalance_decision(oracle._p(tree), oracle._p(counts), 15, 4, oracle._p(ops))
    assert list(ops) == [1, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 8, 8, 8, 8]
    assert conv == 0


def test_rebalance_decision_single_root():
    """csarray.cpp:128-146."""
    tree = np.array([0, ROOT_RANGE], np.uint64)
    counts = np.array([1], np.uint32)
    ops = np.zeros(1, np.int32)
    assert CPU._rebalance_decision(oracle._p(tree), oracle._p(counts), 1, 4, oracle._p(ops)) == 1
    assert list(ops) == [1]


def _check_cornerstone(leaves, counts, keys, bucket):
    assert leaves[0] == 0 and leaves[-1] == ROOT_RANGE
    d = np.diff(leaves.astype(object))
    for r in d:
        assert r > 0 and (r & (r - 1)) == 0 and (r.bit_length() - 1) % 3 == 0  # power of 8
    ref_counts = np.searchsorted(keys, leaves[1:], "left") - np.searchsorted(keys, leaves[:-1], "left")
    assert np.array_equal(counts, ref_counts.astype(np.uint32))
    level = np.array([(63 - (int(r).bit_length() - 1)) // 3 for r in d])
    assert ((counts <= bucket) | (level == 21)).all()
    # the parent of every complete group of 8 sibling leaves holds more than bucket particles
    i = 0
    while i + 8 <= len(d):
        if level[i] > 0 and all(level[i:i + 8] == level[i]) and (int(leaves[i]) >> (3 * (21 - level[i]))) & 7 == 0 \
                and int(leaves[i + 8]) - int(leaves[i]) == 8 * int(d[i]):
            assert counts[i:i + 8].sum() > bucket
        i += 1


@pytest.mark.parametrize("n,bucket,kind", [(0, 64, "u"), (5, 64, "u"), (20000, 64, "u"), (20000, 8, "g"),
                                           (3000, 16, "dup")])
def test_compute_octree_invariants(n, bucket, kind):
    rng = np.random.default_rng(n + bucket)
    if kind == "u":
        keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    elif kind == "g":
        p = np.clip(rng.normal(0.5, 0.05, (3, n)), 0, 1 - 1e-12)
        keys = CPU.sfc_keys(p[0].copy(), p[1].copy(), p[2].copy(), Box.cube(0, 1))
    else:
        keys = rng.integers(0, 10, n).astype(np.uint64) << np.uint64(40)
    keys = np.sort(keys)
    leaves, counts = CPU.compute_octree(keys, bucket)
    _check_cornerstone(leaves, counts, keys, bucket)


# ----------------------------------------------------------------------------- internal octree
def test_octree_4x4x4_known_shape():
    """domain/test/unit/tree/octree.cpp:89-115: 64 level-2 leaves -> 9 internal nodes, 73 nodes."""
    l2 = 1 << 57
    leaves = np.array([c * l2 for c in range(64)] + [ROOT_RANGE], np.uint64)
    o = CPU.build_octree(leaves)
    assert o["numInternalNodes"] == 9 and o["numNodes"] == 73
    lr = o["levelRange"]
    assert lr[1] - lr[0] == 1 and lr[2] - lr[1] == 8 and lr[3] - lr[2] == 64 and lr[-1] == 73
    _check_connectivity(o)


def test_octree_irregular_known_shape():
    """octree.cpp:117-150: 15 leaves -> 2 internal nodes; the root points at the level-1 node."""
    o = CPU.build_octree(_tree_divide_root_and_octant0())
    assert o["numInternalNodes"] == 2 and o["numNodes"] == 17
    co = o["childOffsets"]
    assert co[0] == 1 and co[1] == 9  # root -> level 1, first level-1 node -> its 8 children
    _check_connectivity(o)


def _check_connectivity(o):
    """octree.cpp:27-74: every child of node p has parent p and prefix one level below."""
    pre, co, par = o["prefixes"], o["childOffsets"], o["parents"]
    for p in range(o["numNodes"]):
        if co[p]:
            c0 = co[p]
            assert par[(c0 - 1) // 8] == p
            lp = (63 - int(pre[p]).bit_length() + 1)
            for c in range(8):
                child = int(pre[c0 + c])
                assert child >> 3 == int(pre[p]) and (child & 7) == c


# ----------------------------------------------------------------------------- neighbors
def _brute_force(x, y, z, h, box):
    """all2allNeighbors (domain/test/unit/neighbors/all_to_all.hpp:27-65): FP64 predicate with
    the minimum image in periodic dimensions, sorted neighbor sets."""
    n = len(x)
    P = np.stack([x, y, z], 1)
    L = box.lengths()
    pbc = np.array([box.bx, box.by, box.bz]) == BC_PERIODIC
    out = []
    for i in range(n):
        d = P - P[i]
        for k in range(3):
            if pbc[k]:
                d[:, k] -= L[k] * np.rint(d[:, k] * (1.0 / L[k]))
        d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        nb = np.nonzero(d2 < 4.0 * h[i] * h[i])[0]
        out.append(nb[nb != i])
    return out


@pytest.mark.parametrize("radius,lo,hi,bc", [(0.124, 0.0, 1.0, BC_OPEN), (0.124, 0.0, 1.0, BC_PERIODIC),
                                             (0.0624, -1.0, 2.0, BC_PERIODIC), (0.0624 * 3, -1.0, 2.0, BC_OPEN)])
def test_find_neighbors_brute_force(radius, lo, hi, bc):
    """domain/test/unit/neighbors/findneighbors.cpp:43-134 (2,500 particles)."""
    rng = np.random.default_rng(42)
    n = 2500
    x, y, z = (lo + (hi - lo) * rng.random(n) for _ in range(3))
    box = Box.cube(lo, hi, bc)
    keys = CPU.sfc_keys(x, y, z, box)
    keys, order = CPU.sort_by_key(keys, np.arange(n, dtype=np.uint32))
    x, y, z = x[order], y[order], z[order]
    h = np.full(n, radius / 2)
    t = CPU.tree(keys, box)
    nb, nc = CPU.find_neighbors(x, y, z, h, box, t, n)
    ref = _brute_force(x, y, z, h, box)
    for i in range(n):
        assert nc[i] == len(ref[i])
        assert np.array_equal(nb[i, :nc[i]], ref[i])


# ----------------------------------------------------------------------------- kernel tables
def test_simpson_known_answer():
    """sph/test/table_creation.cpp:17-24: Simpson with 2000 intervals on sin over [0, pi/2]."""
    assert abs(CPU._simpson_sin(0.0, math.pi / 2, 2000) - 1.0) < 1e-14


def test_kernel_normalisation():
    """table_creation.cpp:26-40 (K vs sphynx_3D_k(6) within 1e-4) and the closed-form integral
    1/K = int_0^2 4 pi v^2 sinc(pi v / 2)^6 dv evaluated independently."""
    from scipy.integrate import quad

    wh, wh32, K = CPU.kernel_tables(6.0)
    assert abs(K - CPU.sphynx_k(6.0)) < 1e-4
    f = lambda v: 4 * math.pi * v * v * (math.sin(math.pi * v / 2) / (math.pi * v / 2)) ** 6 if v > 0 else 0.0
    I, _ = quad(f, 0.0, 2.0, epsabs=1e-14, epsrel=1e-14, limit=200)
    assert abs(K - 1.0 / I) < 1e-11
    assert abs(K - 0.790449589432) < 1e-12


def test_table_nodes_and_lookup():
    wh, wh32, K = CPU.kernel_tables(6.0)
    v = np.array([0.0, 0.5, 1.0, 1.5, 1.9999])
    exact = np.array([1.0] + [(math.sin(math.pi * a / 2) / (math.pi * a / 2)) ** 6 for a in v[1:]])
    for a, e in zip(v, exact):
        assert abs(CPU._lookup(oracle._p(wh), a) - e) < 1e-8
    assert CPU._lookup(oracle._p(wh), 2.0) == 0.0 and CPU._lookup(oracle._p(wh), 2.5) == 0.0
    assert wh[0] == 1.0 and abs(wh[-1]) < 1e-30
    assert np.max(np.abs(wh32 - wh)) < 1e-7


# ----------------------------------------------------------------------------- IAD / momentum-energy / density
def _std_fixture():
    """sph/test/std.cpp:45-95: five particles, open box [0, 6], K = sphynx_3D_k(6)."""
    d = dict(x=[1.0, 1.1, 3.2, 1.3, 2.4], y=[1.1, 1.2, 1.3, 4.4, 5.5], z=[1.2, 2.3, 1.4, 1.5, 1.6],
             h=[5.0, 5.1, 5.2, 5.3, 5.4], m=[1.1, 1.2, 1.3, 1.4, 1.5], rho=[0.014, 0.015, 0.016, 0.017, 0.018],
             vx=[0.010, -0.020, 0.030, -0.040, 0.050], vy=[-0.011, 0.021, -0.031, 0.041, -0.051],
             vz=[0.091, -0.081, 0.071, -0.061, 0.055], c=[0.4, 0.5, 0.6, 0.7, 0.8], p=[0.2, 0.3, 0.4, 0.5, 0.6],
             c11=[0.21, 0.27, 0.10, 0.45, 0.46], c12=[-0.22, -0.29, -0.11, -0.44, -0.47],
             c13=[-0.23, -0.31, -0.12, -0.43, -0.48], c22=[0.24, 0.32, 0.13, 0.42, 0.49],
             c23=[-0.25, -0.33, -0.14, -0.41, -0.50], c33=[0.26, 0.34, 0.15, 0.40, 0.51])
    return {k: np.array(v, np.float64) for k, v in d.items()}


def test_iad_known_answer():
    """std.cpp:98-111."""
    f = _std_fixture()
    wh, _, _ = CPU.kernel_tables(6.0)
    K = CPU.sphynx_k(6.0)
    out = np.zeros(6)
    nb = np.array([1, 2, 3, 4], np.uint32)
    CPU._iad_one(0, oracle._p(nb), 4, *[oracle._p(f[k]) for k in ("x", "y", "z", "h", "m", "rho")], oracle._p(wh), K,
                 oracle.C.byref(Box.cube(0.0, 6.0)), oracle._p(out))
    ref = [0.68826690779384281, -0.12963692768970825, -0.20435302538490346, 0.39616100688793993,
           -0.16797800827029263, 1.9055087813473524]
    assert np.max(np.abs(out - ref)) < 1e-8


def _me_one(cnt):
    f = _std_fixture()
    wh, _, _ = CPU.kernel_tables(6.0)
    K = CPU.sphynx_k(6.0)
    out = np.full(5, -1.0)
    nb = np.array([1, 2, 3, 4], np.uint32)
    keys = ("x", "y", "z", "vx", "vy", "vz", "h", "m", "rho", "p", "c", "c11", "c12", "c13", "c22", "c23", "c33")
    CPU._momentum_energy_one(0, oracle._p(nb), cnt, *[oracle._p(f[k]) for k in keys], oracle._p(wh), K,
                             oracle.C.byref(Box.cube(0.0, 6.0)), oracle._p(out))
    return out


def test_momentum_energy_known_answer():
    """std.cpp:113-127 (out = grad_Px, grad_Py, grad_Pz, du, maxvsignal)."""
    gx, gy, gz, du, mv = _me_one(4)
    assert abs(gx - 14.407211846688075) < 1.3e-7
    assert abs(gy + 1.2396802157028355) < 1.4e-7
    assert abs(gz - 15.596554152643426) < 2.15e-7
    assert abs(du + 0.40541191600274296) < 1e-8
    assert abs(mv - 1.4112466828564341) < 1e-10


def test_momentum_energy_no_neighbors():
    """std.cpp:129-143."""
    assert list(_me_one(0)) == [0.0, 0.0, 0.0, 0.0, 0.0]


def test_xmass_known_answer():
    """sph/test/ve.cpp:240-249 on the reference fixture example_data.txt (99 particles)."""
    cols = np.loadtxt(os.path.join(GOLDEN, "example_data.txt"))
    names = ["x", "y", "z", "vx", "vy", "vz", "h", "c", "c11", "c12", "c13", "c22", "c23", "c33", "p", "gradh",
             "rho0", "sumwhrho0", "sumwh", "dvxdx", "dvxdy", "dvxdz", "dvydx", "dvydy", "dvydz", "dvzdx", "dvzdy",
             "dvzdz", "alpha", "u", "divv"]
    f = {k: np.ascontiguousarray(cols[:, c]) for c, k in enumerate(names)}
    mpart = 3.781038064465603e26
    m = np.full(99, mpart)
    wh, _, _ = CPU.kernel_tables(6.0)
    K = CPU.sphynx_k(6.0)
    xm = CPU.xmass_one(0, np.arange(1, 99, dtype=np.uint32), f["x"], f["y"], f["z"], f["h"], m, wh, K,
                       Box.cube(-1e9, 1e9))
    rho0 = m[0] / xm
    assert abs(rho0 - 34.515038498081417) < 7.33e-7
    assert abs(xm - m[0] / f["rho0"][0]) < m[0] / f["rho0"][0] * 1e-7


def test_density_brute_force():
    """rho_i = K / h_i^3 (m_i + sum_j m_j W(r_ij / h_i)) against a direct numpy sum over the
    brute-force neighbor sets with the same interpolated table."""
    rng = np.random.default_rng(5)
    n = 1500
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    box = Box.cube(0.0, 1.0, BC_PERIODIC)
    keys = CPU.sfc_keys(x, y, z, box)
    keys, order = CPU.sort_by_key(keys, np.arange(n, dtype=np.uint32))
    x, y, z = x[order], y[order], z[order]
    h = np.full(n, 0.5 * (60 / (4 / 3 * np.pi * n)) ** (1 / 3))
    m = rng.random(n) + 0.5
    t = CPU.tree(keys, box)
    ngmax = 300
    nb, nc = CPU.find_neighbors(x, y, z, h, box, t, ngmax)
    wh, _, K = CPU.kernel_tables(6.0)
    rho = CPU.density(nb, (nc + 1).astype(np.uint32), ngmax, x, y, z, h, m, wh, K, box)
    ref = _brute_force(x, y, z, h, box)
    idx = np.arange(19999)
    for i in range(0, n, 37):
        j = ref[i]
        d = np.stack([x[j] - x[i], y[j] - y[i], z[j] - z[i]], 1)
        d -= np.rint(d)
        r = np.sqrt((d * d).sum(1)) / h[i]
        k = np.minimum((r * 19999 / 2).astype(int), 19998)
        w = wh[k] + (wh[k + 1] - wh[k]) * (19999 / 2) * (r - k * (2 / 19999))
        expect = K / h[i] ** 3 * (m[i] + np.sum(w * m[j]))
        assert abs(rho[i] - expect) / expect < 1e-12


def test_eos_closed_form():
    u = np.array([0.5, 1.0, 2.0])
    rho = np.array([1.0, 2.0, 3.0])
    p, c = CPU.eos(u, rho, 5.0 / 3.0)
    assert np.allclose(p, rho * u * (2.0 / 3.0), rtol=1e-15, atol=0)
    assert np.allclose(c, np.sqrt(5.0 / 3.0 * 2.0 / 3.0 * u), rtol=1e-15, atol=0)


def test_update_h_closed_form():
    """sph/kernels.hpp:27-33: h * 0.5 * (1 + 1023 ng0 / nc)^(1/10); equal counts leave h ~ unchanged."""
    assert abs(CPU._update_h(100, 100, 1.0) - 0.5 * 1024 ** 0.1) < 1e-15
    assert abs(CPU._update_h(100, 100, 1.0) - 1.0) < 1e-15


# ----------------------------------------------------------------------------- target groups
def _leaves_uniform(level):
    return np.arange(8 ** level + 1, dtype=np.uint64) * np.uint64(1 << (63 - 3 * level))


def _layout_all_in_leaf0(n, num_leaves):
    layout = np.full(num_leaves + 1, n, np.uint32)
    layout[0] = 0
    return layout


def _line(n, step, jumps):
    """Points along x with spacing `step`; jumps[i] = extra offset added before particle i."""
    x = np.zeros(n)
    for i in range(1, n):
        x[i] = x[i - 1] + step + jumps.get(i, 0.0)
    return x, np.full(n, 0.3), np.full(n, 0.3)


def test_groups_split_on_h_criterion():
    """Root leaf (distCrit = 2): only the 2h criterion can split; a 0.51 jump with h = 0.1
    (2h = 0.2) splits, steps of 0.001 do not (groups_gpu.cuh:37-77)."""
    n = 130
    x, y, z = _line(n, 0.001, {21: 0.5, 100: 0.5})
    h = np.full(n, 0.1)
    leaves = _leaves_uniform(0)
    b = CPU.compute_groups(x, y, z, h, leaves, _layout_all_in_leaf0(n, 1), Box.cube(0.0, 1.0))
    assert b.tolist() == [0, 21, 64, 100, 128, 130]
    # without h the threshold is min(distCrit^2, 1) = 1: no split
    b = CPU.compute_groups(x, y, z, None, leaves, _layout_all_in_leaf0(n, 1), Box.cube(0.0, 1.0))
    assert b.tolist() == [0, 64, 128, 130]


def test_groups_leaf_level_sets_distcrit():
    """distCrit = tolFactor * cbrt(leaf volume): level 2 leaves (edge 1/4) with tol 2 give
    distCrit 0.5, so a 0.6 jump splits; level 1 (distCrit 1) does not."""
    n = 64
    x, y, z = _line(n, 0.001, {21: 0.6})
    for level, expect in ((1, [0, 64]), (2, [0, 21, 64])):
        leaves = _leaves_uniform(level)
        b = CPU.compute_groups(x, y, z, None, leaves, _layout_all_in_leaf0(n, len(leaves) - 1), Box.cube(0.0, 1.0))
        assert b.tolist() == expect, level
    # tolFactor scales distCrit: tol 1 at level 1 gives distCrit 0.5 as well
    leaves = _leaves_uniform(1)
    b = CPU.compute_groups(x, y, z, None, leaves, _layout_all_in_leaf0(n, 8), Box.cube(0.0, 1.0), tol=1.0)
    assert b.tolist() == [0, 21, 64]


def test_groups_smallest_leaf_of_the_group_counts():
    """The group's node volume is the minimum over its particles' leaves: one particle in a
    level-2 leaf makes group 0 split at a 0.6 jump; group 1, all in a level-1 leaf, does not."""
    n = 128
    x, y, z = _line(n, 0.001, {21: 0.6, 85: -0.6})
    step1, step2 = 1 << 60, 1 << 57
    leaves = np.array([k * step2 for k in range(8)] + [k * step1 for k in range(1, 9)], np.uint64)
    layout = np.array([0] + [1] * 8 + [n] * 7, np.uint32)
    assert len(layout) == len(leaves)
    b = CPU.compute_groups(x, y, z, None, leaves, layout, Box.cube(0.0, 1.0))
    assert b.tolist() == [0, 21, 64, 128]


def test_groups_box_scaling_and_ragged_range():
    """Coordinates are divided by the box lengths and 2h by the smallest extent; [first, last)
    need not start at 0 and the last group is ragged (its tail lanes repeat last - 1)."""
    n = 160
    x, y, z = _line(n, 0.0005, {30: 0.15, 40: 0.075})
    x, y, z = 2 * x, 2 * y, 2 * z  # box [0, 2]^3: physical jumps 0.3 and 0.15, 0.15 and 0.075 in the unit box
    h = np.full(n, 0.1)            # 2h / minExtent = 0.1 in the unit box
    b = CPU.compute_groups(x, y, z, h, _leaves_uniform(0), _layout_all_in_leaf0(n, 1), Box.cube(0.0, 2.0),
                           first=5, last=150)
    assert b.tolist() == [5, 30, 69, 133, 150]
    assert CPU.compute_groups(x, y, z, h, _leaves_uniform(0), _layout_all_in_leaf0(n, 1), Box.cube(0.0, 2.0),
                              first=7, last=7).tolist() == [7]


def _find_splits_case():
    """The reference's findSplits known answer (domain/test/unit_cuda/traversal/groups.cu:51-112)
    through the group interface: x = y = z = lane, distCrit^2 = 3.01 (tolFactor = sqrt(3.01) on
    a root leaf), 2h = 64 except lane 2 (2h = 0.99 sqrt 3); lane 0 moved to -1, lanes 31 and
    33 moved by -0.5.  Splits after lanes 0, 2, 31 and 33."""
    n = 64
    p = np.arange(n, dtype=np.float64)
    p[0] = -1.0
    p[31] -= 0.5
    p[33] -= 0.5
    h = np.full(n, 32.0)
    h[2] = 0.5 * 0.99 * np.sqrt(3.0)
    return p, h, float(np.float32(np.sqrt(3.01)))


def test_groups_reference_find_splits_known_answer():
    p, h, tol = _find_splits_case()
    b = CPU.compute_groups(p, p.copy(), p.copy(), h, _leaves_uniform(0), _layout_all_in_leaf0(64, 1),
                           Box.cube(0.0, 1.0), tol=tol)
    assert b.tolist() == [0, 1, 3, 32, 34, 64]


# ----------------------------------------------------------------------------- global box
def test_make_global_box_known_answer():
    """makeGlobalBox: open dimensions take the coordinate extrema, periodic and fixed ones keep
    their limits, n = 0 keeps the previous box (box_mpi.hpp:68-107)."""
    x = np.array([0.3, -2.0, 5.5, 1.0])
    y = np.array([7.0, 7.5, -1.0, 0.0])
    z = np.array([0.1, 0.2, 0.3, 0.4])
    prev = Box.make([0.0, 0.0, 0.0], [1.0, 2.0, 3.0], [oracle.BC_OPEN, oracle.BC_PERIODIC, oracle.BC_FIXED])
    b = CPU.make_global_box(x, y, z, prev)
    assert (b.xmin, b.xmax, b.ymin, b.ymax, b.zmin, b.zmax) == (-2.0, 5.5, 0.0, 2.0, 0.0, 3.0)
    assert (b.bx, b.by, b.bz) == (oracle.BC_OPEN, oracle.BC_PERIODIC, oracle.BC_FIXED)
    e = CPU.make_global_box(x[:0], y[:0], z[:0], prev)
    assert (e.xmin, e.xmax) == (0.0, 1.0)
