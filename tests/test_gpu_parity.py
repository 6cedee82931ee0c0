"""GPU parity: every stage of the HIP path against the CPU oracle on the same inputs.

Tolerances (stated per field):
* keys, sort, leaves, counts, link arrays, centres, neighbor counts and lists: bit-exact;
* FP64 with SPHX_OPT_FP64_EXACT: bit-exact for every field;
* FP64 default (FMA contraction): rho <= 1e-13 relative, IAD/ax/du <= 1e-10 of the field's max;
* MIXED: rho <= 1e-5 relative (the target), IAD <= 1e-4, ax/ay/az/du <= 1e-3 of the field's max.
"""
import numpy as np
import pytest
import torch

import oracle
from tests.helpers import oracle_step, rel_err, sedov_lattice, tables, uniform_box

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from sphexa_amd import hydro

    c = hydro.Context(0)
    yield c
    c.close()


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def u64(t):
    return t.cpu().numpy().view(np.uint64)


def u32(t):
    return t.cpu().numpy().view(np.uint32)


def hbox(b):
    from sphexa_amd import hydro

    return hydro.make_box([b.xmin, b.ymin, b.zmin], [b.xmax, b.ymax, b.zmax], [b.bx, b.by, b.bz])


CPU = oracle.Cpu("oracle")


# ----------------------------------------------------------------------------- keys
@pytest.mark.parametrize("periodic", [True, False])
def test_keys_bit_exact(ctx, periodic):
    from sphexa_amd import hydro

    print("case keys periodic", periodic, flush=True)
    inp, box = uniform_box(50000, seed=3, periodic=periodic, lo=-1.7, hi=2.3)
    ref = CPU.sfc_keys(inp["x"], inp["y"], inp["z"], box)
    k = hydro.compute_sfc_keys(ctx, dev(inp["x"]), dev(inp["y"]), dev(inp["z"]), hbox(box))
    assert np.array_equal(u64(k), ref)


def test_keys_skip_remove_key(ctx):
    from sphexa_amd import hydro

    inp, box = uniform_box(1000, seed=4)
    keys0 = np.zeros(1000, np.uint64)
    keys0[::7] = np.uint64(1) << np.uint64(63)
    ref = CPU.sfc_keys(inp["x"], inp["y"], inp["z"], box, keys0)
    k = dev(keys0.view(np.int64))
    hydro.compute_sfc_keys(ctx, dev(inp["x"]), dev(inp["y"]), dev(inp["z"]), hbox(box), k)
    assert np.array_equal(u64(k), ref)


# ----------------------------------------------------------------------------- gather
@pytest.mark.parametrize("n", [1, 1000, 300007])
def test_gather_arrays_equal_take(ctx, n):
    """gatherArrays: every array permuted by one ordering in one launch, bit-exact with
    numpy's take, for 4- and 8-byte arrays (16 at once, the maximum)."""
    from sphexa_amd import hydro

    print("case gather_arrays", n, flush=True)
    rng = np.random.default_rng(n)
    order = rng.permutation(n).astype(np.int32)
    srcs = []
    for k in range(16):
        dt = (np.float64, np.float32, np.int32, np.int64)[k % 4]
        srcs.append(rng.standard_normal(n).astype(dt) if k % 4 < 2 else rng.integers(-2**30, 2**30, n).astype(dt))
    ds = [dev(a) for a in srcs]
    outs = [torch.empty_like(t) for t in ds]
    hydro.gather_arrays(ctx, dev(order), list(zip(ds, outs)))
    for a, o in zip(srcs, outs):
        assert np.array_equal(o.cpu().numpy(), a[order])
    one = torch.empty_like(ds[0])
    hydro.gather_arrays(ctx, dev(order), [(ds[0], one)])
    assert np.array_equal(one.cpu().numpy(), hydro.gather(ctx, dev(order), ds[0]).cpu().numpy())


def test_gather_arrays_rejects_overlap_and_count(ctx):
    from sphexa_amd import hydro
    from sphexa_amd._lib import SphxError

    order = dev(np.arange(64, dtype=np.int32))
    a, b = torch.zeros(64, dtype=torch.float64, device="cuda"), torch.zeros(64, dtype=torch.float64, device="cuda")
    with pytest.raises(SphxError, match="overlaps input"):
        hydro.gather_arrays(ctx, order, [(a, b), (b, torch.empty_like(b))])
    big = torch.zeros(128, dtype=torch.float64, device="cuda")
    with pytest.raises(SphxError, match="outputs 0 and 1 overlap"):
        hydro.gather_arrays(ctx, order, [(a, big[:64]), (b, big[32:96])])
    with pytest.raises(ValueError, match="at most 16"):
        hydro.gather_arrays(ctx, order, [(a, b)] * 17)


# ----------------------------------------------------------------------------- sort
@pytest.mark.parametrize("n,kind", [(1, "rand"), (100, "rand"), (4096, "rand"), (4097, "dup"), (150001, "rand"),
                                    (70000, "dup"), (30000, "lowbits")])
def test_sort_stable_bit_exact(ctx, n, kind):
    from sphexa_amd import hydro

    print("case sort", n, kind, flush=True)
    rng = np.random.default_rng(n)
    if kind == "rand":
        keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    elif kind == "dup":
        keys = rng.integers(0, 50, n, dtype=np.uint64) << np.uint64(40)
    else:
        keys = rng.integers(0, 256, n, dtype=np.uint64)
    vals = rng.permutation(n).astype(np.uint32)
    order = np.argsort(keys, kind="stable")
    k = dev(keys.view(np.int64))
    v = dev(vals.view(np.int32))
    hydro.sort_by_key(ctx, k, v)
    assert np.array_equal(u64(k), keys[order])
    assert np.array_equal(u32(v), vals[order])


def _run_keys(rng, n, maxlen, lowvals):
    """Keys in runs of equal top halves (run lengths 1..maxlen), low halves from lowvals values."""
    lens, tot = [], 0
    while tot < n:
        lens.append(int(rng.integers(1, maxlen + 1)))
        tot += lens[-1]
    top = np.sort(rng.choice(2 ** 31, len(lens), replace=False)).astype(np.uint64)
    hi = np.repeat(top, lens)[:n]
    return (hi << np.uint64(32)) | rng.integers(0, lowvals, n).astype(np.uint64)


# Arrays of at least 2^18 keys sort by their top halves in four passes and fix up the runs of
# equal top halves (at most 64 long); a longer run sends the sort through all eight passes.
# (run65_then_runs last: its fallback makes the context's next sorts take all eight passes)
@pytest.mark.parametrize("kind", ["rand", "runs", "runs_rev", "run64", "ones", "run65_then_runs"])
def test_sort_prefix_and_runs_bit_exact(ctx, kind):
    from sphexa_amd import hydro

    n = 300_007
    print("case prefix sort", kind, flush=True)
    rng = np.random.default_rng(len(kind))
    calls = []
    if kind == "rand":
        calls.append(rng.integers(0, 2 ** 63, n, dtype=np.uint64))
    elif kind == "runs":  # many equal low halves: the stable order of equal keys matters
        calls.append(rng.permutation(_run_keys(rng, n, 64, 8)))
    elif kind == "runs_rev":
        calls.append(_run_keys(rng, n, 64, 2 ** 32)[::-1].copy())
    elif kind == "run64":
        k = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
        k[1000:1064] = (k[1000] & ~np.uint64(0xFFFFFFFF)) | rng.integers(0, 50, 64).astype(np.uint64)
        calls.append(rng.permutation(k))
    elif kind == "run65_then_runs":
        k = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
        k[1000:1065] = (k[1000] & ~np.uint64(0xFFFFFFFF)) | rng.integers(0, 50, 65).astype(np.uint64)
        calls += [rng.permutation(k), rng.permutation(_run_keys(rng, n, 64, 8))]
    else:  # top halves of all ones, and the largest key
        k = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
        k[:40] = (np.uint64(0xFFFFFFFF) << np.uint64(32)) | rng.integers(0, 2 ** 32, 40).astype(np.uint64)
        k[40:43] = np.uint64(2 ** 64 - 1)
        calls.append(rng.permutation(k))
    for keys in calls:
        vals = rng.permutation(n).astype(np.uint32)
        order = np.argsort(keys, kind="stable")
        k = dev(keys.view(np.int64))
        v = dev(vals.view(np.int32))
        hydro.sort_by_key(ctx, k, v)
        assert np.array_equal(u64(k), keys[order])
        assert np.array_equal(u32(v), vals[order])


def _prefix_runs(rng, n, maxlen, lowvals, prefix_bits=24):
    """Keys in runs of equal top prefix_bits bits (run lengths 1..maxlen), the bits below
    from lowvals values."""
    lens, tot = [], 0
    while tot < n:
        lens.append(int(rng.integers(1, maxlen + 1)))
        tot += lens[-1]
    top = np.sort(rng.choice(2 ** (prefix_bits - 1), len(lens), replace=False)).astype(np.uint64)
    return (np.repeat(top, lens)[:n] << np.uint64(64 - prefix_bits)) | rng.integers(0, lowvals, n).astype(np.uint64)


# After a prefix sort whose runs of equal top 24 bits were all short, the context's next sort
# takes three passes (bits 40..63) and fixes up the runs from the work arrays into the
# caller's; a longer 24-bit run there sends it through the eight passes (after a copy back)
# and the sort after that takes four passes again.  The scatter launches show the path.
def test_sort_three_pass_prefix_bit_exact():
    from sphexa_amd import hydro

    n = 300_007
    rng = np.random.default_rng(24)
    c = hydro.Context(0)
    try:
        k65 = rng.integers(0, 2 ** 63, n, dtype=np.uint64)  # one 24-bit run of 65, 32-bit runs short
        k65[1000:1065] = (k65[1000] & ~np.uint64(2 ** 40 - 1)) | (
            rng.choice(2 ** 32, 65, replace=False).astype(np.uint64) << np.uint64(8))
        k64 = rng.integers(0, 2 ** 63, n, dtype=np.uint64)  # one 24-bit run of exactly 64
        k64[5:69] = (k64[5] & ~np.uint64(2 ** 40 - 1)) | rng.integers(0, 40, 64).astype(np.uint64)
        ones = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
        ones[:40] = (np.uint64(0xFFFFFF) << np.uint64(40)) | rng.integers(0, 2 ** 40, 40).astype(np.uint64)
        ones[40:43] = np.uint64(2 ** 64 - 1)
        calls = [  # (keys, prefix passes expected)
            ("rand", rng.integers(0, 2 ** 63, n, dtype=np.uint64), 4),
            ("runs24", rng.permutation(_prefix_runs(rng, n, 64, 8)), 3),
            ("runs24_rev", _prefix_runs(rng, n, 64, 2 ** 40)[::-1].copy(), 3),
            ("run24_64", rng.permutation(k64), 3),
            ("ones24", rng.permutation(ones), 3),
            ("run24_65", rng.permutation(k65), 3),
            ("runs24_after_fallback", rng.permutation(_prefix_runs(rng, n, 64, 8)), 4),
            ("rand_again", rng.integers(0, 2 ** 63, n, dtype=np.uint64), 3),
        ]
        for name, keys, passes in calls:
            print("case three-pass prefix sort", name, flush=True)
            vals = rng.permutation(n).astype(np.uint32)
            order = np.argsort(keys, kind="stable")
            k = dev(keys.view(np.int64))
            v = dev(vals.view(np.int32))
            c.set_timing(True)
            c.timing_summary()
            hydro.sort_by_key(c, k, v)
            t = c.timing_summary()
            c.set_timing(False)
            assert np.array_equal(u64(k), keys[order]), name
            assert np.array_equal(u32(v), vals[order]), name
            scatters = t["radixScatter"][0]
            if name == "run24_65":
                assert scatters > 3, (name, scatters)  # the prefix passes, then the full sort
                assert "runFix" not in t, name
            else:
                assert scatters == passes, (name, scatters)
                assert t["runFix"][0] == 1, name
    finally:
        c.close()


# ----------------------------------------------------------------------------- octree
def _keys(kind, n, seed=5):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        k = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    elif kind == "gauss":
        inp, box = uniform_box(n, seed)
        g = np.clip(rng.normal(0.5, 0.05, (3, n)), 0, 1 - 1e-12)
        k = CPU.sfc_keys(g[0].copy(), g[1].copy(), g[2].copy(), oracle.Box.cube(0, 1))
    elif kind == "dups":
        k = rng.integers(0, 20, n, dtype=np.uint64) << np.uint64(30)
    else:
        k = np.zeros(n, np.uint64)
    return np.sort(k)


@pytest.mark.parametrize("kind,n,bucket", [("uniform", 0, 64), ("uniform", 5, 64), ("uniform", 100000, 64),
                                           ("gauss", 100000, 64), ("gauss", 20000, 8), ("dups", 5000, 16),
                                           ("zeros", 300, 4), ("uniform", 3000, 1)])
def test_octree_leaves_counts_link(ctx, kind, n, bucket):
    from sphexa_amd import hydro

    print("case octree", kind, n, bucket, flush=True)
    keys = _keys(kind, n)
    ref_leaves, ref_counts = CPU.compute_octree(keys, bucket)
    kd = dev(keys.view(np.int64)) if n > 0 else torch.zeros(1, dtype=torch.int64, device="cuda")
    leaves, counts, layout = hydro.compute_octree(ctx, kd[:n] if n > 0 else kd[:0], bucket)
    assert np.array_equal(u64(leaves), ref_leaves)
    assert np.array_equal(u32(counts), ref_counts)
    ref_layout = np.concatenate([[0], np.cumsum(ref_counts)]).astype(np.uint32)
    assert np.array_equal(u32(layout), ref_layout)
    o = hydro.build_octree(ctx, leaves)
    r = CPU.build_octree(ref_leaves)
    assert np.array_equal(u64(o.prefixes), r["prefixes"])
    assert np.array_equal(o.childOffsets.cpu().numpy()[:o.numNodes], r["childOffsets"])
    assert np.array_equal(o.levelRange.cpu().numpy(), r["levelRange"])
    assert np.array_equal(o.internalToLeaf.cpu().numpy(), r["internalToLeaf"])
    assert np.array_equal(o.leafToInternal.cpu().numpy(), r["leafToInternal"])
    if o.numInternalNodes > 0:
        assert np.array_equal(o.parents.cpu().numpy(), r["parents"])
    box = oracle.Box.make([-1, 0, 2], [3, 1, 2.5], [1, 0, 1])
    c, s = hydro.node_fp_centers(ctx, o, hbox(box))
    rc, rs = CPU.node_centers(r["prefixes"], box)
    assert np.array_equal(c.cpu().numpy(), rc.ravel())
    assert np.array_equal(s.cpu().numpy(), rs.ravel())


# ----------------------------------------------------------------------------- neighbor search
def _gpu_tree(ctx, inp, box, bucket=64):
    from sphexa_amd import hydro

    n = len(inp["x"])
    keys = CPU.sfc_keys(inp["x"], inp["y"], inp["z"], box)
    keys, order = CPU.sort_by_key(keys, np.arange(n, dtype=np.uint32))
    s = {k: np.ascontiguousarray(v[order]) for k, v in inp.items()}
    t = hydro.tree_from_keys(ctx, dev(keys.view(np.int64)), hbox(box), bucket)
    return keys, s, t


@pytest.mark.parametrize("periodic,ng,ngmax,n", [(True, 100, 200, 20000), (False, 100, 200, 20000),
                                                 (True, 60, 40, 8000), (True, 800, 1024, 6000)])
def test_find_neighbors_bit_exact(ctx, periodic, ng, ngmax, n):
    from sphexa_amd import hydro

    print("case find_neighbors", periodic, ng, ngmax, n, flush=True)
    inp, box = uniform_box(n, seed=11, ng=ng, periodic=periodic)
    keys, s, t = _gpu_tree(ctx, inp, box)
    rt = CPU.tree(keys, box)
    ref_nb, ref_nc = CPU.find_neighbors(s["x"], s["y"], s["z"], s["h"], box, rt, ngmax)
    nb, nc = hydro.find_neighbors(ctx, dev(s["x"]), dev(s["y"]), dev(s["z"]), dev(s["h"]), hbox(box), t, ngmax)
    nc = u32(nc)
    assert np.array_equal(nc, ref_nc)
    nb = u32(nb.reshape(-1)).reshape(n, ngmax)
    for i in range(0, n, max(1, n // 2000)):
        c = min(nc[i], ngmax)
        assert np.array_equal(nb[i, :c], ref_nb[i, :c]), i


def _walk_counts(rt, p, r):
    """Breadth-first walk of the oracle's tree with the search box of points p (radius r,
    periodic unit box), as the GPU walk does it: (candidate leaves, largest frontier)."""
    co, C, S = rt["childOffsets"], rt["centers"], rt["sizes"]
    glo, ghi = (p - r).min(0), (p + r).max(0)
    gc, gh = 0.5 * (glo + ghi), 0.5 * (ghi - glo)
    front, leaves, widest = list(range(co[0], co[0] + 8)), 0, 8
    while front:
        d = C[front] - gc
        d -= np.rint(d)
        ov = np.all(np.abs(d) <= S[front] + gh, axis=1)
        nxt = []
        for node, o in zip(front, ov):
            if o and co[node] == 0:
                leaves += 1
            elif o:
                nxt += list(range(co[node], co[node] + 8))
        front, widest = nxt, max(widest, len(nxt))
    return leaves, widest


def test_find_neighbors_wide_search_boxes(ctx):
    """A group whose tree walk exceeds the first walk's capacities (512 frontier nodes or
    candidate leaves) is walked again with the larger ones; its lists stay bit-exact."""
    from sphexa_amd import hydro

    print("case find_neighbors wide", flush=True)
    n, ngmax, bucket = 60000, 256, 32
    inp, box = uniform_box(n, seed=13, ng=100, periodic=True)
    keys, s, t = _gpu_tree(ctx, inp, box, bucket=bucket)
    s["h"][:64] = 0.07  # one fixed group with search spheres of radius 0.14
    rt = CPU.tree(keys, box, bucket=bucket)
    p = np.stack([s["x"][:64], s["y"][:64], s["z"][:64]], 1)
    leaves, widest = _walk_counts(rt, p, 0.14)
    assert max(leaves, widest) > 512 and leaves <= 1024 and widest <= 2048, (leaves, widest)
    ref_nb, ref_nc = CPU.find_neighbors(s["x"], s["y"], s["z"], s["h"], box, rt, ngmax)
    nb, nc = hydro.find_neighbors(ctx, dev(s["x"]), dev(s["y"]), dev(s["z"]), dev(s["h"]), hbox(box), t, ngmax)
    nc = u32(nc)
    assert np.array_equal(nc, ref_nc)
    nb = u32(nb.reshape(-1)).reshape(n, ngmax)
    for i in list(range(64)) + list(range(64, n, 97)):
        c = min(nc[i], ngmax)
        assert np.array_equal(nb[i, :c], ref_nb[i, :c]), i


def _run_gpu_step(ctx, inp, box, precision, ng0=100, ngmax=150, exact=False):
    from sphexa_amd import hydro

    n = len(inp["x"])
    ctx.set_fp64_exact(exact)
    step = hydro.HydroStep(ctx, n, hbox(box), precision, ng0=ng0, ngmax=ngmax)
    dt = step.d.dtype
    d_inp = {k: dev(v, torch.float64 if k in ("x", "y", "z") else dt) for k, v in inp.items()}
    d = step.compute_forces(d_inp)
    torch.cuda.synchronize()
    ctx.set_fp64_exact(False)
    return step, d


FIELDS = ("rho", "c11", "c12", "c13", "c22", "c23", "c33", "ax", "ay", "az", "du")
IAD = ("c11", "c12", "c13", "c22", "c23", "c33")
ACC = ("ax", "ay", "az")


def field_scale(ref, f):
    """Error scale of a field: the tensor's diagonal max for IAD components (off-diagonals
    can be pure roundoff, e.g. on a lattice), the max over all three components for
    accelerations, the field's own max otherwise."""
    if f in IAD:
        s = max(np.max(np.abs(ref[g])) for g in ("c11", "c22", "c33"))
    elif f in ACC:
        s = max(np.max(np.abs(ref[g])) for g in ACC)
    else:
        s = np.max(np.abs(ref[f]))
    return s if s > 0 else 1.0


def field_err(got, ref, f):
    return float(np.max(np.abs(got[f] - ref[f])) / field_scale(ref, f))


@pytest.mark.parametrize("case", ["uniform_periodic", "uniform_open", "sedov"])
@pytest.mark.parametrize("mode", ["fp64_exact", "fp64", "mixed"])
def test_hydro_step_vs_oracle(ctx, case, mode):
    from sphexa_amd._lib import FP64, MIXED

    print("case hydro", case, mode, flush=True)
    if case == "sedov":
        inp, box = sedov_lattice(30)
    else:
        inp, box = uniform_box(20000, seed=21, periodic=(case == "uniform_periodic"))
        if case == "uniform_open":
            # the step fits open dimensions to the particles (makeGlobalBox); so does the oracle here
            box = CPU.make_global_box(inp["x"], inp["y"], inp["z"], box)
    ref = oracle_step(inp, box)
    prec = MIXED if mode == "mixed" else FP64
    step, d = _run_gpu_step(ctx, inp, box, prec, exact=(mode == "fp64_exact"))
    assert np.array_equal(u64(step.keys), ref["keys"])
    assert np.array_equal(u32(d.nc), ref["nc"])
    if case == "sedov":
        assert (ref["nc"] == 93).all()
    got = {f: getattr(d, f).double().cpu().numpy() for f in FIELDS}
    errs = {f: field_err(got, ref, f) for f in FIELDS}
    print("errors", {f: f"{e:.2e}" for f, e in errs.items()}, flush=True)
    if mode == "fp64_exact":
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), f
        assert d.minDtCourant == ref["minDt"]
    elif mode == "fp64":
        assert rel_err(got["rho"], ref["rho"]) < 1e-13
        for f in FIELDS[1:]:
            assert errs[f] < 1e-10, f
        assert abs(d.minDtCourant - ref["minDt"]) / ref["minDt"] < 1e-12
    else:
        assert rel_err(got["rho"], ref["rho"]) < 1e-5
        for f in IAD:
            assert errs[f] < 1e-4, f
        for f in ACC + ("du",):
            assert errs[f] < 1e-3, f
        assert abs(d.minDtCourant - ref["minDt"]) / ref["minDt"] < 1e-4


def test_h_iteration(ctx):
    """Start from h giving ~12 neighbors (< ng0/4): findNeighborsSph iterates h.

    Every h update applies the host-tabulated factor pow(1 + 1023 ng0 / nc, 0.1) of
    updateH, so h, the counts and the lists equal the oracle's bit for bit after the
    iterations."""
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64

    inp, box = uniform_box(6000, seed=31, ng=12)
    ref = oracle_step(inp, box)
    step, d = _run_gpu_step(ctx, inp, box, FP64)
    assert np.array_equal(d.h.cpu().numpy(), ref["h"])
    assert np.array_equal(u32(d.nc), ref["nc"])
    nb = u32(hydro.export_neighbors(ctx, d).reshape(-1)).reshape(d.n, d.ngmax)
    for i in range(d.n):
        c = min(int(ref["nc"][i]) - 1, d.ngmax)
        assert np.array_equal(nb[i, :c], ref["nbrs"][i, :c]), i


def test_export_neighbors_cpu_layout(ctx):
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64

    inp, box = uniform_box(5000, seed=41)
    ref = oracle_step(inp, box)
    step, d = _run_gpu_step(ctx, inp, box, FP64)
    nb = u32(hydro.export_neighbors(ctx, d).reshape(-1)).reshape(d.n, d.ngmax)
    nc = ref["nc"]
    for i in range(d.n):
        c = min(int(nc[i]) - 1, d.ngmax)
        assert np.array_equal(nb[i, :c], ref["nbrs"][i, :c]), i


# ----------------------------------------------------------------------------- target groups
@pytest.mark.parametrize("case", ["uniform", "sedov", "clustered", "fine"])
@pytest.mark.parametrize("hdtype", ["f64", "f32", "none"])
def test_compute_groups_bit_exact(ctx, case, hdtype):
    """computeGroupSplits on the GPU against the oracle: identical boundaries.  "fine" uses
    one-particle leaves on clumped particles, so that one 64-particle group spans hundreds
    of leaves, most of them empty."""
    from sphexa_amd import hydro

    print("case groups", case, hdtype, flush=True)
    bucket = 64
    if case == "fine":
        inp, box = uniform_box(4000, seed=54, periodic=False)
        rng = np.random.default_rng(55)
        for a in ("x", "y", "z"):
            inp[a][:2000] = 0.3 + 1e-4 * rng.random(2000)
        bucket = 1
    elif case == "sedov":
        inp, box = sedov_lattice(24)
    elif case == "uniform":
        inp, box = uniform_box(30000, seed=51)
    else:
        # two dense clumps in a sparse background: large SFC jumps between them
        inp, box = uniform_box(20000, seed=52, periodic=False)
        rng = np.random.default_rng(53)
        for k, c in ((0, 0.2), (1, 0.7)):
            sl = slice(k * 6000, (k + 1) * 6000)
            for a in ("x", "y", "z"):
                inp[a][sl] = c + 0.02 * rng.random(6000)
    keys, s, t = _gpu_tree(ctx, inp, box, bucket)
    rt = CPU.tree(keys, box, bucket)
    n = len(s["x"])
    if case == "fine":
        spans = np.searchsorted(rt["layout"][:-1], np.arange(63, n, 64), side="right") - \
            np.searchsorted(rt["layout"][:-1], np.arange(0, n - 63, 64), side="right")
        assert spans.max() > 128  # the leaf walk takes several steps
    if hdtype == "none":
        h_dev, h_ref = None, None
    elif hdtype == "f32":
        h_dev = dev(s["h"], torch.float32)
        h_ref = s["h"].astype(np.float32).astype(np.float64)
    else:
        h_dev, h_ref = dev(s["h"]), s["h"]
    for first, last in ((0, n), (37, n - 101)):
        ref = CPU.compute_groups(s["x"], s["y"], s["z"], h_ref, rt["leaves"], rt["layout"], box, 2.0, first, last)
        g = hydro.compute_groups(ctx, dev(s["x"]), dev(s["y"]), dev(s["z"]), h_dev, t, hbox(box), 2.0, first, last)
        got = u32(g.bounds)
        assert g.numGroups == len(ref) - 1
        assert np.array_equal(got, ref)
        sizes = np.diff(got.astype(np.int64))
        assert sizes.min() >= 1 and sizes.max() <= 64


@pytest.mark.parametrize("case", ["uniform_periodic", "sedov"])
def test_hydro_step_spatial_groups_equal_fixed(ctx, case):
    """Results are per particle: split groups give the same bits as fixed 64-particle groups
    in the FP64-exact loops, and the same neighbor counts."""
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64

    print("case spatial groups", case, flush=True)
    inp, box = sedov_lattice(24) if case == "sedov" else uniform_box(20000, seed=61)
    out = {}
    for spatial in (False, True):
        ctx.set_fp64_exact(True)
        step = hydro.HydroStep(ctx, len(inp["x"]), hbox(box), FP64, spatial_groups=spatial)
        d = step.compute_forces({k: dev(v) for k, v in inp.items()})
        torch.cuda.synchronize()
        ctx.set_fp64_exact(False)
        out[spatial] = {f: getattr(d, f).cpu().numpy().copy() for f in FIELDS + ("nc", "h")}
        if spatial:
            assert step.groups is not None and step.groups.numGroups >= (len(inp["x"]) + 63) // 64
    for f in FIELDS + ("nc", "h"):
        assert np.array_equal(out[True][f], out[False][f]), f


@pytest.mark.parametrize("mode", ["mixed", "fp64", "fp64_exact"])
def test_hydro_std_one_call_equals_four_calls(ctx, mode):
    """sphx_compute_hydro_std (each loop writing the next loop's neighbor records) gives the
    bits of the four calls density, EOS, IAD, momentum-energy (each packing its records)."""
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64, MIXED

    print("case hydro_std one call", mode, flush=True)
    inp, box = uniform_box(30000, seed=63)
    prec = MIXED if mode == "mixed" else FP64
    tin = {k: (dev(v) if (k in ("x", "y", "z") or prec == FP64) else dev(v).float()) for k, v in inp.items()}
    ctx.set_fp64_exact(mode == "fp64_exact")
    try:
        step = hydro.HydroStep(ctx, len(inp["x"]), hbox(box), prec)
        out = {}
        for timed in (True, False):  # timed: the four calls (per-stage events); else the one call
            d = step.compute_forces(tin, timed=timed)
            torch.cuda.synchronize()
            out[timed] = {f: getattr(d, f).cpu().numpy().copy() for f in FIELDS + ("p", "c")}
            out[timed]["dt"] = d.minDtCourant
    finally:
        ctx.set_fp64_exact(False)
    for f in FIELDS + ("p", "c"):
        assert np.array_equal(out[True][f], out[False][f]), f
    assert out[True]["dt"] == out[False]["dt"]


def test_min_dt_deferred(ctx):
    """With the time step deferred (the Python default) the momentum-energy calls return
    without reading minDtCourant: the sphx_particles view holds NaN, sphx_min_dt_courant
    waits for the value, and ParticlesData.minDtCourant reads it when first used.  Values
    equal those read back at once; a pending value is read into its ParticlesData before a
    call on another one replaces it, and when the option is switched off."""
    import ctypes as C
    from sphexa_amd import hydro
    from sphexa_amd._lib import MIXED

    print("case deferred minDtCourant", flush=True)
    inp, box = uniform_box(20000, seed=65)
    inp2, _ = uniform_box(20000, seed=66)
    tin = {k: (dev(v) if k in ("x", "y", "z") else dev(v).float()) for k, v in inp.items()}
    tin2 = {k: (dev(v) if k in ("x", "y", "z") else dev(v).float()) for k, v in inp2.items()}
    want = {}
    try:
        ctx.defer_min_dt(False)
        for name, t in (("a", tin), ("b", tin2)):
            step = hydro.HydroStep(ctx, len(t["x"]), hbox(box), MIXED)
            d = step.compute_forces(t)
            assert d._min_dt_ctx is None
            want[name] = d.minDtCourant
        assert np.isfinite(want["a"]) and want["a"] > 0 and want["a"] != want["b"]

        ctx.defer_min_dt(True)
        sa = hydro.HydroStep(ctx, len(tin["x"]), hbox(box), MIXED)
        sb = hydro.HydroStep(ctx, len(tin2["x"]), hbox(box), MIXED)
        da = sa.compute_forces(tin)
        assert da._min_dt_ctx is ctx  # pending
        assert np.isnan(da.struct().minDtCourant)
        db = sb.compute_forces(tin2)  # reads da's value first
        assert da._min_dt_ctx is None and da.minDtCourant == want["a"]
        assert db._min_dt_ctx is ctx
        v = C.c_double()
        assert ctx.lib.sphx_min_dt_courant(ctx.h, C.byref(v)) == 0 and v.value == want["b"]
        assert db.minDtCourant == want["b"] and db._min_dt_ctx is None
        for _ in range(2):  # repeated calls on one ParticlesData: the latest value
            da = sa.compute_forces(tin)
        ctx.defer_min_dt(False)  # reads the pending value
        assert da._min_dt_ctx is None and da.minDtCourant == want["a"]
    finally:
        ctx.defer_min_dt(True)


def test_hydro_std_one_call_sub_range(ctx):
    """Groups over a sub-range [first, last) (the particles outside are neighbors only): the
    one call packs the records as the four calls do and gives their bits on [first, last)."""
    from sphexa_amd import hydro
    from sphexa_amd._lib import MIXED

    print("case hydro_std one call sub-range", flush=True)
    inp, box = uniform_box(30000, seed=64)
    n = len(inp["x"])
    tin = {k: (dev(v) if k in ("x", "y", "z") else dev(v).float()) for k, v in inp.items()}
    step = hydro.HydroStep(ctx, n, hbox(box), MIXED)
    d = step.compute_forces(tin)  # sorted fields, tree, and every field set once
    first, last = 1000, n - 700
    g = hydro.compute_groups(ctx, d.x, d.y, d.z, d.h, step.octree, step.box, first=first, last=last)
    hydro.find_neighbors_sfc(ctx, d, step.box, step.octree, g)
    out = {}
    for one in (False, True):
        if one:
            hydro.compute_hydro_std(ctx, d, step.box, g)
        else:
            hydro.compute_density(ctx, d, step.box, g)
            hydro.compute_eos_hydro_std(ctx, d, first, last)
            hydro.compute_iad(ctx, d, step.box, g)
            hydro.compute_momentum_energy_std(ctx, d, step.box, g)
        torch.cuda.synchronize()
        out[one] = {f: getattr(d, f).cpu().numpy()[first:last].copy() for f in FIELDS + ("p", "c")}
        out[one]["dt"] = d.minDtCourant
    for f in FIELDS + ("p", "c"):
        assert np.array_equal(out[True][f], out[False][f]), f
    assert out[True]["dt"] == out[False]["dt"]


def test_timing_modes(ctx):
    """Kernel timing: every hot kernel (True), the heavy launches only ("loops": neighbor
    candidates and particle loops), or none; an out-of-range mode is an argument error."""
    from sphexa_amd import _lib, hydro
    from sphexa_amd._lib import MIXED

    print("case timing modes", flush=True)
    inp, box = uniform_box(20000, seed=62)
    step = hydro.HydroStep(ctx, len(inp["x"]), hbox(box), MIXED)
    inp = {k: (dev(v) if k in ("x", "y", "z") else dev(v).float()) for k, v in inp.items()}
    seen = {}
    for mode in (True, "loops", False):
        ctx.set_timing(mode)
        ctx.timing_summary()
        step.compute_forces(inp)
        seen[mode] = set(ctx.timing_summary())
    ctx.set_timing(False)
    heavy = {"nsCandidates", "densityKernel", "iadKernel", "momentumEnergyKernel"}
    assert seen["loops"] == heavy
    assert heavy < seen[True] and "radixScatter" in seen[True]
    assert seen[False] == set()
    assert ctx.lib.sphx_ctx_set_option(ctx.h, _lib.OPT_TIMING, 3) == _lib.SPHX_ERR_ARG


def test_compute_groups_reference_find_splits_known_answer(ctx):
    """The reference's findSplits known answer on the GPU (see tests/test_oracle.py)."""
    from sphexa_amd import hydro
    from tests.test_oracle import _find_splits_case

    print("case groups find_splits", flush=True)
    p, h, tol = _find_splits_case()
    leaves = dev(np.array([0, 1 << 63], dtype=np.uint64).view(np.int64))
    layout = dev(np.array([0, 64], dtype=np.int32))
    t = hydro.build_octree(ctx, leaves)
    t.layout = layout
    g = hydro.compute_groups(ctx, dev(p), dev(p), dev(p), dev(h), t, hydro.make_box(0.0, 1.0, 0), tol)
    assert u32(g.bounds).tolist() == [0, 1, 3, 32, 34, 64]


@pytest.mark.parametrize("bc", ["open", "mixed", "periodic"])
def test_make_global_box_bit_exact(ctx, bc):
    from sphexa_amd import hydro

    print("case global box", bc, flush=True)
    rng = np.random.default_rng(121)
    n = 300_001
    x, y, z = rng.normal(0, 3, n), rng.random(n) * 2 - 1, rng.standard_cauchy(n)
    bcs = {"open": [oracle.BC_OPEN] * 3, "mixed": [oracle.BC_OPEN, oracle.BC_PERIODIC, oracle.BC_FIXED],
           "periodic": [oracle.BC_PERIODIC] * 3}[bc]
    prev = oracle.Box.make([-1.0, -1.0, -1.0], [1.0, 1.0, 1.0], bcs)
    ref = CPU.make_global_box(x, y, z, prev)
    got = hydro.make_global_box(ctx, dev(x), dev(y), dev(z), hbox(prev))
    for f in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "bx", "by", "bz"):
        assert getattr(got, f) == getattr(ref, f), f
