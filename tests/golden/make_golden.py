"""Write the golden fixtures of the hot path from the compiled reference CPU path.

Runs oracle/_ref/libsphexa_ref.so (the reference's own templates, compiled from its
sources by oracle/Makefile) on deterministic inputs and stores inputs' generator
parameters and the outputs as compressed .npz files next to this script:

  sedov30.npz    Sedov lattice 30^3 (27,000 particles), periodic [-0.5, 0.5]^3, FP64
  uniform4k.npz  4,096 uniform random particles, periodic unit box, seed 7, FP64
  open3k.npz     3,000 uniform random particles, open box [-1, 2]^3, seed 11, FP64
  uniform1m.npz  1,000,000 uniform random particles, periodic unit box
                 (sphexa_amd.workloads.uniform_random, seed 102: h as written, ~800
                 neighbors), findNeighborsSph with ng0 100 and ngmax 1024: the SHA-256 of
                 the sorted keys, of nc and of the lists (each particle's stored neighbors,
                 nc - 1 of them, u32 little-endian, particles in order), and 64 full rows

Each file holds, for keys, sort order, leaves, counts, nc, h after the h iteration,
rho, p, c, c11..c33, ax, ay, az, du and minDt: the SHA-256 of the full array
(sha_<name>) and either the full array (small cases) or the rows at 1,500 fixed
sample indices (sample_idx, <name>).  Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.helpers import lists_digest, oracle_step, sedov_lattice, uniform_box  # noqa: E402

CASES = {
    "sedov30": (lambda: sedov_lattice(30), True),
    "uniform4k": (lambda: uniform_box(4096, seed=7, periodic=True), False),
    "open3k": (lambda: uniform_box(3000, seed=11, periodic=False, lo=-1.0, hi=2.0), False),
}
FIELDS = ("keys", "order", "nc", "h", "rho", "p", "c", "c11", "c12", "c13", "c22", "c23", "c33", "ax", "ay", "az",
          "du", "leaves", "counts")


def sha(a):
    import hashlib

    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8)


def uniform1m(n=1_000_000, seed=102, ng0=100, ngmax=1024, chunk=100_000, samples=64):
    """The uniform-random configuration at full size through the compiled reference."""
    import oracle
    from sphexa_amd import workloads

    ref = oracle.Cpu("ref")
    f, _ = workloads.uniform_random(n, seed=seed)
    box = oracle.Box.cube(0.0, 1.0, oracle.BC_PERIODIC)
    keys = ref.sfc_keys(f["x"], f["y"], f["z"], box)
    keys, order = ref.sort_by_key(keys, np.arange(n, dtype=np.uint32))
    s = {k: np.ascontiguousarray(f[k][order]) for k in ("x", "y", "z", "h")}
    t = ref.tree(keys, box)
    idx = np.sort(np.random.default_rng(seed).choice(n, samples, replace=False))
    rows = np.full((samples, ngmax), 0xFFFFFFFF, np.uint32)
    ncs = []

    def chunks():
        for c0 in range(0, n, chunk):
            c1 = min(n, c0 + chunk)
            nb, nc, h2 = ref.find_neighbors_sph(s["x"], s["y"], s["z"], s["h"], box, t, ng0, ngmax, c0, c1)
            assert np.array_equal(h2, s["h"]), "the h iteration fired"
            assert nc.max() - 1 <= ngmax
            for r, i in enumerate(idx):
                if c0 <= i < c1:
                    rows[r, :nc[i - c0] - 1] = nb[i - c0, :nc[i - c0] - 1]
            ncs.append(nc)
            yield nb, nc

    digest = lists_digest(chunks())
    nc = np.concatenate(ncs)
    return {"sha_keys": sha(keys), "sha_nc": sha(nc), "sha_lists": digest, "sample_idx": idx, "sample_rows": rows,
            "sample_nc": nc[idx], "n": np.array([n]), "seed": np.array([seed]), "ng0": np.array([ng0]),
            "ngmax": np.array([ngmax]), "nc_stats": np.array([nc.min(), nc.mean(), nc.max()])}


def main():
    path = os.path.join(HERE, "uniform1m.npz")
    np.savez_compressed(path, **uniform1m())
    print("uniform1m", os.path.getsize(path), "bytes")
    for name, (make, sampled) in CASES.items():
        inp, box = make()
        r = oracle_step(inp, box, kind="ref")
        full = {k: r[k] for k in FIELDS[:17]}
        full["leaves"] = r["tree"]["leaves"]
        full["counts"] = r["tree"]["counts"]
        out = {"sha_" + k: sha(v) for k, v in full.items()}
        out["minDt"] = np.array([r["minDt"]])
        if sampled:
            idx = np.linspace(0, len(r["rho"]) - 1, 1500).astype(np.int64)
            out["sample_idx"] = idx
            for k in FIELDS[:17]:
                out[k] = full[k][idx]
            out["leaves"], out["counts"] = full["leaves"], full["counts"]
        else:
            out.update(full)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(name, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
