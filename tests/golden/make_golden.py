"""Write the golden fixtures of the hot path from the compiled reference CPU path.

Runs oracle/_ref/libsphexa_ref.so (the reference's own templates, compiled from its
sources by oracle/Makefile) on deterministic inputs and stores inputs' generator
parameters and the outputs as compressed .npz files next to this script:

  sedov30.npz    Sedov lattice 30^3 (27,000 particles), periodic [-0.5, 0.5]^3, FP64
  uniform4k.npz  4,096 uniform random particles, periodic unit box, seed 7, FP64
  open3k.npz     3,000 uniform random particles, open box [-1, 2]^3, seed 11, FP64

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

from tests.helpers import oracle_step, sedov_lattice, uniform_box  # noqa: E402

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


def main():
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
