"""The C++ drop-in at run time: tests/compat_driver.cpp (built by build()) runs one std-hydro
time step through include/sphx_dropin.hpp, calling the reference's entry points by their own
names and argument lists (cstone::computeSfcKeys<true>, sortByKey<true>, gatherAcc<true>,
updateOctreeGpu, buildOctreeGpu, nodeFpCenters, sph::computeGroups, findNeighborsSfc, the
loops, computeTimestep, computePositions, updateSmoothingLength) with its own device vectors,
box, octree views, group data and dataset, the way Domain::sync and HydroProp::computeForces
and ::integrate call them; then the CPU-signature entry points on host memory
(computeOctree, updateInternalTree, nodeFpCenters, findNeighbors).
Its results are checked against the oracle on the same input:
* sorted positions, neighbor counts: bit-exact;
* rho <= 1e-13 relative; p, c <= 1e-13 relative; ax, ay, az, du <= 1e-10 of the field's
  largest magnitude (FP64 with FMA contraction, a different summation order);
* Courant and new time step, positions, velocities, u and h after integrate <= 1e-10 relative
  to the field's largest magnitude (they inherit the accelerations' rounding);
* computeOctree on host keys equals updateOctreeGpu's tree; the plain findNeighbors counts over
  host copies of the tree, and the lists of the first 64 particles: bit-exact.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from tests.helpers import oracle_step, rel_err, sedov_lattice, uniform_box

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "_bin", "compat_driver")
CPU = oracle.Cpu("oracle")
NAMES = ("x", "y", "z", "h", "m", "vx", "vy", "vz", "u")


def run_driver(inp, box):
    n = len(inp["x"])
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            np.array([n], np.uint64).tofile(f)
            for k in NAMES:
                np.ascontiguousarray(inp[k], np.float64).tofile(f)
            for lo, hi, bc in ((box.xmin, box.xmax, box.bx), (box.ymin, box.ymax, box.by),
                               (box.zmin, box.zmax, box.bz)):
                np.array([lo, hi], np.float64).tofile(f)
                np.array([bc], np.int32).tofile(f)
        r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(fout, "rb").read()
    out, o = {}, 0
    for k in ("x", "y", "z", "h", "rho", "p", "c", "ax", "ay", "az", "du"):
        out[k] = np.frombuffer(raw, np.float64, n, o)
        o += 8 * n
    out["nc"] = np.frombuffer(raw, np.uint32, n, o)
    o += 4 * n
    out["minDtCourant"], out["minDt"] = np.frombuffer(raw, np.float64, 2, o)
    o += 16
    for k in ("x1", "y1", "z1", "vx1", "u1", "h1"):
        out[k] = np.frombuffer(raw, np.float64, n, o)
        o += 8 * n
    out["same_tree"] = int(np.frombuffer(raw, np.int32, 1, o)[0])
    o += 4
    out["plain_nc"] = np.frombuffer(raw, np.uint32, n, o)
    o += 4 * n
    rows = min(64, n)
    out["plain_nb"] = np.frombuffer(raw, np.uint32, rows * 150, o).reshape(rows, 150)
    return out


def scale(g, w):
    s = np.max(np.abs(w))
    return float(np.max(np.abs(g - w)) / (s if s > 0 else 1.0))


@pytest.mark.parametrize("case", ["uniform", "sedov"])
def test_compat_driver_time_step(case):
    assert os.path.exists(DRIVER), "tests/_bin/compat_driver not built: run __graft_entry__.build()"
    print("case compat driver", case, flush=True)
    inp, box = uniform_box(4096, seed=51) if case == "uniform" else sedov_lattice(16)
    got = run_driver(inp, box)
    want = oracle_step(inp, box)
    s = want["sorted"]
    for k in ("x", "y", "z"):
        assert np.array_equal(got[k], s[k]), k
    assert np.array_equal(got["nc"], want["nc"])
    assert np.array_equal(got["h"], want["h"])
    assert rel_err(got["rho"], want["rho"]) <= 1e-13
    assert rel_err(got["p"], want["p"]) <= 1e-13 and rel_err(got["c"], want["c"]) <= 1e-13
    for k in ("ax", "ay", "az", "du"):
        assert scale(got[k], want[k]) <= 1e-10, k
    assert got["minDtCourant"] == pytest.approx(want["minDt"], rel=1e-12)
    # integrate: computeTimestep from minDt = minDt_m1 = 1e-6, then positions and h
    ts, _ = CPU.compute_timestep([1e-6, 1e-6, 0.0], min_dt_courant=want["minDt"])
    assert got["minDt"] == pytest.approx(ts[0], rel=1e-12)
    st = dict(s, h=want["h"], ax=want["ax"], ay=want["ay"], az=want["az"], du=want["du"])
    for k in ("x_m1", "y_m1", "z_m1", "du_m1"):
        st[k] = np.zeros(len(s["x"]))
    st = {k: st[k] for k in ("x", "y", "z", "x_m1", "y_m1", "z_m1", "vx", "vy", "vz", "ax", "ay", "az", "h", "u",
                             "du", "du_m1")}
    w1 = CPU.compute_positions(st, ts[0], ts[1], box)
    h1, _, _ = CPU.update_smoothing_length(100, want["nc"], want["h"])
    for g, w in (("x1", "x"), ("y1", "y"), ("z1", "z"), ("vx1", "vx"), ("u1", "u")):
        assert scale(got[g], w1[w]) <= 1e-10, g
    assert rel_err(got["h1"], h1) <= 1e-12
    # the CPU-signature entry points on host memory
    assert got["same_tree"] == 1
    pnb, pnc = CPU.find_neighbors(s["x"], s["y"], s["z"], want["h"], box, want["tree"], 150)
    assert np.array_equal(got["plain_nc"], pnc)
    for i in range(64):
        c = min(int(pnc[i]), 150)
        assert np.array_equal(got["plain_nb"][i, :c], pnb[i, :c]), i
