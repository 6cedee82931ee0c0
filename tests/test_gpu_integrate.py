"""GPU parity of the integration step (computeTimestep, computePositions, updateSmoothingLength)
and of repeated time steps, against the oracle on the same inputs.

Tolerances:
* positions, x_m1, velocities, u and du_m1 in FP64: bit-exact (the kernel rounds every
  operation like the reference CPU loop), except the two places where a device libm
  function replaces glibc: exp in the u < 0 branch and tanh in the fixed-boundary factor,
  <= 1e-14 of the field's largest magnitude there (a 1-ulp factor difference, amplified
  where the velocity update cancels);
* the acceleration time step: bit-exact (the minimum is exact, the root is taken on the host);
* updateH: device pow against glibc pow, <= 4.5e-16 relative (2 ulp); removeKey marks bit-exact;
* MIXED (FP32 state fields): FP64 arithmetic on the FP32 values, so the FP64 oracle's
  result on the same values, rounded to FP32, bit for bit (positions FP64, bit for bit);
* three full time steps of a Sedov lattice (FP64-exact loops): x, v, u, h within 1e-12
  relative of the oracle's three steps (h differs by device pow at 1 ulp, which can move a
  smoothing length across a table node; the loops themselves are bit-identical).
"""
import numpy as np
import pytest
import torch

import oracle
from tests.helpers import oracle_step, rel_err, sedov_lattice

pytestmark = pytest.mark.gpu

CPU = oracle.Cpu("oracle")
FIELDS = ("x", "y", "z", "x_m1", "y_m1", "z_m1", "vx", "vy", "vz", "ax", "ay", "az", "h", "u", "du", "du_m1")


@pytest.fixture(scope="module")
def ctx():
    from sphexa_amd import hydro

    c = hydro.Context(0)
    yield c
    c.close()


def state(n, seed=0, lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    st = {k: lo + (hi - lo) * rng.random(n) for k in ("x", "y", "z")}
    for k in ("x_m1", "y_m1", "z_m1"):
        st[k] = rng.normal(0, 2e-3, n)
    for k in ("vx", "vy", "vz", "ax", "ay", "az", "du", "du_m1"):
        st[k] = rng.normal(0, 1.0, n)
    st["h"] = np.full(n, 0.02) * (1 + 0.5 * rng.random(n))
    st["u"] = 0.5 + rng.random(n)
    return st


def to_dataset(st, precision):
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64

    n = len(st["x"])
    d = hydro.ParticlesData(n, precision)
    dt = d.dtype
    for k in FIELDS:
        t = torch.from_numpy(np.ascontiguousarray(st[k])).to("cuda")
        setattr(d, k, t if k in ("x", "y", "z") else t.to(dt))
    d.nc = torch.zeros(n, dtype=torch.int32, device="cuda")
    return d


def from_dataset(d):
    return {k: getattr(d, k).double().cpu().numpy() for k in FIELDS}


def hbox(b):
    from sphexa_amd import hydro

    return hydro.make_box([b.xmin, b.ymin, b.zmin], [b.xmax, b.ymax, b.zmax], [b.bx, b.by, b.bz])


@pytest.mark.parametrize("bc", [oracle.BC_PERIODIC, oracle.BC_OPEN, oracle.BC_FIXED])
def test_positions_fp64(ctx, bc):
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64

    print("case positions fp64 bc", bc, flush=True)
    n = 100_003
    st = state(n, seed=21)
    st["x"][:300] = np.linspace(0.0, 0.1, 300)  # near the low x wall
    st["du"][300:400] = -1e4                     # u < 0 branch
    st["u"][300:400] = 1.0
    box = oracle.Box.cube(0.0, 1.0, bc)
    want = CPU.compute_positions(st, 2e-3, 1.7e-3, box)
    d = to_dataset(st, FP64)
    ts = hydro.TimestepState(minDt=2e-3, minDt_m1=1.7e-3)
    hydro.compute_positions(ctx, d, hbox(box), ts)
    got = from_dataset(d)
    exact = np.ones(n, bool)
    exact[300:400] = False  # device exp
    if bc == oracle.BC_FIXED:  # device tanh within 5 h of a wall, before or after the move
        for X in (st, want):
            for k in ("x", "y", "z"):
                exact &= np.minimum(np.abs(X[k]), np.abs(1.0 - X[k])) / st["h"] >= 5.0
    assert exact.sum() > n // 3
    for k in FIELDS:
        assert np.array_equal(got[k][exact], want[k][exact]), k
        g, w = got[k][~exact], want[k][~exact]
        assert np.max(np.abs(g - w), initial=0.0) <= 1e-14 * np.max(np.abs(want[k])), k


def test_positions_mixed(ctx):
    from sphexa_amd import hydro
    from sphexa_amd._lib import MIXED

    print("case positions mixed", flush=True)
    n = 65_536
    st = state(n, seed=22)
    # the FP32 fields hold FP32 values on both sides
    for k in FIELDS:
        if k not in ("x", "y", "z"):
            st[k] = st[k].astype(np.float32).astype(np.float64)
    box = oracle.Box.cube(0.0, 1.0, oracle.BC_PERIODIC)
    want = CPU.compute_positions(st, 2e-3, 1.7e-3, box)
    d = to_dataset(st, MIXED)
    hydro.compute_positions(ctx, d, hbox(box), hydro.TimestepState(minDt=2e-3, minDt_m1=1.7e-3))
    got = from_dataset(d)
    # FP64 arithmetic on the FP32 values, stored to FP32: the oracle's result rounded
    for k in ("x", "y", "z"):
        assert np.array_equal(got[k], want[k]), k
    for k in ("x_m1", "y_m1", "z_m1", "vx", "vy", "vz", "u", "du_m1"):
        assert np.array_equal(got[k], want[k].astype(np.float32).astype(np.float64)), k


@pytest.mark.parametrize("precision", [0, 1])
def test_timestep_acceleration_limit(ctx, precision):
    from sphexa_amd import hydro

    print("case timestep precision", precision, flush=True)
    rng = np.random.default_rng(23)
    n = 1_000_001
    st = state(n, seed=23)
    st["ax"] = rng.normal(0, 10, n)
    st["ax"][7] = 0.0  # zero acceleration: h^2 / 0 = inf
    d = to_dataset(st, precision)
    dd = from_dataset(d)  # the field values the kernel sees
    d.minDtCourant = 4e-3
    ts = hydro.TimestepState(minDt=1e-3, minDt_m1=9e-4, ttot=0.25, g=1.0, etaAcc=0.2)
    hydro.compute_timestep(ctx, d, ts)
    if precision == 0:
        want_min = CPU.min_h2_a2(dd["h"], dd["ax"], dd["ay"], dd["az"])
    else:  # h^2 in FP32, as the reference's HydroType product
        hh = (dd["h"].astype(np.float32) * dd["h"].astype(np.float32)).astype(np.float64)
        want_min = np.min(hh / (dd["ax"] ** 2 + (dd["ay"] ** 2 + dd["az"] ** 2)))
    want_acc = 0.2 * want_min ** 0.25
    assert ts.minDtAcc == want_acc
    assert ts.minDt == min(want_acc, 4e-3, 1.1e-3)
    assert ts.minDt_m1 == 1e-3 and ts.ttot == 0.25 + ts.minDt
    # g == 0: no acceleration limit, nothing launched
    ts2 = hydro.TimestepState(minDt=1e-3, minDt_m1=9e-4)
    hydro.compute_timestep(ctx, d, ts2)
    assert ts2.minDt == 1.1e-3 and ts2.minDtAcc == np.inf


@pytest.mark.parametrize("precision", [0, 1])
def test_update_smoothing_length(ctx, precision):
    from sphexa_amd import hydro

    print("case update h precision", precision, flush=True)
    rng = np.random.default_rng(24)
    n = 200_000
    h = 0.01 + rng.random(n) * 0.02
    nc = rng.integers(2, 400, n).astype(np.uint32)
    keys = rng.integers(0, 2 ** 62, n).astype(np.uint64)
    d = hydro.ParticlesData(n, precision)
    d.h = torch.from_numpy(h).to("cuda").to(d.dtype)
    hd = d.h.double().cpu().numpy()
    d.nc = torch.from_numpy(nc.view(np.int32)).to("cuda")
    kd = torch.from_numpy(keys.view(np.int64)).to("cuda")
    assert not hydro.update_smoothing_length(ctx, d, kd)
    want, _, _ = CPU.update_smoothing_length(100, nc, hd)
    got = d.h.double().cpu().numpy()
    tol = 4.5e-16 if precision == 0 else 6e-8
    assert rel_err(got, want if precision == 0 else want.astype(np.float32)) <= tol
    assert np.array_equal(kd.cpu().numpy().view(np.uint64), keys)
    nc[[5, 99_999]] = [1, 0]
    d.nc = torch.from_numpy(nc.view(np.int32)).to("cuda")
    assert hydro.update_smoothing_length(ctx, d, kd)
    k = kd.cpu().numpy().view(np.uint64)
    assert k[5] == k[99_999] == 2 ** 63 and np.array_equal(np.delete(k, [5, 99_999]), np.delete(keys, [5, 99_999]))


def oracle_time_steps(inp, box, steps, ts):
    """The reference's propagator on the CPU: sync (sort), computeForces, integrate."""
    st = {k: v.copy() for k, v in inp.items()}
    for _ in range(steps):
        r = oracle_step(st, box)
        s = {k: v.copy() for k, v in r["sorted"].items()}
        s.update(h=r["h"], ax=r["ax"], ay=r["ay"], az=r["az"], du=r["du"])
        t, _ = CPU.compute_timestep(ts, min_dt_courant=r["minDt"])
        ts = list(t)
        s = CPU.compute_positions(s, ts[0], ts[1], box)
        s["h"], _, removed = CPU.update_smoothing_length(100, r["nc"], s["h"])
        assert not removed
        st = {k: s[k] for k in ("x", "y", "z", "h", "m", "vx", "vy", "vz", "u", "x_m1", "y_m1", "z_m1", "du_m1")}
    return st, ts


def test_sedov_time_steps_fp64_exact(ctx):
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64

    print("case sedov 20^3 three time steps", flush=True)
    inp, box = sedov_lattice(20)
    n = len(inp["x"])
    for k in ("x_m1", "y_m1", "z_m1", "du_m1"):
        inp[k] = np.zeros(n)
    ts0 = [1e-6, 1e-6, 0.0]
    want, ts_want = oracle_time_steps(inp, box, 3, ts0)

    ctx.set_fp64_exact(True)
    try:
        state = {k: torch.from_numpy(v).to("cuda") for k, v in inp.items()}
        sim = hydro.Simulation(ctx, state, hbox(box), FP64,
                               ts=hydro.TimestepState(minDt=ts0[0], minDt_m1=ts0[1], ttot=ts0[2]))
        for _ in range(3):
            sim.step()
        torch.cuda.synchronize()
    finally:
        ctx.set_fp64_exact(False)
    assert sim.ts.minDt == pytest.approx(ts_want[0], rel=1e-12)
    assert sim.ts.ttot == pytest.approx(ts_want[2], rel=1e-12)
    got = {k: v.double().cpu().numpy() for k, v in sim.state.items()}
    for k in ("x", "y", "z", "h", "u", "x_m1", "du_m1"):
        assert rel_err(got[k], want[k], floor=1e-12 * np.max(np.abs(want[k]))) <= 1e-12, k
    for k in ("vx", "vy", "vz"):
        assert np.max(np.abs(got[k] - want[k])) <= 1e-12 * np.max(np.abs(want[k])), k
    assert np.max(np.abs(want["vx"])) > 0  # the blast has started to move particles


def oracle_ve_time_steps(inp, box, steps, ts, alpha):
    """The reference's VE propagator on the CPU: sync (sort), computeForces (AV switches over
    the current minDt, rhoTimestep), integrate."""
    from tests.helpers import oracle_ve_step

    st = {k: v.copy() for k, v in inp.items()}
    st["alpha"] = alpha.copy()
    for _ in range(steps):
        a = st.pop("alpha")
        r = oracle_ve_step(st, box, a, dt=ts[0])
        s = {k: np.ascontiguousarray(v[r["order"]]) for k, v in st.items()}
        s.update(h=r["h"], ax=r["ax"], ay=r["ay"], az=r["az"], du=r["du"])
        min_dt_rho = 0.06 / abs(np.max(r["divv"]))
        t, _ = CPU.compute_timestep(ts, min_dt_courant=r["minDt"], min_dt_rho=min_dt_rho)
        ts = list(t)
        s = CPU.compute_positions(s, ts[0], ts[1], box)
        s["h"], _, removed = CPU.update_smoothing_length(100, r["nc"], s["h"])
        assert not removed
        st = {k: s[k] for k in ("x", "y", "z", "h", "m", "vx", "vy", "vz", "u", "x_m1", "y_m1", "z_m1", "du_m1")}
        st["alpha"] = r["alpha"]
    return st, ts


def test_ve_time_steps_fp64_exact(ctx):
    """Three VE time steps (HydroVeProp: computeForces + integrate) against the oracle's; the
    loops are bit-identical in FP64-exact mode, h differs by device pow (1 ulp): 1e-12."""
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64

    print("case ve sedov 16^3 three time steps", flush=True)
    inp, box = sedov_lattice(16)
    n = len(inp["x"])
    inp["vx"] = 0.1 * np.sin(2 * np.pi * inp["x"])
    for k in ("x_m1", "y_m1", "z_m1", "du_m1"):
        inp[k] = np.zeros(n)
    alpha = np.full(n, 0.5)
    ts0 = [1e-6, 1e-6, 0.0]
    want, ts_want = oracle_ve_time_steps(inp, box, 3, ts0, alpha)

    ctx.set_fp64_exact(True)
    try:
        state = {k: torch.from_numpy(v).to("cuda") for k, v in inp.items()}
        state["alpha"] = torch.from_numpy(alpha).to("cuda")
        sim = hydro.Simulation(ctx, state, hbox(box), FP64, propagator="ve",
                               ts=hydro.TimestepState(minDt=ts0[0], minDt_m1=ts0[1], ttot=ts0[2]))
        # the derivative table is an input: the oracle's, as in the VE parity tests
        sim.step_obj.ve.whd = torch.from_numpy(CPU.kernel_derivative_table()).to("cuda")
        for _ in range(3):
            sim.step()
        torch.cuda.synchronize()
    finally:
        ctx.set_fp64_exact(False)
    assert sim.ts.minDt == pytest.approx(ts_want[0], rel=1e-12)
    assert sim.ts.ttot == pytest.approx(ts_want[2], rel=1e-12)
    got = {k: v.double().cpu().numpy() for k, v in sim.state.items()}
    for k in ("x", "y", "z", "h", "u", "x_m1", "du_m1", "alpha"):
        assert rel_err(got[k], want[k], floor=1e-12 * np.max(np.abs(want[k]))) <= 1e-12, k
    for k in ("vx", "vy", "vz"):
        assert np.max(np.abs(got[k] - want[k])) <= 1e-12 * np.max(np.abs(want[k])), k


def test_integration_edge_sizes(ctx):
    """One particle and an empty range: the kernels handle ragged and empty launches."""
    from sphexa_amd import hydro
    from sphexa_amd._lib import FP64

    print("case integration edge sizes", flush=True)
    st = state(1, seed=25)
    box = oracle.Box.cube(0.0, 1.0, oracle.BC_PERIODIC)
    want = CPU.compute_positions(st, 1e-3, 1e-3, box)
    d = to_dataset(st, FP64)
    hydro.compute_positions(ctx, d, hbox(box), hydro.TimestepState(minDt=1e-3, minDt_m1=1e-3))
    got = from_dataset(d)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k
    ts = hydro.TimestepState(minDt=1e-3, minDt_m1=1e-3, g=1.0)
    hydro.compute_timestep(ctx, d, ts, first=0, last=0)  # empty range: no acceleration limit
    assert ts.minDtAcc == np.inf and ts.minDt == 1.1e-3
    e = hydro.ParticlesData(0, FP64)
    e.h = torch.zeros(0, dtype=torch.float64, device="cuda")
    e.nc = torch.zeros(0, dtype=torch.int32, device="cuda")
    assert not hydro.update_smoothing_length(ctx, e)
