"""Pins of the oracle's integration step (computeTimestep, computePositions, updateSmoothingLength).

Each check is one the mathematics or the reference fixes independently of the oracle's code:
* positionUpdate reproduces constant-acceleration kinematics exactly (any dt, dt_m1) and is
  time-reversible (positions.hpp:86-99 documents both);
* energyUpdate is the 2nd-order Adams-Bashforth step of a linearly extrapolated du, with
  the exponential branch for u < 0;
* putInBox folds once by the box length in periodic dimensions only;
* fixed-boundary factors are 1 beyond 5 h and scale motion near the wall;
* computeTimestep takes the minimum of its candidate limits; updateH leaves h unchanged at
  nc == ng0 (1024^0.1 = 2) and marks nc <= 1 with removeKey;
* equality, bit for bit, with the compiled reference CPU templates (oracle/_ref) where built.
"""
import numpy as np
import pytest

import oracle

CPU = oracle.Cpu("oracle")
REF = oracle.Cpu("ref") if oracle.available("ref") else None
need_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (reference sources absent)")

FIELDS = ("x", "y", "z", "x_m1", "y_m1", "z_m1", "vx", "vy", "vz", "ax", "ay", "az", "h", "u", "du", "du_m1")


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


def test_constant_acceleration_kinematics():
    # x(t) = x0 + v0 t + a t^2 / 2 sampled at t_n - dt_m1, t_n: one update lands on x(t_n + dt)
    rng = np.random.default_rng(4)
    n = 500
    x0, v0, a = (rng.normal(0, 1, n) for _ in range(3))
    tn = 0.37
    for dt, dt_m1 in ((1e-3, 1e-3), (2e-3, 7e-4), (5e-4, 3e-3)):
        pos = lambda t: x0 + v0 * t + 0.5 * a * t * t
        st = state(n)
        st.update(x=pos(tn), y=pos(tn), z=pos(tn), x_m1=pos(tn) - pos(tn - dt_m1), ax=a, ay=a, az=a)
        st["y_m1"], st["z_m1"] = st["x_m1"].copy(), st["x_m1"].copy()
        box = oracle.Box.cube(-1e3, 1e3, oracle.BC_OPEN)
        o = CPU.compute_positions(st, dt, dt_m1, box)
        want_x = pos(tn + dt)
        want_v = v0 + a * (tn + dt)
        assert np.max(np.abs(o["x"] - want_x)) < 1e-12
        assert np.max(np.abs(o["vx"] - want_v)) < 1e-9
        assert np.max(np.abs(o["x_m1"] - (want_x - pos(tn)))) < 1e-12


def test_time_reversal():
    # positionUpdate(-dt, dt_m1, X_n+1, A_n, dX_n) back-propagates X_n+1 to X_n
    n = 400
    st = state(n, seed=2)
    box = oracle.Box.cube(-10.0, 10.0, oracle.BC_OPEN)
    fwd = CPU.compute_positions(st, 1e-3, 8e-4, box)
    back_in = dict(st)
    back_in.update(x=fwd["x"], y=fwd["y"], z=fwd["z"])
    back = CPU.compute_positions(back_in, -1e-3, 8e-4, box)
    for k in ("x", "y", "z"):
        assert np.max(np.abs(back[k] - st[k])) < 4e-16 * 2
    # the velocity of the backward step is V_n - A dt: not V_n+1
    assert not np.allclose(back["vx"], fwd["vx"])


def test_energy_update_closed_forms():
    n = 300
    st = state(n, seed=5)
    dt, dt_m1 = 2e-3, 1.5e-3
    box = oracle.Box.cube(0.0, 1.0, oracle.BC_PERIODIC)
    o = CPU.compute_positions(st, dt, dt_m1, box)
    du, dum1, u = st["du"], st["du_m1"], st["u"]
    want = u + du * dt + 0.5 * (du - dum1) / dt_m1 * dt * dt
    assert np.max(np.abs(o["u"] - want) / np.abs(want)) < 1e-15
    assert np.array_equal(o["du_m1"], du)
    # constant du: plain Euler
    st2 = dict(st, du_m1=st["du"].copy())
    o2 = CPU.compute_positions(st2, dt, dt_m1, box)
    assert np.max(np.abs(o2["u"] - (u + du * dt))) < 1e-15
    # u_new < 0: u_old exp(u_new dt / u_old)
    st3 = dict(st, du=np.full(n, -1e4), du_m1=np.full(n, -1e4), u=np.full(n, 1.0))
    o3 = CPU.compute_positions(st3, dt, dt_m1, box)
    un = 1.0 - 1e4 * dt
    assert un < 0
    assert np.allclose(o3["u"], np.exp(un * dt), rtol=1e-15, atol=0)


def test_put_in_box():
    n = 6
    st = state(n, seed=1)
    for k in ("x_m1", "y_m1", "z_m1", "ax", "ay", "az"):
        st[k] = np.zeros(n)
    st["x"] = np.array([0.999, 0.001, 0.5, 0.999, 0.001, 0.5])
    st["x_m1"] = np.array([0.004, -0.004, 0.0, 0.004, -0.004, 0.0])  # 4e-3 per dt_m1 = dt
    dt = 1.0
    per = CPU.compute_positions(st, dt, dt, oracle.Box.cube(0.0, 1.0, oracle.BC_PERIODIC))
    opn = CPU.compute_positions(st, dt, dt, oracle.Box.cube(0.0, 1.0, oracle.BC_OPEN))
    assert np.allclose(per["x"][:2], [0.003, 0.997], atol=1e-15)
    assert np.allclose(opn["x"][:2], [1.003, -0.003], atol=1e-15)
    assert per["x"][2] == opn["x"][2] == 0.5


def test_fixed_boundary_factors():
    n = 200
    st = state(n, seed=8, lo=0.2, hi=0.8)  # every particle >= 0.2 from the walls, h <= 0.03
    fixed = oracle.Box.cube(0.0, 1.0, oracle.BC_FIXED)
    opn = oracle.Box.cube(0.0, 1.0, oracle.BC_OPEN)
    a = CPU.compute_positions(st, 1e-3, 1e-3, fixed)
    b = CPU.compute_positions(st, 1e-3, 1e-3, opn)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # on the wall the factor is 0.5 (tanh(-6) + 1) = 6.1e-6: the particle barely moves
    st2 = state(4, seed=9)
    st2["x"] = np.zeros(4)
    c = CPU.compute_positions(st2, 1e-3, 1e-3, fixed)
    f = 0.5 * (np.tanh(-6.0) + 1)
    moved = np.abs(c["x"] - st2["x"])
    free = np.abs(CPU.compute_positions(st2, 1e-3, 1e-3, opn)["x"] - st2["x"])
    assert np.allclose(moved, f * free, rtol=1e-9)


def test_timestep_rule():
    ts = [1e-3, 8e-4, 0.5]
    out, acc = CPU.compute_timestep(ts, min_dt_courant=2e-3)
    assert acc == np.inf
    assert out[0] == 1.1 * 1e-3 and out[1] == 1e-3 and out[2] == 0.5 + 1.1e-3
    out, _ = CPU.compute_timestep(ts, min_dt_courant=5e-4, min_dt_rho=7e-4)
    assert out[0] == 5e-4 and out[2] == 0.5 + 5e-4
    out, _ = CPU.compute_timestep(ts, min_dt_courant=5e-4, min_dt_rho=3e-4)
    assert out[0] == 3e-4
    rng = np.random.default_rng(3)
    n = 1000
    h = 0.01 + 0.01 * rng.random(n)
    a = rng.normal(0, 10, (3, n))
    out, acc = CPU.compute_timestep(ts, h, a[0], a[1], a[2], g=1.0, eta_acc=0.2, min_dt_courant=1.0)
    want = 0.2 * np.min(h * h / np.sum(a * a, axis=0)) ** 0.25
    assert acc == pytest.approx(want, rel=1e-15)
    assert out[0] == min(acc, 1.1e-3)


def test_update_h_closed_form():
    rng = np.random.default_rng(6)
    n = 500
    h = 0.01 + rng.random(n) * 0.02
    nc = np.full(n, 100, np.uint32)
    h2, _, removed = CPU.update_smoothing_length(100, nc, h)
    assert np.max(np.abs(h2 - h) / h) < 4e-16 and not removed
    nc = rng.integers(2, 400, n).astype(np.uint32)
    nc[[3, 17]] = [1, 0]
    keys = rng.integers(0, 2 ** 62, n).astype(np.uint64)
    h3, k3, removed = CPU.update_smoothing_length(100, nc, h, keys)
    ok = nc > 0
    want = h[ok] * 0.5 * (1 + 1023 * 100 / nc[ok].astype(np.float64)) ** 0.1
    assert np.max(np.abs(h3[ok] - want) / want) < 4e-16
    assert removed and k3[3] == k3[17] == 2 ** 63
    assert np.array_equal(np.delete(k3, [3, 17]), np.delete(keys, [3, 17]))


@need_ref
@pytest.mark.parametrize("bc", [oracle.BC_PERIODIC, oracle.BC_OPEN])
def test_positions_equal_reference(bc):
    st = state(3000, seed=11)
    st["du"][:50] = -1e4  # the exponential branch
    st["u"][:50] = 1.0
    box = oracle.Box.cube(0.0, 1.0, bc)
    a = CPU.compute_positions(st, 2e-3, 1.7e-3, box)
    b = REF.compute_positions(st, 2e-3, 1.7e-3, box)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


@need_ref
def test_positions_fixed_boundaries_equal_reference():
    # particles near the walls; the reference loop shares its factor variable between OpenMP
    # threads, so it is called one particle at a time
    st = state(300, seed=12)
    st["x"][:100] = np.linspace(0.0, 0.1, 100)
    box = oracle.Box.cube(0.0, 1.0, oracle.BC_FIXED)
    a = CPU.compute_positions(st, 1e-3, 1e-3, box)
    b = dict(st)
    for i in range(300):
        b = REF.compute_positions(b, 1e-3, 1e-3, box, first=i, last=i + 1)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


@need_ref
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_update_h_equal_reference(dtype):
    rng = np.random.default_rng(13)
    n = 2000
    h = 0.01 + rng.random(n) * 0.02
    nc = rng.integers(1, 400, n).astype(np.uint32)
    keys = rng.integers(0, 2 ** 62, n).astype(np.uint64)
    a = CPU.update_smoothing_length(100, nc, h, keys, dtype=dtype)
    b = REF.update_smoothing_length(100, nc, h, keys, dtype=dtype)
    assert a[0].dtype == dtype
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_update_h_f32_closed_form():
    """The FP32 instantiation: h unchanged at nc = ng0 (2 * 0.5 exactly, 1024^0.1 = 2 is exact in
    powf too), and within a few float ulp of the double closed form elsewhere."""
    rng = np.random.default_rng(8)
    n = 1000
    h = (0.01 + rng.random(n) * 0.02).astype(np.float32)
    h2, _, _ = CPU.update_smoothing_length(100, np.full(n, 100, np.uint32), h, dtype=np.float32)
    assert np.array_equal(h2, h)
    nc = rng.integers(2, 4000, n).astype(np.uint32)
    h3, _, _ = CPU.update_smoothing_length(100, nc, h, dtype=np.float32)
    want = h.astype(np.float64) * 0.5 * (1 + 1023 * 100 / nc.astype(np.float64)) ** 0.1
    assert np.max(np.abs(h3 - want) / want) < 3e-7
