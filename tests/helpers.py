"""Shared input generators and oracle pipelines for the tests."""
import numpy as np

import oracle

TABLES = {}


def tables():
    if "t" not in TABLES:
        TABLES["t"] = oracle.Cpu("oracle").kernel_tables()
    return TABLES["t"]


def uniform_box(n, seed=1, ng=100, periodic=True, lo=0.0, hi=1.0):
    """i.i.d. uniform positions in a cube, h for ~ng neighbors, v ~ N(0, 0.1^2), u = 1."""
    rng = np.random.default_rng(seed)
    L = hi - lo
    x, y, z = (lo + L * rng.random(n) for _ in range(3))
    h = np.full(n, 0.5 * L * (ng / (4 / 3 * np.pi * n)) ** (1 / 3))
    m = np.full(n, 1.0 / n)
    v = [rng.normal(0, 0.1, n) for _ in range(3)]
    u = np.ones(n)
    bc = oracle.BC_PERIODIC if periodic else oracle.BC_OPEN
    return dict(x=x, y=y, z=z, h=h, m=m, vx=v[0], vy=v[1], vz=v[2], u=u), oracle.Box.cube(lo, hi, bc)


def sedov_lattice(side=30):
    """Sedov blast-wave lattice: side^3 particles at cell centres of [-0.5, 0.5]^3, periodic,
    h for ~100 neighbors, u = 1e-10 except particles within r < 0.1 sharing total energy 1."""
    n = side ** 3
    c = -0.5 + (np.arange(side) + 0.5) / side
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    x, y, z = X.ravel().copy(), Y.ravel().copy(), Z.ravel().copy()
    m = np.full(n, 1.0 / n)
    h = np.full(n, 0.5 * (3 * 100 / (4 * np.pi * n)) ** (1 / 3))
    r = np.sqrt(x * x + y * y + z * z)
    u = np.full(n, 1e-10)
    inner = r < 0.1
    u[inner] = 1.0 / (inner.sum() * m[0])
    zero = np.zeros(n)
    return dict(x=x, y=y, z=z, h=h, m=m, vx=zero.copy(), vy=zero.copy(), vz=zero.copy(), u=u), \
        oracle.Box.cube(-0.5, 0.5, oracle.BC_PERIODIC)


def oracle_step(inp, box, ng0=100, ngmax=150, bucket=64, kind="oracle", gamma=5.0 / 3.0, Kcour=0.2):
    """The reference stage order on the CPU: returns a dict of SFC-sorted results."""
    cpu = oracle.Cpu(kind)
    n = len(inp["x"])
    keys = cpu.sfc_keys(inp["x"], inp["y"], inp["z"], box)
    keys, order = cpu.sort_by_key(keys, np.arange(n, dtype=np.uint32))
    s = {k: np.ascontiguousarray(v[order]) for k, v in inp.items()}
    t = cpu.tree(keys, box, bucket)
    wh, _, K = cpu.kernel_tables()
    nb, nc, h = cpu.find_neighbors_sph(s["x"], s["y"], s["z"], s["h"], box, t, ng0, ngmax)
    rho = cpu.density(nb, nc, ngmax, s["x"], s["y"], s["z"], h, s["m"], wh, K, box)
    p, c = cpu.eos(s["u"], rho, gamma)
    cs = cpu.iad(nb, nc, ngmax, s["x"], s["y"], s["z"], h, s["m"], rho, wh, K, box)
    ax, ay, az, du, dt = cpu.momentum_energy(nb, nc, ngmax, s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], h,
                                             s["m"], rho, p, c, cs, wh, K, Kcour, box)
    return dict(keys=keys, order=order, sorted=s, tree=t, nbrs=nb, nc=nc, h=h, rho=rho, p=p, c=c, c11=cs[0],
                c12=cs[1], c13=cs[2], c22=cs[3], c23=cs[4], c33=cs[5], ax=ax, ay=ay, az=az, du=du, minDt=dt, K=K,
                wh=wh)


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.abs(b), floor)
    scale[scale == 0] = 1.0
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


def oracle_ve_step(inp, box, alpha, dt, ng0=100, ngmax=150, bucket=64, gamma=5.0 / 3.0, Kcour=0.2, grad_v=True):
    """The reference VE stage order on the CPU (HydroVeProp::computeForces, avClean off):
    returns a dict of SFC-sorted results."""
    cpu = oracle.Cpu("oracle")
    n = len(inp["x"])
    keys = cpu.sfc_keys(inp["x"], inp["y"], inp["z"], box)
    keys, order = cpu.sort_by_key(keys, np.arange(n, dtype=np.uint32))
    s = {k: np.ascontiguousarray(v[order]) for k, v in inp.items()}
    t = cpu.tree(keys, box, bucket)
    wh, _, K = cpu.kernel_tables()
    nb, nc, h = cpu.find_neighbors_sph(s["x"], s["y"], s["z"], s["h"], box, t, ng0, ngmax)
    f = dict(s, h=h, wh=wh, whd=cpu.kernel_derivative_table(), alpha=np.ascontiguousarray(alpha[order]))
    f["xm"] = cpu.ve_xmass(f, K, nb, nc, ngmax, box)
    f["kx"], f["gradh"] = cpu.ve_def_gradh(f, K, nb, nc, ngmax, box)
    e = cpu.ve_eos(f["u"], f["m"], f["kx"], f["xm"], f["gradh"], gamma)
    f.update(prho=e["prho"], c=e["c"], rho=e["rho"], p=e["p"])
    f.update(cpu.ve_iad_divv_curlv(f, K, nb, nc, ngmax, box, grad_v=grad_v))
    f["alpha"] = cpu.ve_av_switches(f, K, nb, nc, ngmax, box, dt)
    f["ax"], f["ay"], f["az"], f["du"], f["minDt"] = cpu.ve_momentum_energy(f, K, nb, nc, ngmax, box, Kcour=Kcour)
    f.update(nc=nc, nbrs=nb, order=order, K=K)
    return f


def lists_digest(chunks):
    """SHA-256 over the stored neighbors of each particle in order, nc - 1 per particle (the
    lists are complete: ngmax exceeds every count), as u32 little-endian.  `chunks` yields
    (rows [k, ngmax], nc [k]) for consecutive particles, as numpy arrays."""
    import hashlib

    hl = hashlib.sha256()
    for nb, nc in chunks:
        keep = np.arange(nb.shape[1])[None, :] < (nc.astype(np.int64) - 1)[:, None]
        hl.update(np.ascontiguousarray(nb[keep], dtype="<u4").tobytes())
    return np.frombuffer(hl.digest(), np.uint8)
