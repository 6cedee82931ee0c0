"""CPU checker for the SPH hot path (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package.  It binds two shared libraries with ctypes:

* ``liboracle.so``: the C++ restatement of the reference CPU path
  (oracle/sphx_oracle.cpp), always built by oracle/Makefile;
* ``_ref/libsphexa_ref.so``: the reference's own CPU templates compiled from its
  headers (oracle/ref_driver.cpp); present only where the reference sources were
  available at build time.

``Cpu(lib)`` wraps either library behind one numpy interface, so a test can run
the same call on both and compare.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsphexa_ref.so")

TABLE_SIZE = 20000
BC_OPEN, BC_PERIODIC, BC_FIXED = 0, 1, 2


class Box(C.Structure):
    """sphx_box (include/sphx.h)."""

    _fields_ = [("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
                ("zmin", C.c_double), ("zmax", C.c_double), ("bx", C.c_int32), ("by", C.c_int32),
                ("bz", C.c_int32), ("reserved", C.c_int32)]

    @classmethod
    def cube(cls, lo, hi, bc=BC_OPEN):
        return cls(lo, hi, lo, hi, lo, hi, bc, bc, bc, 0)

    @classmethod
    def make(cls, lo3, hi3, bc3):
        return cls(lo3[0], hi3[0], lo3[1], hi3[1], lo3[2], hi3[2], bc3[0], bc3[1], bc3[2], 0)

    def lengths(self):
        return np.array([self.xmax - self.xmin, self.ymax - self.ymin, self.zmax - self.zmin])


def build(quiet=True):
    """Compile liboracle.so (and _ref when the reference sources are present)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


class VeFields(C.Structure):
    """Input fields of the VE loops (OrcVe in sphx_oracle.cpp): float64 arrays or None."""

    NAMES = ("x", "y", "z", "vx", "vy", "vz", "h", "m", "c", "xm", "kx", "gradh", "prho", "divv", "alpha", "c11",
             "c12", "c13", "c22", "c23", "c33", "dV11", "dV12", "dV13", "dV22", "dV23", "dV33", "wh", "whd")
    _fields_ = [(nm, C.c_void_p) for nm in NAMES] + [("K", C.c_double)]

    @classmethod
    def of(cls, fields: dict, K: float):
        keep = {}
        v = cls()
        for nm in cls.NAMES:
            a = fields.get(nm)
            if a is not None:
                a = np.ascontiguousarray(a, np.float64)
                keep[nm] = a
                setattr(v, nm, a.ctypes.data)
        v.K = K
        v._keep = keep  # keeps the arrays alive with the structure
        return v


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


_D = C.c_double
_U = C.c_uint32
_I = C.c_int
_V = C.c_void_p
_U64 = C.c_uint64


class Cpu:
    """numpy front end for liboracle.so (prefix 'orc') or libsphexa_ref.so (prefix 'ref')."""

    def __init__(self, kind="oracle"):
        path = ORACLE_SO if kind == "oracle" else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.kind = kind
        self.pre = "orc" if kind == "oracle" else "ref"
        self.lib = C.CDLL(path)
        f = self._f
        f("kernel_tables", None, [_D, _V, _V, _V])
        f("sphynx_k", _D, [_D])
        f("ihilbert", _U64, [_U, _U, _U])
        f("decode_hilbert", None, [_U64, _V])
        f("compute_sfc_keys", None, [_V, _V, _V, _V, _U64, _V])
        f("sort_by_key", None, [_V, _V, _U64])
        f("compute_octree", _I, [_V, _U64, _U, _U, _V, _V, _I])
        f("build_octree", None, [_V, _I, _V, _V, _V, _V, _V, _V])
        f("node_centers", None, [_V, _I, _V, _V, _V])
        if kind == "oracle":
            f("find_neighbors", None, [_V, _V, _V, _V, _U, _U, _V, _V, _V, _V, _V, _V, _V, _U, _V, _V])
            f("find_neighbors_sph", None, [_V, _V, _V, _V, _U, _U, _V, _V, _V, _V, _V, _V, _V, _U, _U, _V, _V])
            f("density", None, [_U, _U, _V, _V, _U, _V, _V, _V, _V, _V, _V, _D, _V, _V])
            f("eos", None, [_U, _U, _V, _V, _D, _V, _V])
            f("iad", None, [_U, _U, _V, _V, _U] + [_V] * 7 + [_D, _V] + [_V] * 6)
            f("momentum_energy", _D, [_U, _U, _V, _V, _U] + [_V] * 18 + [_D, C.c_float, _V] + [_V] * 4)
            f("make_global_box", None, [_V, _V, _V, _U64, _V, _V])
            f("compute_groups", _I, [_U, _U, _V, _V, _V, _V, _V, _I, _V, _V, C.c_float, _V, _I])
            f("iad_one", None, [_U, _V, _U] + [_V] * 7 + [_D, _V, _V])
            f("momentum_energy_one", None, [_U, _V, _U] + [_V] * 18 + [_D, _V, _V])
            f("rebalance_decision", _I, [_V, _V, _I, _U, _V])
            f("update_octree", _I, [_V, _U64, _U, _V, _V, _V, _I])
            f("update_h", _D, [_U, _U, _D])
            f("simpson_sin", _D, [_D, _D, _U64])
            f("lookup", _D, [_V, _D])
        else:
            f("find_neighbors", None, [_V, _V, _V, _V, _U, _U, _V, _I, _V, _V, _V, _V, _V, _V, _V, _V, _U, _V, _V])
            f("find_neighbors_sph", None,
              [_V, _V, _V, _V, _U, _U, _V, _I, _V, _V, _V, _V, _V, _V, _V, _V, _U, _U, _V, _V])
            f("density", None, [_U64, _U, _U, _V, _V, _U, _V, _V, _V, _V, _V, _V, _D, _V, _V])
            f("eos", None, [_U64, _U, _U, _V, _V, _D, _V, _V])
            f("iad", None, [_U64, _U, _U, _V, _V, _U] + [_V] * 7 + [_D, _V] + [_V] * 6)
            f("momentum_energy", _D, [_U64, _U, _U, _V, _V, _U] + [_V] * 18 + [_D, C.c_float, _V] + [_V] * 4)
            f("hydro_step", _D, [_U64] + [_V] * 9 + [_V, _U, _U, _U, _D, _D, C.c_float] + [_V] * 5)
        f("xmass_one", _D, [_U, _V, _U, _V, _V, _V, _V, _V, _V, _D, _V])
        if kind == "oracle":
            f("min_h2_a2", _D, [_U, _U, _V, _V, _V, _V])
            f("compute_timestep", _D, [_U, _U, _V, _V, _V, _V, _D, _D, _D, _D, _V])
            f("compute_positions", None, [_U, _U] + [_V] * 16 + [_D, _D, _V])
        else:
            f("compute_positions", None, [_U64, _U, _U] + [_V] * 16 + [_D, _D, _V])
        f("update_smoothing_length", _I, [_U, _U, _U, _V, _V, _V])
        f("update_smoothing_length_f32", _I, [_U, _U, _U, _V, _V, _V])
        pre = [] if kind == "oracle" else [_U64]
        loop = pre + [_U, _U, _V, _V, _U, _V, _V]
        f("ve_xmass", None, loop + [_V])
        f("ve_def_gradh", None, loop + [_V, _V])
        f("ve_eos", None, pre + [_U, _U, _V, _V, _V, _V, _V, _D, _V, _V, _V, _V])
        f("ve_iad_divv_curlv", None, loop + [_I] + [_V] * 14)
        f("ve_av_switches", None, loop + [_D, _D, _D, _D, _V])
        f("ve_momentum_energy", _D, loop + [_I, _D, _D, _D, C.c_float, _V, _V, _V, _V])
        f("kernel_derivative_table", None, [_D, _V])
        f("limit_box_shrinking", None, [_V, _V, _D, _V])

    def _f(self, name, res, args):
        fn = getattr(self.lib, f"{self.pre}_{name}")
        fn.restype = res
        fn.argtypes = args
        setattr(self, "_" + name, fn)

    # ---------------------------------------------------------------- tables
    def kernel_tables(self, sinc_index=6.0):
        wh64 = np.zeros(TABLE_SIZE, np.float64)
        wh32 = np.zeros(TABLE_SIZE, np.float32)
        K = np.zeros(1)
        self._kernel_tables(sinc_index, _p(wh64), _p(wh32), _p(K))
        return wh64, wh32, float(K[0])

    def sphynx_k(self, n):
        return self._sphynx_k(n)

    # ---------------------------------------------------------------- keys
    def ihilbert(self, x, y, z):
        return int(self._ihilbert(x, y, z))

    def decode_hilbert(self, key):
        out = np.zeros(3, np.uint32)
        self._decode_hilbert(key, _p(out))
        return tuple(int(v) for v in out)

    def sfc_keys(self, x, y, z, box, keys=None):
        n = len(x)
        keys = np.zeros(n, np.uint64) if keys is None else keys.copy()
        self._compute_sfc_keys(_p(x), _p(y), _p(z), _p(keys), n, C.byref(box))
        return keys

    def sort_by_key(self, keys, vals):
        k = np.ascontiguousarray(keys, np.uint64).copy()
        v = np.ascontiguousarray(vals, np.uint32).copy()
        self._sort_by_key(_p(k), _p(v), len(k))
        return k, v

    # ---------------------------------------------------------------- tree
    def compute_octree(self, keys, bucket, max_count=0xFFFFFFFF):
        keys = np.ascontiguousarray(keys, np.uint64)
        cap = max(16, len(keys) + 8)
        while True:
            leaves = np.zeros(cap, np.uint64)
            counts = np.zeros(cap, np.uint32)
            L = self._compute_octree(_p(keys), len(keys), bucket, max_count, _p(leaves), _p(counts), cap)
            if L > 0:
                return leaves[:L + 1].copy(), counts[:L].copy()
            cap = -L + 8

    def build_octree(self, leaves):
        L = len(leaves) - 1
        ni = (L - 1) // 7
        nn = L + ni
        o = dict(prefixes=np.zeros(nn, np.uint64), childOffsets=np.zeros(nn + 1, np.int32),
                 parents=np.zeros(max(1, (nn - 1) // 8), np.int32), levelRange=np.zeros(23, np.int32),
                 internalToLeaf=np.zeros(nn, np.int32), leafToInternal=np.zeros(nn, np.int32))
        self._build_octree(_p(np.ascontiguousarray(leaves, np.uint64)), L, _p(o["prefixes"]), _p(o["childOffsets"]),
                           _p(o["parents"]), _p(o["levelRange"]), _p(o["internalToLeaf"]), _p(o["leafToInternal"]))
        o["childOffsets"] = o["childOffsets"][:nn]
        o.update(numLeafNodes=L, numInternalNodes=ni, numNodes=nn)
        return o

    def node_centers(self, prefixes, box):
        nn = len(prefixes)
        c = np.zeros((nn, 3))
        s = np.zeros((nn, 3))
        self._node_centers(_p(np.ascontiguousarray(prefixes, np.uint64)), nn, C.byref(box), _p(c), _p(s))
        return c, s

    def tree(self, keys, box, bucket=64):
        """The single-rank harness (findneighbors.cpp:59-83): leaves, internal tree, layout, centres."""
        leaves, counts = self.compute_octree(keys, bucket)
        o = self.build_octree(leaves)
        layout = np.zeros(len(leaves), np.uint32)
        layout[1:] = np.cumsum(counts, dtype=np.uint64).astype(np.uint32)
        c, s = self.node_centers(o["prefixes"], box)
        o.update(leaves=leaves, counts=counts, layout=layout, centers=c, sizes=s)
        return o

    def compute_groups(self, x, y, z, h, leaves, layout, box, tol=2.0, first=0, last=None):
        """Target-group boundaries (numGroups + 1 entries) of computeGroupSplits with groupSize 64."""
        last = len(x) if last is None else last
        cap = 64 * ((last - first + 63) // 64) + 2
        b = np.zeros(cap, np.uint32)
        hh = None if h is None else np.ascontiguousarray(h, np.float64)
        ng = self._compute_groups(first, last, _p(x), _p(y), _p(z), _p(hh), _p(np.ascontiguousarray(leaves, np.uint64)),
                                  len(leaves) - 1, _p(np.ascontiguousarray(layout, np.uint32)), C.byref(box), tol,
                                  _p(b), cap)
        assert ng >= 0
        return b[:ng + 1]

    def make_global_box(self, x, y, z, prev):
        out = Box()
        self._make_global_box(_p(x), _p(y), _p(z), len(x), C.byref(prev), C.byref(out))
        return out

    def limit_box_shrinking(self, fitting, previous, shrink_limit=0.05):
        out = Box()
        self._limit_box_shrinking(C.byref(fitting), C.byref(previous), shrink_limit, C.byref(out))
        return out

    # ---------------------------------------------------------------- neighbors
    def find_neighbors(self, x, y, z, h, box, t, ngmax, first=0, last=None, lists=True):
        last = len(x) if last is None else last
        n = last - first
        nc = np.zeros(n, np.uint32)
        nb = np.zeros(n * ngmax, np.uint32) if lists else None
        if self.kind == "oracle":
            self._find_neighbors(_p(x), _p(y), _p(z), _p(h), first, last, C.byref(box), _p(t["childOffsets"]),
                                 _p(t["parents"]), _p(t["internalToLeaf"]), _p(t["layout"]), _p(t["centers"]),
                                 _p(t["sizes"]), ngmax, _p(nb), _p(nc))
        else:
            if nb is None:
                nb = np.zeros(n * ngmax, np.uint32)
            self._find_neighbors(_p(x), _p(y), _p(z), _p(h), first, last, C.byref(box), t["numLeafNodes"],
                                 _p(t["prefixes"]), _p(t["childOffsets"]), _p(t["parents"]), _p(t["internalToLeaf"]),
                                 _p(t["levelRange"]), _p(t["layout"]), _p(t["centers"]), _p(t["sizes"]), ngmax,
                                 _p(nb), _p(nc))
        return (nb.reshape(n, ngmax) if nb is not None else None), nc

    def find_neighbors_sph(self, x, y, z, h, box, t, ng0, ngmax, first=0, last=None):
        """Returns (lists[n, ngmax], nc with self, h updated copy)."""
        last = len(x) if last is None else last
        n = last - first
        h = np.array(h, np.float64, copy=True)
        nc = np.zeros(n, np.uint32)
        nb = np.zeros(n * ngmax, np.uint32)
        if self.kind == "oracle":
            self._find_neighbors_sph(_p(x), _p(y), _p(z), _p(h), first, last, C.byref(box), _p(t["childOffsets"]),
                                     _p(t["parents"]), _p(t["internalToLeaf"]), _p(t["layout"]), _p(t["centers"]),
                                     _p(t["sizes"]), ng0, ngmax, _p(nb), _p(nc))
        else:
            self._find_neighbors_sph(_p(x), _p(y), _p(z), _p(h), first, last, C.byref(box), t["numLeafNodes"],
                                     _p(t["prefixes"]), _p(t["childOffsets"]), _p(t["parents"]),
                                     _p(t["internalToLeaf"]), _p(t["levelRange"]), _p(t["layout"]), _p(t["centers"]),
                                     _p(t["sizes"]), ng0, ngmax, _p(nb), _p(nc))
        return nb.reshape(n, ngmax), nc, h

    # ---------------------------------------------------------------- SPH loops (FP64)
    def _n(self, *a):
        return [] if self.kind == "oracle" else [a[0]]

    def density(self, nb, nc, ngmax, x, y, z, h, m, wh, K, box, first=0, last=None):
        n = len(x)
        last = n if last is None else last
        rho = np.zeros(n)
        self._density(*self._n(n), first, last, _p(nb), _p(nc), ngmax, _p(x), _p(y), _p(z), _p(h), _p(m), _p(wh), K,
                      C.byref(box), _p(rho))
        return rho

    def eos(self, u, rho, gamma, first=0, last=None):
        n = len(u)
        last = n if last is None else last
        p = np.zeros(n)
        c = np.zeros(n)
        self._eos(*self._n(n), first, last, _p(u), _p(rho), gamma, _p(p), _p(c))
        return p, c

    def iad(self, nb, nc, ngmax, x, y, z, h, m, rho, wh, K, box, first=0, last=None):
        n = len(x)
        last = n if last is None else last
        cs = [np.zeros(n) for _ in range(6)]
        self._iad(*self._n(n), first, last, _p(nb), _p(nc), ngmax, _p(x), _p(y), _p(z), _p(h), _p(m), _p(rho),
                  _p(wh), K, C.byref(box), *[_p(a) for a in cs])
        return cs

    def momentum_energy(self, nb, nc, ngmax, x, y, z, vx, vy, vz, h, m, rho, p, c, cs, wh, K, Kcour, box, first=0,
                        last=None):
        n = len(x)
        last = n if last is None else last
        out = [np.zeros(n) for _ in range(4)]
        dt = self._momentum_energy(*self._n(n), first, last, _p(nb), _p(nc), ngmax, _p(x), _p(y), _p(z), _p(vx),
                                   _p(vy), _p(vz), _p(h), _p(m), _p(rho), _p(p), _p(c), *[_p(a) for a in cs], _p(wh),
                                   K, Kcour, C.byref(box), *[_p(a) for a in out])
        return out[0], out[1], out[2], out[3], dt

    def xmass_one(self, i, nbrs, x, y, z, h, m, wh, K, box):
        nb = np.ascontiguousarray(nbrs, np.uint32)
        return self._xmass_one(i, _p(nb), len(nb), _p(x), _p(y), _p(z), _p(h), _p(m), _p(wh), K, C.byref(box))


    # ---------------------------------------------------------------- integration step
    def compute_timestep(self, ts, h=None, ax=None, ay=None, az=None, g=0.0, eta_acc=0.2, min_dt_courant=np.inf,
                         min_dt_rho=np.inf, first=0, last=None):
        """computeTimestep on one rank; ts = [minDt, minDt_m1, ttot] (updated copy returned)
        with minDtAcc."""
        t = np.array(ts, dtype=np.float64)
        n = 0 if h is None else len(h)
        last = n if last is None else last
        acc = self._compute_timestep(first, last, _p(h), _p(ax), _p(ay), _p(az), g, eta_acc, min_dt_courant,
                                     min_dt_rho, _p(t))
        return t, acc

    def min_h2_a2(self, h, ax, ay, az):
        return self._min_h2_a2(0, len(h), _p(h), _p(ax), _p(ay), _p(az))

    def compute_positions(self, st, dt, dt_m1, box, first=0, last=None):
        """updatePositionsHost + updateIntEnergyHost on copies of the state dict st (keys x, y,
        z, x_m1, y_m1, z_m1, vx, vy, vz, ax, ay, az, h, u, du, du_m1, all float64); returns
        the updated copy."""
        o = {k: np.array(v, dtype=np.float64) for k, v in st.items()}
        n = len(o["x"])
        last = n if last is None else last
        args = [_p(o[k]) for k in ("x", "y", "z", "x_m1", "y_m1", "z_m1", "vx", "vy", "vz", "ax", "ay", "az", "h",
                                   "u", "du", "du_m1")]
        self._compute_positions(*self._n(n), first, last, *args, dt, dt_m1, C.byref(box))
        return o

    def update_smoothing_length(self, ng0, nc, h, keys=None, first=0, last=None, dtype=np.float64):
        """updateSmoothingLengthCpu on copies: (h, keys or None, keysRemoved); dtype float32
        runs the FP32 instantiation (the reference app's h type)."""
        h = np.array(h, dtype=dtype)
        nc = np.ascontiguousarray(nc, np.uint32)
        k = None if keys is None else np.array(keys, dtype=np.uint64)
        last = len(h) if last is None else last
        fn = self._update_smoothing_length_f32 if h.dtype == np.float32 else self._update_smoothing_length
        r = fn(first, last, ng0, _p(nc), _p(h), _p(k))
        return h, k, bool(r)


    # ---------------------------------------------------------------- VE loops (FP64)
    def kernel_derivative_table(self, sinc_index=6.0):
        whd = np.zeros(TABLE_SIZE, np.float64)
        self._kernel_derivative_table(sinc_index, _p(whd))
        return whd

    def _ve(self, fields, K, nb, nc, ngmax, first, last):
        n = len(fields["x"])
        last = n if last is None else last
        v = VeFields.of(fields, K)
        nb = np.ascontiguousarray(nb, np.uint32)
        nc = np.ascontiguousarray(nc, np.uint32)
        return n, last, v, [*self._n(n), first, last, _p(nb), _p(nc), ngmax, C.byref(v)], (nb, nc)

    def ve_xmass(self, fields, K, nb, nc, ngmax, box, first=0, last=None):
        n, last, v, a, keep = self._ve(fields, K, nb, nc, ngmax, first, last)
        xm = np.zeros(n)
        self._ve_xmass(*a, C.byref(box), _p(xm))
        return xm

    def ve_def_gradh(self, fields, K, nb, nc, ngmax, box, first=0, last=None):
        n, last, v, a, keep = self._ve(fields, K, nb, nc, ngmax, first, last)
        kx, gradh = np.zeros(n), np.zeros(n)
        self._ve_def_gradh(*a, C.byref(box), _p(kx), _p(gradh))
        return kx, gradh

    def ve_eos(self, u, m, kx, xm, gradh, gamma, first=0, last=None):
        n = len(u)
        last = n if last is None else last
        out = [np.zeros(n) for _ in range(4)]
        arrs = [np.ascontiguousarray(a, np.float64) for a in (u, m, kx, xm, gradh)]
        self._ve_eos(*self._n(n), first, last, *[_p(a) for a in arrs], gamma, *[_p(o) for o in out])
        return dict(prho=out[0], c=out[1], rho=out[2], p=out[3])

    def ve_iad_divv_curlv(self, fields, K, nb, nc, ngmax, box, first=0, last=None, grad_v=True, given_iad=False):
        """IAD (c11..c33) then divv, curlv and the symmetrised velocity gradient; given_iad
        (oracle only) runs the divv part alone on the c11..c33 of `fields`."""
        n, last, v, a, keep = self._ve(fields, K, nb, nc, ngmax, first, last)
        names = ("c11", "c12", "c13", "c22", "c23", "c33", "divv", "curlv", "dV11", "dV12", "dV13", "dV22", "dV23",
                 "dV33")
        out = {nm: (np.zeros(n) if (grad_v or not nm.startswith("dV")) else None) for nm in names}
        self._ve_iad_divv_curlv(*a, C.byref(box), int(given_iad), *[_p(out[nm]) for nm in names])
        return {k: o for k, o in out.items() if o is not None}

    def ve_av_switches(self, fields, K, nb, nc, ngmax, box, dt, alphamin=0.05, alphamax=1.0, decay_constant=0.2,
                       first=0, last=None):
        n, last, v, a, keep = self._ve(fields, K, nb, nc, ngmax, first, last)
        alpha = np.array(fields["alpha"], dtype=np.float64)
        self._ve_av_switches(*a, C.byref(box), dt, alphamin, alphamax, decay_constant, _p(alpha))
        return alpha

    def ve_momentum_energy(self, fields, K, nb, nc, ngmax, box, av_clean=False, Atmin=0.1, Atmax=0.2, ramp=None,
                           Kcour=0.2, first=0, last=None):
        n, last, v, a, keep = self._ve(fields, K, nb, nc, ngmax, first, last)
        ramp = 1.0 / (Atmax - Atmin) if ramp is None else ramp
        out = [np.zeros(n) for _ in range(4)]
        dt = self._ve_momentum_energy(*a, C.byref(box), int(av_clean), Atmin, Atmax, ramp, Kcour,
                                      *[_p(o) for o in out])
        return out[0], out[1], out[2], out[3], dt


def available(kind):
    return os.path.exists(ORACLE_SO if kind == "oracle" else REF_SO)
