"""Host side of the MI355X SPH hot path: a Python mirror of the reference's C++
entry points over torch device tensors, calling libsphx.so through its C ABI.

Names follow the reference interface:

=============================  ==============================================================
this module                    reference entry point
=============================  ==============================================================
``compute_sfc_keys``           cstone::computeSfcKeys<true> (domain/include/cstone/sfc/sfc_gpu.h:26)
``sort_by_key``                cstone::sortByKey<true> (primitives/primitives_acc.hpp:115)
``gather``                     cstone::gather (primitives/gather.hpp:90)
``gather_arrays``              cstone::gatherArrays (domain/layout.hpp:231)
``compute_octree``             cstone::computeOctree (tree/csarray.hpp:430)
``build_octree``               cstone::updateInternalTree / buildOctreeGpu (tree/octree.hpp:361)
``node_fp_centers``            cstone::nodeFpCenters (focus/source_center.hpp:146)
``find_neighbors``             cstone::findNeighbors (findneighbors.hpp:149)
``find_neighbors_sfc``         sph::findNeighborsSfc (sph/include/sph/find_neighbors.hpp:44)
``compute_density``            sph::computeDensity (sph/include/sph/hydro_std/density.hpp:55)
``compute_eos_hydro_std``      sph::computeEOS_HydroStd (hydro_std/eos.hpp:144)
``compute_iad``                sph::computeIAD (hydro_std/iad.hpp:74)
``compute_momentum_energy_std``sph::computeMomentumEnergySTD (hydro_std/momentum_energy.hpp:96)
``compute_timestep``           sph::computeTimestep (sph/include/sph/ts_global.hpp:101)
``compute_positions``          sph::computePositions (sph/include/sph/positions.hpp:208)
``update_smoothing_length``    sph::updateSmoothingLength (sph/include/sph/update_h.hpp:35)
``HydroStep.compute_forces``   HydroProp::computeForces (main/src/propagator/std_hydro.hpp:119)
``HydroStep.integrate``        HydroProp::integrate (main/src/propagator/std_hydro.hpp:170)
``Simulation.step``            one time step: sync, computeForces, integrate
``compute_xmass``              sph::computeXMass (sph/include/sph/hydro_ve/xmass.hpp:69)
``compute_ve_def_gradh``       sph::computeVeDefGradh (hydro_ve/ve_def_gradh.hpp:80)
``compute_eos_ve``             sph::computeEOS, ideal gas (hydro_ve/eos.hpp:192)
``compute_iad_divv_curlv``     sph::computeIadDivvCurlv (hydro_ve/iad_divv_curlv.hpp:92)
``compute_av_switches``        sph::computeAVswitches (hydro_ve/av_switches.hpp:81)
``compute_momentum_energy_ve`` sph::computeMomentumEnergy<false> (hydro_ve/momentum_energy.hpp:104)
``VeHydroStep.compute_forces`` HydroVeProp::computeForces (main/src/propagator/ve_hydro.hpp:126)
=============================  ==============================================================

Errors raise ``SphxError`` carrying the library's status code; there is no CPU
fallback.  Device memory comes from torch; the library's context launches on
torch's current stream.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import math
import weakref
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import Box, Groups, Octree, Particles, SphxError, Timestep, VeFields, FP64, MIXED, BC_OPEN, BC_PERIODIC, \
    BC_FIXED

_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpyAsync.restype = C.c_int
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return _hip


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _dev(t, dtypes, n, name, optional=False, on_device=True):
    """Checked device pointer of a tensor the library reads or writes n elements of: the
    dtype must be one of ``dtypes`` (the kernels reinterpret the memory), the tensor
    contiguous (the kernels index it densely), at least n long, and on a GPU (unless
    ``on_device`` is False).  None is allowed only when ``optional``.  Raises
    TypeError/ValueError before any device call."""
    if t is None:
        if optional:
            return None
        raise ValueError(f"{name}: missing")
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    dtypes = dtypes if isinstance(dtypes, tuple) else (dtypes,)
    if t.dtype not in dtypes:
        raise TypeError(f"{name}: dtype {t.dtype}, expected {' or '.join(str(d) for d in dtypes)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: not contiguous (pass .contiguous())")
    if t.numel() < n:
        raise ValueError(f"{name}: {t.numel()} elements, at least {n} needed")
    if on_device and not t.is_cuda:
        raise ValueError(f"{name}: on {t.device}, expected a GPU tensor")
    return C.c_void_p(t.data_ptr())


_F64, _F32, _I32, _I64 = torch.float64, torch.float32, torch.int32, torch.int64


def make_box(lo, hi, bc=BC_OPEN):
    """cstone::Box<double>: lo/hi scalars or 3-sequences, bc scalar or 3-sequence."""
    lo3 = [lo] * 3 if np.isscalar(lo) else list(lo)
    hi3 = [hi] * 3 if np.isscalar(hi) else list(hi)
    bc3 = [bc] * 3 if np.isscalar(bc) else list(bc)
    return Box(lo3[0], hi3[0], lo3[1], hi3[1], lo3[2], hi3[2], bc3[0], bc3[1], bc3[2], 0)


class Context:
    """Owns a sphx_ctx bound to torch's current stream on ``device``."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("sphexa_amd needs a HIP device (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", device)
        self.stream = torch.cuda.current_stream(self.device)
        h = C.c_void_p()
        st = self.lib.sphx_ctx_create(device, C.c_void_p(self.stream.cuda_stream), C.byref(h))
        if st != 0:
            raise SphxError(st, "sphx_ctx_create failed")
        self.h = h
        self._min_dt_owner = None  # weak reference to the ParticlesData whose time step is pending
        self.defer_min_dt(True)

    def check(self, st):
        if st != 0:
            raise SphxError(st, self.lib.sphx_last_error(self.h).decode())

    def defer_min_dt(self, on: bool):
        """Momentum-energy calls leave their Courant time step on the device (on, the default
        here) and ParticlesData.minDtCourant waits for it when read, or each call reads it
        back at once (off, the library's default)."""
        self._resolve_min_dt()
        self.check(self.lib.sphx_ctx_set_option(self.h, _lib.OPT_DEFER_MIN_DT, int(bool(on))))

    def min_dt_courant(self) -> float:
        """The Courant time step of the context's last momentum-energy call (waits for it)."""
        v = C.c_double()
        self.check(self.lib.sphx_min_dt_courant(self.h, C.byref(v)))
        return v.value

    def _resolve_min_dt(self):
        """Read a pending time step into its ParticlesData before another call replaces it."""
        owner = self._min_dt_owner() if self._min_dt_owner is not None else None
        self._min_dt_owner = None
        if owner is not None and owner._min_dt_ctx is self:
            owner.minDtCourant = self.min_dt_courant()

    def _min_dt_call(self, d, call):
        """Run a momentum-energy call (returning its sphx_particles view) and give d its time
        step: the value, or the context to wait on when the call deferred it."""
        owner = self._min_dt_owner() if self._min_dt_owner is not None else None
        if owner is not None and owner is not d:
            self._resolve_min_dt()
        s = call()
        if math.isnan(s.minDtCourant):
            d._min_dt, d._min_dt_ctx = math.nan, self
            self._min_dt_owner = weakref.ref(d)
        else:
            d.minDtCourant = s.minDtCourant

    def set_fp64_exact(self, on: bool):
        self.check(self.lib.sphx_ctx_set_option(self.h, _lib.OPT_FP64_EXACT, int(bool(on))))

    def set_timing(self, on):
        """Record HIP events around every hot kernel launch (True), around the neighbor-candidate
        and SPH particle-loop launches only ("loops"), or none (False); read with timing_summary."""
        mode = 2 if on == "loops" else int(bool(on))
        self.check(self.lib.sphx_ctx_set_option(self.h, _lib.OPT_TIMING, mode))

    def timing_summary(self):
        """{kernel: (launches, total ms)} since the last call; synchronises the stream."""
        cap = 64
        names = (C.c_char_p * cap)()
        launches = (C.c_int * cap)()
        ms = (C.c_double * cap)()
        num = C.c_int()
        self.check(self.lib.sphx_timing_summary(self.h, cap, names, launches, ms, C.byref(num)))
        return {names[k].decode(): (launches[k], ms[k]) for k in range(min(num.value, cap))}

    def set_loop_shape(self, precision: int, shape: int):
        """Particle-loop kernel shape: 0 LDS-staged sub-groups, 1 one lane per target."""
        opt = _lib.OPT_LOOP_SHAPE_FP64 if precision == FP64 else _lib.OPT_LOOP_SHAPE_MIXED
        self.check(self.lib.sphx_ctx_set_option(self.h, opt, int(shape)))

    def candidate_tests(self) -> float:
        """(target, candidate) pairs tested by the last neighbor search."""
        v = C.c_double()
        self.check(self.lib.sphx_get_stat(self.h, _lib.STAT_CANDIDATE_TESTS, C.byref(v)))
        return v.value

    def synchronize(self):
        self.check(self.lib.sphx_ctx_synchronize(self.h))

    @contextlib.contextmanager
    def on_stream(self, stream):
        """Launch on ``stream`` (a default-priority torch stream of this device) inside the
        block, then on the context's own stream again."""
        self.check(self.lib.sphx_ctx_set_stream(self.h, C.c_void_p(stream.cuda_stream)))
        try:
            yield
        finally:
            self.check(self.lib.sphx_ctx_set_stream(self.h, C.c_void_p(self.stream.cuda_stream)))

    def close(self):
        if self.h:
            self._resolve_min_dt()  # a pending time step stays readable
            self.lib.sphx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def d2d(self, dst: torch.Tensor, src_ptr: int, nbytes: int):
        st = _hiprt().hipMemcpyAsync(C.c_void_p(dst.data_ptr()), C.c_void_p(src_ptr), nbytes, 3,
                                     C.c_void_p(self.stream.cuda_stream))
        if st != 0:
            raise RuntimeError(f"hipMemcpyAsync failed ({st})")


def kernel_tables(sinc_index: float = 6.0):
    """(wh64 numpy, wh32 numpy, K): ParticlesData::createTables (particles_data.hpp:377-387)."""
    lib = _lib.load()
    wh64 = np.zeros(_lib.TABLE_SIZE, np.float64)
    wh32 = np.zeros(_lib.TABLE_SIZE, np.float32)
    K = C.c_double()
    lib.sphx_kernel_tables(sinc_index, wh64.ctypes.data_as(C.c_void_p), wh32.ctypes.data_as(C.c_void_p),
                           C.byref(K))
    return wh64, wh32, K.value


# ------------------------------------------------------------------------------------------- SFC, sort
def make_global_box(ctx: Context, x, y, z, previous: Box) -> Box:
    """cstone::makeGlobalBox on one rank (sfc/box_mpi.hpp:68-107): open dimensions take
    the coordinate extrema, periodic and fixed ones keep the limits of `previous`."""
    out = Box()
    n = x.numel()
    ctx.check(ctx.lib.sphx_make_global_box(ctx.h, _dev(x, _F64, n, "x"), _dev(y, _F64, n, "y"), _dev(z, _F64, n, "z"),
                                           n, C.byref(previous), C.byref(out)))
    return out


def limit_box_shrinking(fitting: Box, previous: Box, shrink_limit: float = 0.05) -> Box:
    """cstone::limitBoxShrinking (domain/include/cstone/sfc/box.hpp:397-414): each side of the
    fitted box moves inward by at most shrink_limit of the previous box's length (Box's own
    xmax - xmin); the boundary types are the previous box's.  Host arithmetic in double."""
    lx, ly, lz = previous.xmax - previous.xmin, previous.ymax - previous.ymin, previous.zmax - previous.zmin
    return Box(min(fitting.xmin, previous.xmin + shrink_limit * lx), max(fitting.xmax, previous.xmax - shrink_limit * lx),
               min(fitting.ymin, previous.ymin + shrink_limit * ly), max(fitting.ymax, previous.ymax - shrink_limit * ly),
               min(fitting.zmin, previous.zmin + shrink_limit * lz), max(fitting.zmax, previous.zmax - shrink_limit * lz),
               previous.bx, previous.by, previous.bz, 0)


def compute_sfc_keys(ctx: Context, x, y, z, box: Box, keys=None):
    n = x.numel()
    if keys is None:
        keys = torch.zeros(n, dtype=torch.int64, device=x.device)
    ctx.check(ctx.lib.sphx_compute_sfc_keys(ctx.h, _dev(x, _F64, n, "x"), _dev(y, _F64, n, "y"),
                                            _dev(z, _F64, n, "z"), _dev(keys, _I64, n, "keys"), n, C.byref(box)))
    return keys


def sequence(ctx: Context, n: int, first: int = 0, device=None):
    v = torch.empty(n, dtype=torch.int32, device=device or ctx.device)
    ctx.check(ctx.lib.sphx_sequence(ctx.h, _dev(v, _I32, n, "sequence"), n, first))
    return v


def sort_by_key(ctx: Context, keys, values):
    """Stable in-place sort of (int64-stored uint64 keys, int32-stored uint32 values)."""
    n = keys.numel()
    ctx.check(ctx.lib.sphx_sort_by_key(ctx.h, _dev(keys, _I64, n, "keys"), _dev(values, _I32, n, "values", True),
                                       n))


def gather(ctx: Context, order, src, dst=None):
    """dst[i] = src[order[i]] for i < src.numel(); dst has src's dtype."""
    fns = {torch.float64: ctx.lib.sphx_gather_f64, torch.float32: ctx.lib.sphx_gather_f32,
           torch.int32: ctx.lib.sphx_gather_u32}
    if src.dtype not in fns:
        raise TypeError(f"gather: src dtype {src.dtype}, expected float64, float32 or int32")
    dst = torch.empty_like(src) if dst is None else dst
    n = src.numel()
    ctx.check(fns[src.dtype](ctx.h, _dev(order, _I32, n, "order"), _dev(src, src.dtype, n, "src"),
                             _dev(dst, src.dtype, n, "dst"), n))
    return dst


_GATHER_BYTES = {torch.float64: 8, torch.int64: 8, torch.float32: 4, torch.int32: 4}


def gather_arrays(ctx: Context, order, pairs):
    """dst[i] = src[order[i]] for every (src, dst) in ``pairs`` (at most 16, all of
    order's length, 4- or 8-byte dtypes, dst of src's dtype), in one launch."""
    pairs = list(pairs)
    if not pairs:
        return
    if len(pairs) > _lib.GATHER_MAX_ARRAYS:
        raise ValueError(f"gather_arrays: {len(pairs)} arrays, at most {_lib.GATHER_MAX_ARRAYS}")
    n = order.numel()
    k = len(pairs)
    src_p, dst_p, size = (C.c_void_p * k)(), (C.c_void_p * k)(), (C.c_uint32 * k)()
    for j, (src, dst) in enumerate(pairs):
        if src.dtype not in _GATHER_BYTES:
            raise TypeError(f"gather_arrays: array {j} has dtype {src.dtype}, expected a 4- or 8-byte type")
        src_p[j] = _dev(src, src.dtype, n, f"src {j}")
        dst_p[j] = _dev(dst, src.dtype, n, f"dst {j}")
        size[j] = _GATHER_BYTES[src.dtype]
    ctx.check(ctx.lib.sphx_gather_arrays(ctx.h, _dev(order, _I32, n, "order"), n, k, src_p, dst_p, size))


# ------------------------------------------------------------------------------------------- octree
def compute_octree(ctx: Context, keys, bucket_size: int = 64, max_count: int = 0xFFFFFFFF):
    """(leaves int64[L+1], counts int32[L], layout int32[L+1]) as new device tensors."""
    L = C.c_int32()
    pl, pc, py = C.c_void_p(), C.c_void_p(), C.c_void_p()
    ctx.check(ctx.lib.sphx_compute_octree(ctx.h, _dev(keys, _I64, keys.numel(), "keys"), keys.numel(), bucket_size,
                                          max_count, C.byref(L),
                                          C.byref(pl), C.byref(pc), C.byref(py)))
    n = L.value
    leaves = torch.empty(n + 1, dtype=torch.int64, device=keys.device)
    counts = torch.empty(n, dtype=torch.int32, device=keys.device)
    layout = torch.empty(n + 1, dtype=torch.int32, device=keys.device)
    ctx.d2d(leaves, pl.value, 8 * (n + 1))
    ctx.d2d(counts, pc.value, 4 * n)
    ctx.d2d(layout, py.value, 4 * (n + 1))
    return leaves, counts, layout


@dataclass
class OctreeData:
    """Device arrays of the internal octree + neighbor-search view (OctreeData, OctreeNsView)."""

    numLeafNodes: int
    numInternalNodes: int
    numNodes: int
    prefixes: torch.Tensor
    childOffsets: torch.Tensor
    parents: torch.Tensor
    levelRange: torch.Tensor
    internalToLeaf: torch.Tensor
    leafToInternal: torch.Tensor
    leaves: torch.Tensor
    layout: torch.Tensor | None = None
    centers: torch.Tensor | None = None
    sizes: torch.Tensor | None = None

    def view(self) -> Octree:
        return Octree(self.numLeafNodes, self.numInternalNodes, self.numNodes, 0, _ptr(self.prefixes),
                      _ptr(self.childOffsets), _ptr(self.parents), _ptr(self.levelRange), _ptr(self.internalToLeaf),
                      _ptr(self.leafToInternal), _ptr(self.leaves), _ptr(self.layout), _ptr(self.centers),
                      _ptr(self.sizes))


def build_octree(ctx: Context, leaves):
    L = leaves.numel() - 1
    ni = (L - 1) // 7
    nn = L + ni
    dev = leaves.device
    o = OctreeData(L, ni, nn, torch.empty(nn, dtype=torch.int64, device=dev),
                   torch.empty(nn + 1, dtype=torch.int32, device=dev),
                   torch.empty(max(1, (nn - 1) // 8), dtype=torch.int32, device=dev),
                   torch.empty(23, dtype=torch.int32, device=dev), torch.empty(nn, dtype=torch.int32, device=dev),
                   torch.empty(nn, dtype=torch.int32, device=dev), leaves)
    v = o.view()
    ctx.check(ctx.lib.sphx_build_octree(ctx.h, _ptr(leaves), C.byref(v)))
    return o


def node_fp_centers(ctx: Context, octree: OctreeData, box: Box):
    nn = octree.numNodes
    octree.centers = torch.empty(nn * 3, dtype=torch.float64, device=octree.prefixes.device)
    octree.sizes = torch.empty(nn * 3, dtype=torch.float64, device=octree.prefixes.device)
    ctx.check(ctx.lib.sphx_node_centers(ctx.h, _ptr(octree.prefixes), nn, C.byref(box), _ptr(octree.centers),
                                        _ptr(octree.sizes)))
    return octree.centers, octree.sizes


def tree_from_keys(ctx: Context, keys, box: Box, bucket_size: int = 64):
    """Single-rank harness: leaves, internal tree, layout, centres (findneighbors.cpp:59-83)."""
    leaves, counts, layout = compute_octree(ctx, keys, bucket_size)
    o = build_octree(ctx, leaves)
    o.layout = layout
    o.counts = counts
    node_fp_centers(ctx, o, box)
    return o


# ------------------------------------------------------------------------------------------- neighbors
def find_neighbors(ctx: Context, x, y, z, h, box: Box, octree: OctreeData, ngmax: int, first=0, last=None,
                   lists=True):
    """cstone::findNeighbors: (nbrs int32[n, ngmax] or None, nc int32[n]) for [first, last)."""
    last = x.numel() if last is None else last
    n = last - first
    nc = torch.zeros(n, dtype=torch.int32, device=x.device)
    nb = torch.zeros(n * ngmax, dtype=torch.int32, device=x.device) if lists else None
    v = octree.view()
    m = x.numel()
    ctx.check(ctx.lib.sphx_find_neighbors(ctx.h, _dev(x, _F64, m, "x"), _dev(y, _F64, m, "y"), _dev(z, _F64, m, "z"),
                                          _dev(h, _F64, m, "h"), first, last, C.byref(box), C.byref(v), ngmax,
                                          _ptr(nb), _ptr(nc)))
    return (nb.view(n, ngmax) if nb is not None else None), nc


@dataclass
class ParticlesData:
    """Device particle fields (sphexa::ParticlesData field contract, particles_data.hpp:218-246)."""

    n: int
    precision: int = FP64
    ng0: int = 100
    ngmax: int = 150
    Kcour: float = 0.2
    gamma: float = 5.0 / 3.0
    sincIndex: float = 6.0
    K: float = 0.0
    x: torch.Tensor | None = None
    y: torch.Tensor | None = None
    z: torch.Tensor | None = None
    h: torch.Tensor | None = None
    m: torch.Tensor | None = None
    vx: torch.Tensor | None = None
    vy: torch.Tensor | None = None
    vz: torch.Tensor | None = None
    u: torch.Tensor | None = None
    rho: torch.Tensor | None = None
    p: torch.Tensor | None = None
    c: torch.Tensor | None = None
    c11: torch.Tensor | None = None
    c12: torch.Tensor | None = None
    c13: torch.Tensor | None = None
    c22: torch.Tensor | None = None
    c23: torch.Tensor | None = None
    c33: torch.Tensor | None = None
    ax: torch.Tensor | None = None
    ay: torch.Tensor | None = None
    az: torch.Tensor | None = None
    du: torch.Tensor | None = None
    nc: torch.Tensor | None = None
    keys: torch.Tensor | None = None
    wh: torch.Tensor | None = None
    x_m1: torch.Tensor | None = None
    y_m1: torch.Tensor | None = None
    z_m1: torch.Tensor | None = None
    du_m1: torch.Tensor | None = None
    # the Courant time step of the last momentum-energy call (minDtCourant), and the context
    # to read it from while that call's value is still on the device
    _min_dt: float = field(default=math.inf, init=False, repr=False)
    _min_dt_ctx: object = field(default=None, init=False, repr=False, compare=False)

    @property
    def minDtCourant(self) -> float:
        if self._min_dt_ctx is not None:
            self._min_dt, self._min_dt_ctx = self._min_dt_ctx.min_dt_courant(), None
        return self._min_dt

    @minDtCourant.setter
    def minDtCourant(self, v: float):
        self._min_dt, self._min_dt_ctx = float(v), None

    HYDRO = ("h", "m", "vx", "vy", "vz", "u", "rho", "p", "c", "c11", "c12", "c13", "c22", "c23", "c33", "ax",
             "ay", "az", "du")
    # integration state, allocated by allocate(integration=True)
    INTEGRATION = ("x_m1", "y_m1", "z_m1", "du_m1")

    @property
    def dtype(self):
        return torch.float64 if self.precision == FP64 else torch.float32

    def allocate(self, device, integration=False):
        """Allocate every missing field (and the integration state if asked), make the kernel
        table and K."""
        for f in ("x", "y", "z"):
            if getattr(self, f) is None:
                setattr(self, f, torch.zeros(self.n, dtype=torch.float64, device=device))
        for f in self.HYDRO + (self.INTEGRATION if integration else ()):
            if getattr(self, f) is None:
                setattr(self, f, torch.zeros(self.n, dtype=self.dtype, device=device))
        if self.nc is None:
            self.nc = torch.zeros(self.n, dtype=torch.int32, device=device)
        if self.wh is None:
            wh64, wh32, K = kernel_tables(self.sincIndex)
            self.K = K
            self.wh = torch.from_numpy(wh64 if self.precision == FP64 else wh32).to(device)
        return self

    def struct(self) -> Particles:
        """The sphx_particles view; every present field is checked (dtype, contiguity, length,
        device) because the kernels reinterpret the raw pointers."""
        s = Particles()
        s.n, s.precision, s.ng0, s.ngmax, s.Kcour, s.K, s.gamma = (self.n, self.precision, self.ng0, self.ngmax,
                                                                   self.Kcour, self.K, self.gamma)
        key = (self.precision, self.n)
        if getattr(self, "_spec_key", None) != key:
            fields = ("x", "y", "z", "nc", "wh") + self.HYDRO + self.INTEGRATION
            self._spec = tuple((f, _F64 if f in ("x", "y", "z") else (_I32 if f == "nc" else self.dtype),
                                _lib.TABLE_SIZE if f == "wh" else self.n) for f in fields)
            self._spec_key = key
        # one pass over the fields; the careful pass below only runs to name what is wrong
        for f, want, n in self._spec:
            t = getattr(self, f)
            if t is None:
                setattr(s, f, None)
            elif t.dtype is want and t.is_cuda and t.numel() >= n and t.is_contiguous():
                setattr(s, f, t.data_ptr())
            else:
                for g, w, m in self._spec:  # layout of every field first, then the device
                    _dev(getattr(self, g), w, m, f"ParticlesData.{g}", optional=True, on_device=False)
                for g, w, m in self._spec:
                    _dev(getattr(self, g), w, m, f"ParticlesData.{g}", optional=True)
        s.minDtCourant = self._min_dt if self._min_dt_ctx is None else math.nan
        return s


@dataclass
class TargetGroups:
    """Target groups over [first, last): group g is [bounds[g], bounds[g+1]) (cstone::GroupData,
    domain/include/cstone/traversal/groups.hpp)."""

    first: int
    last: int
    numGroups: int
    bounds: torch.Tensor  # int32[numGroups + 1]

    def struct(self) -> Groups:
        p = self.bounds.data_ptr()
        return Groups(self.first, self.last, self.numGroups, p, p + 4)


def compute_groups(ctx: Context, x, y, z, h, octree: OctreeData, box: Box, tol_factor: float = 2.0, first=0,
                   last=None) -> TargetGroups:
    """cstone::computeGroupSplits with 64-particle groups (traversal/groups_gpu.cu:108-137), as
    called by sph::computeSpatialGroups with tolFactor 2 (sph/groups.hpp). h may be None
    (the spatial criterion alone); its dtype selects FP64 or mixed fields."""
    last = x.numel() if last is None else last
    prec = MIXED if (h is not None and h.dtype == torch.float32) else FP64
    v = octree.view()
    g = Groups()
    m = x.numel()
    ctx.check(ctx.lib.sphx_compute_groups(ctx.h, first, last, _dev(x, _F64, m, "x"), _dev(y, _F64, m, "y"),
                                          _dev(z, _F64, m, "z"), _dev(h, (_F64, _F32), m, "h", True), prec,
                                          C.byref(v), C.byref(box), tol_factor, C.byref(g)))
    bounds = torch.empty(g.numGroups + 1, dtype=torch.int32, device=x.device)
    if g.numGroups > 0:
        ctx.d2d(bounds, g.groupStart, 4 * (g.numGroups + 1))
    else:
        bounds.fill_(last)
    return TargetGroups(first, last, g.numGroups, bounds)


def _groups_struct(groups, d: ParticlesData):
    if groups is None:
        return Groups(0, d.n, (d.n + 63) // 64, None, None)
    if isinstance(groups, TargetGroups):
        return groups.struct()
    return groups


def find_neighbors_sfc(ctx: Context, d: ParticlesData, box: Box, octree: OctreeData, groups=None):
    """sph::findNeighborsSfc: neighbor search with h iteration; d.h and d.nc are updated."""
    s = d.struct()
    g = _groups_struct(groups, d)
    v = octree.view()
    ctx.check(ctx.lib.sphx_find_neighbors_sph(ctx.h, C.byref(g), C.byref(s), C.byref(box), C.byref(v)))


def export_neighbors(ctx: Context, d: ParticlesData):
    out = torch.zeros(d.n * d.ngmax, dtype=torch.int32, device=d.x.device)
    ctx.check(ctx.lib.sphx_export_neighbors(ctx.h, _ptr(out)))
    return out.view(d.n, d.ngmax)


def compute_density(ctx: Context, d: ParticlesData, box: Box, groups=None):
    s = d.struct()
    ctx.check(ctx.lib.sphx_compute_density(ctx.h, C.byref(_groups_struct(groups, d)), C.byref(s), C.byref(box)))


def compute_eos_hydro_std(ctx: Context, d: ParticlesData, first=0, last=None):
    s = d.struct()
    ctx.check(ctx.lib.sphx_compute_eos(ctx.h, first, d.n if last is None else last, C.byref(s)))


def compute_iad(ctx: Context, d: ParticlesData, box: Box, groups=None):
    s = d.struct()
    ctx.check(ctx.lib.sphx_compute_iad(ctx.h, C.byref(_groups_struct(groups, d)), C.byref(s), C.byref(box)))


def compute_momentum_energy_std(ctx: Context, d: ParticlesData, box: Box, groups=None):
    """Sets d.minDtCourant (read from the device when first used)."""
    def call():
        s = d.struct()
        ctx.check(ctx.lib.sphx_compute_momentum_energy(ctx.h, C.byref(_groups_struct(groups, d)), C.byref(s),
                                                       C.byref(box)))
        return s
    ctx._min_dt_call(d, call)


def compute_hydro_std(ctx: Context, d: ParticlesData, box: Box, groups=None):
    """Density, EOS, IAD and momentum-energy in one call, the order of HydroProp::computeForces
    (main/src/propagator/std_hydro.hpp:139-153); same results as the four calls, with each
    loop writing the next loop's neighbor records when the groups cover every particle.  Sets
    d.minDtCourant (read from the device when first used)."""
    def call():
        s = d.struct()
        ctx.check(ctx.lib.sphx_compute_hydro_std(ctx.h, C.byref(_groups_struct(groups, d)), C.byref(s),
                                                 C.byref(box)))
        return s
    ctx._min_dt_call(d, call)


# ------------------------------------------------------------------------------------------- integration
@dataclass
class TimestepState:
    """The dataset's global time-step scalars (particles_data.hpp:93-111): minDt, minDt_m1 and
    ttot evolve; g != 0 adds the acceleration limit etaAcc (min h^2/|a|^2)^(1/4)."""

    minDt: float = 1e-12
    minDt_m1: float = 1e-12
    ttot: float = 0.0
    g: float = 0.0
    etaAcc: float = 0.2
    minDtRho: float = math.inf
    minDtAcc: float = math.inf

    def struct(self, min_dt_courant: float) -> Timestep:
        return Timestep(self.minDt, self.minDt_m1, self.ttot, min_dt_courant, self.minDtRho, self.g, self.etaAcc,
                        self.minDtAcc)


def compute_timestep(ctx: Context, d: ParticlesData, ts: TimestepState, first=0, last=None):
    """sph::computeTimestep on one rank: ts.minDt becomes min(acceleration limit if g != 0,
    d.minDtCourant, ts.minDtRho, 1.1 ts.minDt); ts.minDt_m1 and ts.ttot advance."""
    s = d.struct()
    t = ts.struct(d.minDtCourant)
    ctx.check(ctx.lib.sphx_compute_timestep(ctx.h, first, d.n if last is None else last, C.byref(s), C.byref(t)))
    ts.minDt, ts.minDt_m1, ts.ttot, ts.minDtAcc = t.minDt, t.minDt_m1, t.ttot, t.minDtAcc
    return ts.minDt


def compute_positions(ctx: Context, d: ParticlesData, box: Box, ts: TimestepState, groups=None):
    """sph::computePositions (CPU semantics): Press position update, Adams-Bashforth energy
    update, du_m1 = du, with ts.minDt and ts.minDt_m1."""
    s = d.struct()
    t = ts.struct(d.minDtCourant)
    ctx.check(ctx.lib.sphx_compute_positions(ctx.h, C.byref(_groups_struct(groups, d)), C.byref(s), C.byref(box),
                                             C.byref(t)))


def update_smoothing_length(ctx: Context, d: ParticlesData, keys=None, groups=None) -> bool:
    """sph::updateSmoothingLength: h from nc; particles with nc <= 1 get removeKey in keys.
    Returns whether any key was removed."""
    s = d.struct()
    r = C.c_int()
    ctx.check(ctx.lib.sphx_update_smoothing_length(ctx.h, C.byref(_groups_struct(groups, d)), C.byref(s),
                                                   _dev(keys, _I64, d.n, "keys", True), C.byref(r)))
    return bool(r.value)


# ------------------------------------------------------------------------------------------- VE hydro
def kernel_derivative_tables(sinc_index: float = 6.0):
    """(whd64 numpy, whd32 numpy): the kernel-derivative table of createTables."""
    lib = _lib.load()
    whd64 = np.zeros(_lib.TABLE_SIZE, np.float64)
    whd32 = np.zeros(_lib.TABLE_SIZE, np.float32)
    lib.sphx_kernel_derivative_tables(sinc_index, whd64.ctypes.data_as(C.c_void_p), whd32.ctypes.data_as(C.c_void_p))
    return whd64, whd32


@dataclass
class VeData:
    """Device fields of the VE loops that ParticlesData does not hold (particles_data.hpp:218-246)
    and the VE constants (alphamin, alphamax, decay_constant :127-129; Atmin, Atmax, ramp
    :134-136).  grad_v allocates the symmetrised velocity gradient dV11..dV33."""

    n: int
    precision: int = FP64
    alphamin: float = 0.05
    alphamax: float = 1.0
    decay: float = 0.2
    Atmin: float = 0.1
    Atmax: float = 0.2
    ramp: float = 1.0 / (0.2 - 0.1)
    av_clean: bool = False
    xm: torch.Tensor | None = None
    kx: torch.Tensor | None = None
    gradh: torch.Tensor | None = None
    prho: torch.Tensor | None = None
    divv: torch.Tensor | None = None
    curlv: torch.Tensor | None = None
    alpha: torch.Tensor | None = None
    dV11: torch.Tensor | None = None
    dV12: torch.Tensor | None = None
    dV13: torch.Tensor | None = None
    dV22: torch.Tensor | None = None
    dV23: torch.Tensor | None = None
    dV33: torch.Tensor | None = None
    whd: torch.Tensor | None = None

    FIELDS = ("xm", "kx", "gradh", "prho", "divv", "curlv", "alpha")
    GRAD_V = ("dV11", "dV12", "dV13", "dV22", "dV23", "dV33")

    @property
    def dtype(self):
        return torch.float64 if self.precision == FP64 else torch.float32

    def allocate(self, device, grad_v=False, sinc_index=6.0):
        for f in self.FIELDS + (self.GRAD_V if grad_v else ()):
            if getattr(self, f) is None:
                setattr(self, f, torch.zeros(self.n, dtype=self.dtype, device=device))
        if self.alpha is not None and not bool(self.alpha.any()):
            self.alpha.fill_(self.alphamin)
        if self.whd is None:
            w64, w32 = kernel_derivative_tables(sinc_index)
            self.whd = torch.from_numpy(w64 if self.precision == FP64 else w32).to(device)
        return self

    def struct(self) -> VeFields:
        v = VeFields()
        for f in self.FIELDS + self.GRAD_V + ("whd",):
            p = _dev(getattr(self, f), self.dtype, _lib.TABLE_SIZE if f == "whd" else self.n, f"VeData.{f}",
                     optional=True)
            setattr(v, f, p.value if p is not None else None)
        v.alphamin, v.alphamax, v.decay = self.alphamin, self.alphamax, self.decay
        v.Atmin, v.Atmax, v.ramp = self.Atmin, self.Atmax, self.ramp
        v.avClean = int(bool(self.av_clean))
        return v


def _ve_call(ctx, fn, d: ParticlesData, ve: VeData, box: Box, groups, *extra):
    s = d.struct()
    v = ve.struct()
    ctx.check(fn(ctx.h, C.byref(_groups_struct(groups, d)), C.byref(s), C.byref(v), C.byref(box), *extra))
    return s


def compute_xmass(ctx: Context, d: ParticlesData, ve: VeData, box: Box, groups=None):
    _ve_call(ctx, ctx.lib.sphx_compute_xmass, d, ve, box, groups)


def compute_ve_def_gradh(ctx: Context, d: ParticlesData, ve: VeData, box: Box, groups=None):
    _ve_call(ctx, ctx.lib.sphx_compute_ve_def_gradh, d, ve, box, groups)


def compute_eos_ve(ctx: Context, d: ParticlesData, ve: VeData, first=0, last=None):
    s = d.struct()
    v = ve.struct()
    ctx.check(ctx.lib.sphx_compute_eos_ve(ctx.h, first, d.n if last is None else last, C.byref(s), C.byref(v)))


def compute_iad_divv_curlv(ctx: Context, d: ParticlesData, ve: VeData, box: Box, groups=None):
    _ve_call(ctx, ctx.lib.sphx_compute_iad_divv_curlv, d, ve, box, groups)


def compute_av_switches(ctx: Context, d: ParticlesData, ve: VeData, box: Box, dt: float, groups=None):
    _ve_call(ctx, ctx.lib.sphx_compute_av_switches, d, ve, box, groups, C.c_double(dt))


def rho_timestep(ctx: Context, d: ParticlesData, ve: VeData, Krho: float = 0.06, first=0, last=None) -> float:
    """sph::rhoTimestep (sph/include/sph/ts_global.hpp:73-98): Krho / |max divv|."""
    s = d.struct()
    v = ve.struct()
    out = C.c_double()
    ctx.check(ctx.lib.sphx_rho_timestep(ctx.h, first, d.n if last is None else last, C.byref(s), C.byref(v), Krho,
                                        C.byref(out)))
    return out.value


def compute_momentum_energy_ve(ctx: Context, d: ParticlesData, ve: VeData, box: Box, groups=None):
    """Sets d.minDtCourant (read from the device when first used)."""
    ctx._min_dt_call(d, lambda: _ve_call(ctx, ctx.lib.sphx_compute_momentum_energy_ve, d, ve, box, groups))


# ------------------------------------------------------------------------------------------- adjoints
def density_adjoint(ctx: Context, d: ParticlesData, box: Box, g_rho, groups=None, want_h=True):
    """Adjoint of compute_density on the context's lists: (dL/dx, dL/dy, dL/dz[, dL/dh]) for
    g_rho = dL/drho (no counterpart in the reference)."""
    dev = d.x.device
    g_rho = g_rho.to(d.dtype).contiguous()
    gx, gy, gz = (torch.empty(d.n, dtype=torch.float64, device=dev) for _ in range(3))
    gh = torch.empty(d.n, dtype=d.dtype, device=dev) if want_h else None
    s = d.struct()
    ctx.check(ctx.lib.sphx_density_adjoint(ctx.h, C.byref(_groups_struct(groups, d)), C.byref(s), C.byref(box),
                                           _dev(g_rho, d.dtype, d.n, "g_rho"), _ptr(gx), _ptr(gy), _ptr(gz),
                                           _ptr(gh)))
    return (gx, gy, gz, gh) if want_h else (gx, gy, gz)


def momentum_energy_adjoint(ctx: Context, d: ParticlesData, box: Box, g_ax=None, g_ay=None, g_az=None, g_du=None,
                            groups=None):
    """Adjoint of compute_momentum_energy_std with respect to x, y, z (v, h, m, rho, p, c and
    the IAD tensors held fixed): (dL/dx, dL/dy, dL/dz) for the given output gradients."""
    dev = d.x.device
    conv = lambda t: None if t is None else t.to(d.dtype).contiguous()
    g_ax, g_ay, g_az, g_du = conv(g_ax), conv(g_ay), conv(g_az), conv(g_du)
    gx, gy, gz = (torch.empty(d.n, dtype=torch.float64, device=dev) for _ in range(3))
    s = d.struct()
    ctx.check(ctx.lib.sphx_momentum_energy_adjoint(ctx.h, C.byref(_groups_struct(groups, d)), C.byref(s),
                                                   C.byref(box), *(_dev(g, d.dtype, d.n, nm, True) for g, nm in
                                                                   ((g_ax, "g_ax"), (g_ay, "g_ay"), (g_az, "g_az"),
                                                                    (g_du, "g_du"))),
                                                   _ptr(gx), _ptr(gy), _ptr(gz)))
    return gx, gy, gz


# ------------------------------------------------------------------------------------------- the step
STAGES = ("box", "keys", "sort", "gather", "octree", "link", "centers", "groups", "neighbors", "density", "eos",
          "iad", "momentum_energy")


class HydroStep:
    """One single-rank std-hydro force evaluation from unsorted particles, in the stage
    order of HydroProp::computeForces (main/src/propagator/std_hydro.hpp:119-168) with
    the single-rank domain sync (keys, sort, permute, cornerstone tree, link, centres).

    ``compute_forces(inputs)`` takes unsorted device tensors x, y, z (float64) and
    h, m, vx, vy, vz, u (precision dtype), and leaves the SFC-sorted results in
    ``self.d`` (a ParticlesData), the tree in ``self.octree`` and the target groups in
    ``self.groups``.  ``spatial_groups`` selects computeSpatialGroups' split groups (the
    reference's GPU path, sph/groups.hpp:42-67) over plain 64-particle groups.
    """

    def __init__(self, ctx: Context, n: int, box: Box, precision=FP64, bucket_size=64, ng0=100, ngmax=150,
                 sinc_index=6.0, gamma=5.0 / 3.0, Kcour=0.2, spatial_groups=True):
        self.ctx, self.n, self.box, self.bucket = ctx, n, box, bucket_size
        self.spatial_groups = spatial_groups
        self.groups = None
        self.d = ParticlesData(n, precision, ng0, ngmax, Kcour, gamma, sinc_index).allocate(ctx.device)
        self.keys = torch.empty(n, dtype=torch.int64, device=ctx.device)
        self.order = torch.empty(n, dtype=torch.int32, device=ctx.device)
        self.octree = None
        self.events = None
        self.box_calls = 0
        self.side = torch.cuda.Stream(device=ctx.device)  # default priority

    def gather_inputs(self, pairs, overlap: bool):
        """The permutation of the inputs (gatherArrays).  overlap: on a second stream, so
        that it runs beside the tree build (which reads only the keys); the returned
        callable makes the context's stream wait for it before the fields are read."""
        ctx = self.ctx
        if not overlap:
            gather_arrays(ctx, self.order, pairs)
            return lambda: None
        sorted_ = torch.cuda.Event()
        sorted_.record(ctx.stream)
        self.side.wait_event(sorted_)
        with ctx.on_stream(self.side):
            gather_arrays(ctx, self.order, pairs)
        gathered = torch.cuda.Event()
        gathered.record(self.side)
        return lambda: ctx.stream.wait_event(gathered)

    def sync_box(self, inp: dict) -> Box:
        """The box of this call, as Domain::sync on one rank (domain/assignment.hpp:102-106):
        periodic and fixed dimensions keep their limits; open dimensions take the particles'
        extrema (makeGlobalBox) on the first call, limited by limitBoxShrinking against the
        previous call's box afterwards."""
        box = self.box
        if not (box.bx and box.by and box.bz):
            fitted = make_global_box(self.ctx, inp["x"], inp["y"], inp["z"], box)
            box = self.box = fitted if self.box_calls == 0 else limit_box_shrinking(fitted, box)
        self.box_calls += 1
        return box

    def compute_forces(self, inp: dict, timed=False):
        ctx, d = self.ctx, self.d
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(STAGES) + 1)] if timed else None
        mark = (lambda k: ev[k].record(ctx.stream)) if timed else (lambda k: None)
        mark(0)
        box = self.sync_box(inp)
        mark(1)
        self.keys.zero_()
        compute_sfc_keys(ctx, inp["x"], inp["y"], inp["z"], box, self.keys)
        mark(2)
        ctx.check(ctx.lib.sphx_sequence(ctx.h, _ptr(self.order), self.n, 0))
        sort_by_key(ctx, self.keys, self.order)
        mark(3)
        join = self.gather_inputs([(inp[f], getattr(d, f)) for f in
                                   ("x", "y", "z", "h", "m", "vx", "vy", "vz", "u") + ParticlesData.INTEGRATION
                                   if f in inp], overlap=not timed)
        mark(4)
        leaves, counts, layout = compute_octree(ctx, self.keys, self.bucket)
        mark(5)
        o = build_octree(ctx, leaves)
        o.layout, o.counts = layout, counts
        mark(6)
        node_fp_centers(ctx, o, box)
        self.octree = o
        join()
        mark(7)
        g = compute_groups(ctx, d.x, d.y, d.z, d.h, o, box) if self.spatial_groups else None
        self.groups = g
        mark(8)
        find_neighbors_sfc(ctx, d, box, o, g)
        mark(9)
        if timed:  # the four calls, for the per-stage times
            compute_density(ctx, d, box, g)
            mark(10)
            compute_eos_hydro_std(ctx, d)
            mark(11)
            compute_iad(ctx, d, box, g)
            mark(12)
            compute_momentum_energy_std(ctx, d, box, g)
            mark(13)
        else:
            compute_hydro_std(ctx, d, box, g)
        self.events = ev
        return d

    def integrate(self, ts: TimestepState, remove_unconverged=False):
        """HydroProp::integrate (main/src/propagator/std_hydro.hpp:170-186): time step, positions
        and energy, smoothing lengths, on the SFC-ordered fields of the last compute_forces
        (which must have received x_m1, y_m1, z_m1 and du_m1).  Raises, as the reference does,
        when a particle is left with at most itself as neighbor, unless remove_unconverged."""
        ctx, d = self.ctx, self.d
        compute_timestep(ctx, d, ts)
        compute_positions(ctx, d, self.box, ts, self.groups)
        removed = update_smoothing_length(ctx, d, self.keys, self.groups)
        if removed and not remove_unconverged:
            raise RuntimeError("Neighbor search did not converge")
        return ts

    def momentum_energy_vjp(self, g_ax=None, g_ay=None, g_az=None, g_du=None, recompute_lists=True):
        """dL/d(x, y, z), in SFC order, of the last compute_forces' momentum-energy outputs for the
        given output gradients (rho, p, c and the IAD tensors held fixed).  recompute_lists re-runs
        the neighbor search from the final smoothing lengths first, as a backward pass that does
        not keep the forward's lists would; it reproduces the forward's lists unless a particle
        ended the forward's h iteration unconverged."""
        if recompute_lists:
            find_neighbors_sfc(self.ctx, self.d, self.box, self.octree, self.groups)
        return momentum_energy_adjoint(self.ctx, self.d, self.box, g_ax, g_ay, g_az, g_du, self.groups)

    def stage_times_ms(self):
        """Per-stage milliseconds of the last timed compute_forces (synchronises)."""
        if not self.events:
            return {}
        self.events[-1].synchronize()
        return {s: self.events[k].elapsed_time(self.events[k + 1]) for k, s in enumerate(STAGES)}


VE_STAGES = ("box", "keys", "sort", "gather", "octree", "link", "centers", "groups", "neighbors", "xmass",
             "gradh", "eos", "iad_divv_curlv", "av_switches", "momentum_energy")


class VeHydroStep(HydroStep):
    """One single-rank force evaluation of the VE propagator (HydroVeProp::computeForces,
    main/src/propagator/ve_hydro.hpp:126-205): the same sync, tree, groups and neighbor
    search as HydroStep, then xmass, veDefGradh, EOS, iadDivvCurlv with rhoTimestep
    (``self.minDtRho``), AVswitches (over the time step ``dt``) and the VE momentum-energy
    loop (with AV cleaning when ``av_clean``).  ``self.ve`` holds the VE fields;
    alpha is conserved: pass it in ``inputs`` (unsorted) or it starts at alphamin."""

    def __init__(self, ctx: Context, n: int, box: Box, precision=FP64, grad_v=False, av_clean=False, **kw):
        super().__init__(ctx, n, box, precision, **kw)
        self.ve = VeData(n, precision, av_clean=av_clean).allocate(ctx.device, grad_v=grad_v or av_clean)
        self.minDtRho = math.inf

    def compute_forces(self, inp: dict, timed=False, dt: float = 1e-6):
        ctx, d = self.ctx, self.d
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(VE_STAGES) + 1)] if timed else None
        mark = (lambda k: ev[k].record(ctx.stream)) if timed else (lambda k: None)
        mark(0)
        box = self.sync_box(inp)
        mark(1)
        self.keys.zero_()
        compute_sfc_keys(ctx, inp["x"], inp["y"], inp["z"], box, self.keys)
        mark(2)
        ctx.check(ctx.lib.sphx_sequence(ctx.h, _ptr(self.order), self.n, 0))
        sort_by_key(ctx, self.keys, self.order)
        mark(3)
        pairs = [(inp[f], getattr(d, f)) for f in ("x", "y", "z", "h", "m", "vx", "vy", "vz", "u") +
                 ParticlesData.INTEGRATION if f in inp]
        if "alpha" in inp:
            pairs.append((inp["alpha"], self.ve.alpha))
        join = self.gather_inputs(pairs, overlap=not timed)
        mark(4)
        leaves, counts, layout = compute_octree(ctx, self.keys, self.bucket)
        mark(5)
        o = build_octree(ctx, leaves)
        o.layout, o.counts = layout, counts
        mark(6)
        node_fp_centers(ctx, o, box)
        self.octree = o
        join()
        mark(7)
        g = compute_groups(ctx, d.x, d.y, d.z, d.h, o, box) if self.spatial_groups else None
        self.groups = g
        mark(8)
        find_neighbors_sfc(ctx, d, box, o, g)
        mark(9)
        ve = self.ve
        compute_xmass(ctx, d, ve, box, g)
        mark(10)
        compute_ve_def_gradh(ctx, d, ve, box, g)
        mark(11)
        compute_eos_ve(ctx, d, ve)
        mark(12)
        compute_iad_divv_curlv(ctx, d, ve, box, g)
        self.minDtRho = rho_timestep(ctx, d, ve)
        mark(13)
        compute_av_switches(ctx, d, ve, box, dt, g)
        mark(14)
        compute_momentum_energy_ve(ctx, d, ve, box, g)
        mark(15)
        self.events = ev
        return d

    def stage_times_ms(self):
        if not self.events:
            return {}
        self.events[-1].synchronize()
        return {s: self.events[k].elapsed_time(self.events[k + 1]) for k, s in enumerate(VE_STAGES)}


class Simulation:
    """Repeated single-rank time steps of the std-hydro propagator: each ``step()`` re-sorts
    the particles (keys, sort, permutation of every conserved field, as Domain::sync does on one
    rank, main/src/propagator/std_hydro.hpp:103-116 with ConservedFields :72), evaluates the
    forces and integrates.  The state stays in HBM in SFC order between steps; two buffer sets
    alternate so that the permutation never gathers in place."""

    CONSERVED = ("x", "y", "z", "h", "m", "vx", "vy", "vz", "u") + ParticlesData.INTEGRATION

    def __init__(self, ctx: Context, state: dict, box: Box, precision=FP64, ts: TimestepState | None = None,
                 propagator: str = "std", **step_kw):
        """propagator "std" (HydroProp) or "ve" (HydroVeProp: alpha is conserved too, and the
        time step includes rhoTimestep).  Both carry the specific internal energy u (the VE
        propagator's temperature is u / cv with a constant cv)."""
        n = state["x"].numel()
        if propagator not in ("std", "ve"):
            raise ValueError(f"unknown propagator {propagator!r}")
        self.ve = propagator == "ve"
        self.step_obj = (VeHydroStep if self.ve else HydroStep)(ctx, n, box, precision, **step_kw)
        if self.ve:
            self.CONSERVED = Simulation.CONSERVED + ("alpha",)
        d = self.step_obj.d
        d.allocate(ctx.device, integration=True)
        dt = d.dtype
        self.state = {}
        for f in self.CONSERVED:
            want = torch.float64 if f in ("x", "y", "z") else dt
            src = state.get(f)
            if src is None and f == "alpha":
                src = torch.full((n,), self.step_obj.ve.alphamin, dtype=want, device=ctx.device)
            self.state[f] = (torch.zeros(n, dtype=want, device=ctx.device) if src is None
                             else src.to(device=ctx.device, dtype=want).clone())
        self.ts = ts if ts is not None else TimestepState()
        self.steps = 0

    @property
    def d(self) -> ParticlesData:
        return self.step_obj.d

    def step(self):
        st = self.step_obj
        if self.ve:
            # AV switches over the current step (d.minDt, ve_hydro.hpp:172); rhoTimestep (:166) feeds
            # computeTimestep
            st.compute_forces(self.state, dt=self.ts.minDt)
            self.ts.minDtRho = st.minDtRho
        else:
            st.compute_forces(self.state)
        st.integrate(self.ts)
        # the integrated SFC-ordered fields become the state; the old state buffers receive the
        # next step's permutation
        for f in self.CONSERVED:
            holder = st.ve if f == "alpha" else st.d
            cur = getattr(holder, f)
            setattr(holder, f, self.state[f])
            self.state[f] = cur
        self.steps += 1
        return self.ts
