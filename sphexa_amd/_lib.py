"""ctypes declarations of libsphx.so (include/sphx.h).

The library is built in-tree (sphexa_amd/libsphx.so) by ``__graft_entry__.build()``
or ``make -C sphexa_amd/csrc``.  Loading fails loudly when it is missing: there is
no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; SPHX_LIBRARY names another build of the same library (e.g. a
# variant compiled with different kernel options) to load instead
LIB_PATH = os.environ.get("SPHX_LIBRARY") or os.path.join(HERE, "libsphx.so")

SPHX_OK, SPHX_ERR_HIP, SPHX_ERR_ARG, SPHX_ERR_STACK, SPHX_ERR_NGMAX, SPHX_ERR_STATE, SPHX_ERR_NOMEM, \
    SPHX_ERR_NODEV = range(8)
BC_OPEN, BC_PERIODIC, BC_FIXED = 0, 1, 2
FP64, MIXED = 0, 1
OPT_FP64_EXACT = 1
OPT_TIMING = 2
OPT_LOOP_SHAPE_FP64 = 3
OPT_LOOP_SHAPE_MIXED = 4
OPT_DEFER_MIN_DT = 5
STAT_CANDIDATE_TESTS = 1
STAT_SCRATCH_BYTES = 2
GATHER_MAX_ARRAYS = 16
TABLE_SIZE = 20000
MAX_LEVEL = 21

ERROR_NAMES = {SPHX_ERR_HIP: "HIP", SPHX_ERR_ARG: "ARG", SPHX_ERR_STACK: "STACK", SPHX_ERR_NGMAX: "NGMAX",
               SPHX_ERR_STATE: "STATE", SPHX_ERR_NOMEM: "NOMEM", SPHX_ERR_NODEV: "NODEV"}

# every function the header declares, with its ctypes signature
_V, _U32, _U64, _I32, _D, _F = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_double, C.c_float


class Box(C.Structure):
    _fields_ = [("xmin", _D), ("xmax", _D), ("ymin", _D), ("ymax", _D), ("zmin", _D), ("zmax", _D),
                ("bx", _I32), ("by", _I32), ("bz", _I32), ("reserved", _I32)]


class Octree(C.Structure):
    _fields_ = [("numLeafNodes", _I32), ("numInternalNodes", _I32), ("numNodes", _I32), ("reserved", _I32),
                ("prefixes", _V), ("childOffsets", _V), ("parents", _V), ("levelRange", _V),
                ("internalToLeaf", _V), ("leafToInternal", _V), ("leaves", _V), ("layout", _V),
                ("centers", _V), ("sizes", _V)]


class Groups(C.Structure):
    _fields_ = [("firstBody", _U32), ("lastBody", _U32), ("numGroups", _U32), ("groupStart", _V),
                ("groupEnd", _V)]


class Particles(C.Structure):
    _fields_ = [("n", _U64), ("precision", _I32), ("ng0", _U32), ("ngmax", _U32), ("Kcour", _F), ("K", _D),
                ("gamma", _D), ("x", _V), ("y", _V), ("z", _V), ("h", _V), ("m", _V), ("vx", _V), ("vy", _V),
                ("vz", _V), ("u", _V), ("rho", _V), ("p", _V), ("c", _V), ("c11", _V), ("c12", _V), ("c13", _V),
                ("c22", _V), ("c23", _V), ("c33", _V), ("ax", _V), ("ay", _V), ("az", _V), ("du", _V), ("nc", _V),
                ("wh", _V), ("minDtCourant", _D), ("x_m1", _V), ("y_m1", _V), ("z_m1", _V), ("du_m1", _V)]


class VeFields(C.Structure):
    _fields_ = [(nm, _V) for nm in ("xm", "kx", "gradh", "prho", "divv", "curlv", "alpha", "dV11", "dV12", "dV13",
                                    "dV22", "dV23", "dV33", "whd")] + \
               [(nm, _D) for nm in ("alphamin", "alphamax", "decay", "Atmin", "Atmax", "ramp")] + \
               [("avClean", _I32), ("reserved", _I32)]


class Timestep(C.Structure):
    _fields_ = [("minDt", _D), ("minDt_m1", _D), ("ttot", _D), ("minDtCourant", _D), ("minDtRho", _D), ("g", _D),
                ("etaAcc", _D), ("minDtAcc", _D)]


SIGNATURES = {
    "sphx_ctx_create": (C.c_int, [C.c_int, _V, C.POINTER(_V)]),
    "sphx_ctx_destroy": (C.c_int, [_V]),
    "sphx_ctx_set_stream": (C.c_int, [_V, _V]),
    "sphx_ctx_synchronize": (C.c_int, [_V]),
    "sphx_ctx_set_option": (C.c_int, [_V, C.c_int, C.c_int]),
    "sphx_last_error": (C.c_char_p, [_V]),
    "sphx_version": (C.c_char_p, []),
    "sphx_kernel_tables": (C.c_int, [_D, _V, _V, _V]),
    "sphx_compute_sfc_keys": (C.c_int, [_V, _V, _V, _V, _V, _U64, C.POINTER(Box)]),
    "sphx_sort_by_key": (C.c_int, [_V, _V, _V, _U64]),
    "sphx_sort_by_key_noalloc": (C.c_int, [_V, _V, _V, _U64]),
    "sphx_sequence": (C.c_int, [_V, _V, _U64, _U32]),
    "sphx_gather_f64": (C.c_int, [_V, _V, _V, _V, _U64]),
    "sphx_gather_f32": (C.c_int, [_V, _V, _V, _V, _U64]),
    "sphx_gather_u32": (C.c_int, [_V, _V, _V, _V, _U64]),
    "sphx_gather_arrays": (C.c_int, [_V, _V, _U64, C.c_int32, _V, _V, _V]),
    "sphx_compute_octree": (C.c_int, [_V, _V, _U64, _U32, _U32, C.POINTER(_I32), C.POINTER(_V), C.POINTER(_V),
                                      C.POINTER(_V)]),
    "sphx_octree_sizes": (C.c_int, [_I32, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32)]),
    "sphx_build_octree": (C.c_int, [_V, _V, C.POINTER(Octree)]),
    "sphx_node_centers": (C.c_int, [_V, _V, _I32, C.POINTER(Box), _V, _V]),
    "sphx_compute_groups": (C.c_int, [_V, _U32, _U32, _V, _V, _V, _V, _I32, C.POINTER(Octree), C.POINTER(Box),
                                      _F, C.POINTER(Groups)]),
    "sphx_find_neighbors": (C.c_int, [_V, _V, _V, _V, _V, _U32, _U32, C.POINTER(Box), C.POINTER(Octree), _U32,
                                      _V, _V]),
    "sphx_find_neighbors_sph": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(Box),
                                          C.POINTER(Octree)]),
    "sphx_export_neighbors": (C.c_int, [_V, _V]),
    "sphx_compute_density": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(Box)]),
    "sphx_compute_eos": (C.c_int, [_V, _U32, _U32, C.POINTER(Particles)]),
    "sphx_compute_iad": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(Box)]),
    "sphx_compute_momentum_energy": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(Box)]),
    "sphx_compute_hydro_std": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(Box)]),
    "sphx_min_dt_courant": (C.c_int, [_V, C.POINTER(_D)]),
    "sphx_density_adjoint": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(Box), _V, _V, _V,
                                       _V, _V]),
    "sphx_momentum_energy_adjoint": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(Box), _V,
                                               _V, _V, _V, _V, _V, _V]),
    "sphx_compute_timestep": (C.c_int, [_V, _U32, _U32, C.POINTER(Particles), C.POINTER(Timestep)]),
    "sphx_compute_positions": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(Box),
                                         C.POINTER(Timestep)]),
    "sphx_update_smoothing_length": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), _V,
                                               C.POINTER(C.c_int)]),
    "sphx_kernel_derivative_tables": (C.c_int, [_D, _V, _V]),
    "sphx_smoothing_length_factors": (C.c_int, [_U32, _U32, _V, _V]),
    "sphx_compute_xmass": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(VeFields),
                                     C.POINTER(Box)]),
    "sphx_compute_ve_def_gradh": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(VeFields),
                                            C.POINTER(Box)]),
    "sphx_compute_eos_ve": (C.c_int, [_V, _U32, _U32, C.POINTER(Particles), C.POINTER(VeFields)]),
    "sphx_compute_iad_divv_curlv": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(VeFields),
                                              C.POINTER(Box)]),
    "sphx_compute_av_switches": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles), C.POINTER(VeFields),
                                           C.POINTER(Box), _D]),
    "sphx_compute_momentum_energy_ve": (C.c_int, [_V, C.POINTER(Groups), C.POINTER(Particles),
                                                  C.POINTER(VeFields), C.POINTER(Box)]),
    "sphx_rho_timestep": (C.c_int, [_V, _U32, _U32, C.POINTER(Particles), C.POINTER(VeFields), _D,
                                    C.POINTER(_D)]),
    "sphx_make_global_box": (C.c_int, [_V, _V, _V, _V, _U64, C.POINTER(Box), C.POINTER(Box)]),
    "sphx_get_stat": (C.c_int, [_V, C.c_int, C.POINTER(_D)]),
    "sphx_timing_summary": (C.c_int, [_V, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_int), C.POINTER(_D),
                                      C.POINTER(C.c_int)]),
}

_lib = None


class SphxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sphx error {ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load libsphx.so once; raise if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built: run __graft_entry__.build() or make -C sphexa_amd/csrc")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols(path: str = LIB_PATH):
    """Names of the SIGNATURES entries that the library exports (no GPU needed)."""
    lib = C.CDLL(path)
    return [n for n in SIGNATURES if hasattr(lib, n)]
