"""CPU checks of the drop-in boundary: libsphx.so loads and exports every function that
include/sphx.h declares, the host-side tables equal the oracle's bit for bit, argument
errors come back as status codes (no GPU call), and include/sphx_compat.hpp compiles
against mock types and against the reference's own types."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from sphexa_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sphx.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sphx_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libsphx.so not built"
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    unbound = [n for n in names if n not in _lib.SIGNATURES]
    assert not unbound, f"declared but missing from the ctypes table: {unbound}"


def test_kernel_tables_match_oracle():
    from sphexa_amd.hydro import kernel_tables

    wh64, wh32, K = kernel_tables(6.0)
    o64, o32, oK = oracle.Cpu("oracle").kernel_tables(6.0)
    assert np.array_equal(wh64, o64) and np.array_equal(wh32, o32) and K == oK


def test_kernel_derivative_tables_match_oracle():
    """The library's host table of the kernel derivative equals the oracle's bit for bit, and
    the compiled reference's where oracle/_ref is built (its sin/cos come from the sincos
    call that g++ makes of them); the FP32 table is the FP64 one rounded."""
    from sphexa_amd.hydro import kernel_derivative_tables

    w64, w32 = kernel_derivative_tables(6.0)
    o64 = oracle.Cpu("oracle").kernel_derivative_table(6.0)
    assert np.array_equal(w64, o64)
    if oracle.available("ref"):
        assert np.array_equal(w64, oracle.Cpu("ref").kernel_derivative_table(6.0))
    assert np.max(np.abs(w32 - o64)) < 1e-6 * np.max(np.abs(o64))


@pytest.mark.parametrize("ng0", [100, 0, 37, 2047])
def test_smoothing_length_factors_match_reference(ng0):
    """The host factors every h update of the library applies, h' = (h * 0.5) * f[nc], give the
    oracle's (and the compiled reference's) updateSmoothingLengthCpu bit for bit, in double
    and in float, for every count of the table."""
    lib = _lib.load()
    cap = 1 << 17
    f64 = np.zeros(cap, np.float64)
    f32 = np.zeros(cap, np.float32)
    assert lib.sphx_smoothing_length_factors(ng0, cap, f64.ctypes.data, f32.ctypes.data) == 0
    nc = np.arange(1, cap, dtype=np.uint32)
    rng = np.random.default_rng(ng0)
    h = 0.01 + rng.random(len(nc)) * 0.02
    libs = [oracle.Cpu("oracle")] + ([oracle.Cpu("ref")] if oracle.available("ref") else [])
    for dtype, f in ((np.float64, f64), (np.float32, f32)):
        hv = h.astype(dtype)
        mine = (hv * dtype(0.5)) * f[1:]
        assert mine.dtype == dtype
        for cpu in libs:
            want, _, _ = cpu.update_smoothing_length(ng0, nc, hv, np.zeros(len(nc), np.uint64), dtype=dtype)
            assert np.array_equal(mine, want), (cpu.pre, dtype)
    # nc = 0 (never produced by a search: nc counts the particle itself): the reference's
    # division by zero, inf for ng0 > 0 and NaN for ng0 = 0, in both tables
    assert (np.isinf(f64[0]) and np.isinf(f32[0])) if ng0 else (np.isnan(f64[0]) and np.isnan(f32[0]))


def test_octree_sizes():
    lib = _lib.load()
    ni, nn, npar = C.c_int32(), C.c_int32(), C.c_int32()
    assert lib.sphx_octree_sizes(15, C.byref(ni), C.byref(nn), C.byref(npar)) == 0
    assert (ni.value, nn.value, npar.value) == (2, 17, 2)
    assert lib.sphx_octree_sizes(1, C.byref(ni), C.byref(nn), C.byref(npar)) == 0
    assert (ni.value, nn.value, npar.value) == (0, 1, 1)
    assert lib.sphx_octree_sizes(0, None, None, None) == _lib.SPHX_ERR_ARG


def test_null_context_is_an_argument_error():
    lib = _lib.load()
    assert lib.sphx_compute_sfc_keys(None, None, None, None, None, 10, None) == _lib.SPHX_ERR_ARG
    assert lib.sphx_sort_by_key(None, None, None, 10) == _lib.SPHX_ERR_ARG
    assert lib.sphx_ctx_destroy(None) == 0
    assert lib.sphx_last_error(None) == b"null context"
    assert lib.sphx_version().startswith(b"sphx")


def test_context_without_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.sphx_ctx_create(0, None, C.byref(h)) == _lib.SPHX_ERR_NODEV


MOCK = r"""
#include <cstddef>
#include <vector>
#include <array>
#include "sphx_compat.hpp"
enum class BT : char { open = 0, periodic = 1 };
struct MockBox {
    double xmin() const { return 0; } double xmax() const { return 1; }
    double ymin() const { return 0; } double ymax() const { return 1; }
    double zmin() const { return 0; } double zmax() const { return 1; }
    BT boundaryX() const { return BT::periodic; } BT boundaryY() const { return BT::periodic; }
    BT boundaryZ() const { return BT::periodic; }
};
template<class T> struct DevVec { T* p = nullptr; size_t n = 0; T* data() { return p; } size_t size() const { return n; } };
struct Data {
    DevVec<double> x, y, z; DevVec<float> h, m, vx, vy, vz, u, rho, p, c, c11, c12, c13, c22, c23, c33, ax, ay, az, du, wh;
    DevVec<unsigned> nc; unsigned ng0 = 100, ngmax = 150; double K = 0.79, gamma = 5.0 / 3.0, minDtCourant = 0; float Kcour = 0.2f;
    DevVec<float> x_m1, y_m1, z_m1, du_m1; DevVec<uint64_t> keys;
    double minDt = 1e-12, minDt_m1 = 1e-12, ttot = 0, minDtRho = 1e30, g = 0, etaAcc = 0.2;
    DevVec<float> xm, kx, gradh, prho, divv, curlv, alpha, whd, dV11, dV12, dV13, dV22, dV23, dV33;
    float alphamin = 0.05f, alphamax = 1.0f, decay_constant = 0.2f, Atmin = 0.1f, Atmax = 0.2f, ramp = 10.0f;
    double Krho = 0.06;
};
struct Group { unsigned firstBody, lastBody, numGroups; const unsigned* groupStart; const unsigned* groupEnd; };
void use(Data& d, const Group& g) {
    MockBox box;
    sphx_compat::computeDensity(g, d, box);
    sphx_compat::computeEOS_HydroStd(0, 10, d);
    sphx_compat::computeIAD(g, d, box);
    sphx_compat::computeMomentumEnergySTD(g, d, box);
    sphx_compat::computeTimestep(0, 10, d);
    sphx_compat::computePositions(g, d, box, float(d.minDt), std::array<float, 1>{float(d.minDt_m1)});
    (void)sphx_compat::updateSmoothingLength(g, d);
    sphx_compat::computeXMass(g, d, box);
    sphx_compat::computeVeDefGradh(g, d, box);
    sphx_compat::computeEOS(0, 10, d);
    sphx_compat::computeIadDivvCurlv(g, d, box);
    sphx_compat::computeAVswitches(g, d, box);
    sphx_compat::computeMomentumEnergy<false>(g, nullptr, d, box);
    sphx_compat::computeMomentumEnergy<true>(g, nullptr, d, box);
    (void)sphx_compat::rhoTimestep(0, 10, d);
    sphx_compat::computeSfcKeys(nullptr, nullptr, nullptr, nullptr, 0, box);
    (void)sphx_compat::computeGroupSplits(0u, 0u, nullptr, nullptr, nullptr, (const double*)nullptr, nullptr, 0,
                                          nullptr, box);
}
"""

REFERENCE_TYPES = r"""
#include <utility>
#include "cstone/findneighbors.hpp"
#include "cstone/traversal/groups.hpp"
#include "sphx_compat.hpp"
void use(const cstone::OctreeNsView<double, uint64_t>& view, const cstone::GroupView& g) {
    cstone::Box<double> box(0, 1, cstone::BoundaryType::periodic);
    box = sphx_compat::makeGlobalBox(nullptr, nullptr, nullptr, 0, box);
    sphx_compat::computeSfcKeys(nullptr, nullptr, nullptr, nullptr, 0, box);
    sphx_compat::findNeighbors(nullptr, nullptr, nullptr, nullptr, 0, 0, box, view, 0, 150, nullptr, nullptr);
    sphx_compat::nodeFpCenters<cstone::Vec3<double>>(nullptr, 0, nullptr, nullptr, box);
    (void)sphx_compat::toGroups(g);
    auto s = sphx_compat::computeGroupSplits(0u, 0u, nullptr, nullptr, nullptr, (const float*)nullptr, nullptr, 0,
                                             nullptr, box);
    (void)sphx_compat::toGroups(s);
}
"""


DROPIN_WITH_REFERENCE = r"""
#include <utility>
#include "cstone/findneighbors.hpp"
#include "cstone/tree/csarray.hpp"
#include "cstone/tree/octree.hpp"
#include "cstone/tree/octree_gpu.h"
#include "cstone/focus/source_center.hpp"
#include "cstone/primitives/primitives_acc.hpp"
#include "cstone/sfc/sfc_gpu.h"
#include "cstone/traversal/groups.hpp"
#include "sph/hydro_std/density.hpp"
#include "sph/hydro_std/iad.hpp"
#include "sph/hydro_std/momentum_energy.hpp"
#include "sph/hydro_std/eos.hpp"
#include "sph/hydro_ve/xmass.hpp"
#include "sph/hydro_ve/ve_def_gradh.hpp"
#include "sph/hydro_ve/eos.hpp"
#include "sph/hydro_ve/iad_divv_curlv.hpp"
#include "sph/hydro_ve/av_switches.hpp"
#include "sph/hydro_ve/momentum_energy.hpp"
#include "sph/find_neighbors.hpp"
#include "sph/groups.hpp"
#include "sph/update_h.hpp"
#include "sph/positions.hpp"
#include "sph/particles_data.hpp"
#include "sphx_dropin.hpp"
using Data = sphexa::ParticlesData<cstone::GpuTag>;
void use(Data& d, cstone::GroupData<cstone::GpuTag>& groups, const cstone::OctreeNsView<double, uint64_t>& view,
         cstone::OctreeView<uint64_t> ov, std::span<uint64_t> keys, std::span<unsigned> vals,
         cstone::DeviceVector<uint64_t>& tree, cstone::DeviceVector<unsigned>& counts,
         cstone::DeviceVector<uint64_t>& tmp, cstone::DeviceVector<int>& work, cstone::Vec3<double>* centers,
         const double* x, unsigned* nb, unsigned* nc)
{
    cstone::Box<double> box(0, 1, cstone::BoundaryType::periodic);
    cstone::computeSfcKeys<true>(x, x, x, keys.data(), keys.size(), box);
    cstone::DeviceVector<uint64_t> kb;
    cstone::DeviceVector<unsigned> vb;
    cstone::sortByKey<true>(keys, vals, kb, vb, 1.05);
    cstone::gatherAcc<true>(std::span<const unsigned>(vals), x, const_cast<double*>(x));
    auto [t, c] = cstone::computeOctree(std::span<const uint64_t>(keys), 64);
    (void)cstone::updateOctreeGpu(std::span<const uint64_t>(keys), 64, tree, counts, tmp, work);
    cstone::buildOctreeGpu(t.data(), ov);
    cstone::updateInternalTree(std::span<const uint64_t>(t), ov);
    cstone::nodeFpCenters(std::span<const uint64_t>(ov.prefixes, 10), centers, centers, box);
    cstone::findNeighbors(x, x, x, x, 0u, 10u, box, view, 150u, nb, nc);
    sph::computeGroups(0, 10, d, box, groups);
    sph::findNeighborsSfc(0, 10, d, box);
    sph::computeDensity(groups.view(), d, box);
    sph::computeEOS_HydroStd(0, 10, d);
    sph::computeIAD(groups.view(), d, box);
    sph::computeMomentumEnergySTD(groups.view(), d, box);
    sph::computeXMass(groups.view(), d, box);
    sph::computeVeDefGradh(groups.view(), d, box);
    sph::computeEOS(0, 10, d);
    sph::computeIadDivvCurlv(groups.view(), d, box);
    sph::computeAVswitches(groups.view(), d, box);
    sph::computeMomentumEnergy<false>(groups.view(), nullptr, d, box);
    (void)sph::rhoTimestep(0, 10, d);
    sph::computeTimestep(0, 10, d);
    sph::computePositions(groups.view(), d, box, 1e-3f, util::array<float, 4>{});
    (void)sph::updateSmoothingLength(groups.view(), d);
}
"""

# every entry point of sphx_dropin.hpp, by the C ABI function it must reach
DROPIN_CALLS = ("sphx_compute_sfc_keys", "sphx_sort_by_key", "sphx_gather_f64", "sphx_compute_octree",
                "sphx_build_octree", "sphx_node_centers", "sphx_find_neighbors", "sphx_compute_groups",
                "sphx_find_neighbors_sph", "sphx_compute_density", "sphx_compute_eos", "sphx_compute_iad",
                "sphx_compute_momentum_energy", "sphx_compute_xmass", "sphx_compute_ve_def_gradh",
                "sphx_compute_eos_ve", "sphx_compute_iad_divv_curlv", "sphx_compute_av_switches",
                "sphx_compute_momentum_energy_ve", "sphx_rho_timestep", "sphx_compute_timestep",
                "sphx_compute_positions", "sphx_update_smoothing_length")


def test_dropin_header_overrides_reference_entry_points():
    """include/sphx_dropin.hpp included after the reference's own headers: every call with the
    reference's own types and argument lists (cstone::..., sph::..., ParticlesData<GpuTag>,
    DeviceVector, OctreeNsView, GroupData) resolves, unambiguously, to the drop-in overload: the
    object file calls the library's C ABI for each, and instantiates none of the reference's
    implementations (no GPU kernels' host stubs, no CPU search, no MPI reduction).  The
    reference's sph/ts_global.hpp needs mpi.h, which this image lacks, so computeTimestep and
    rhoTimestep meet only the drop-in's declarations here."""
    ref = "/root/reference"
    if not os.path.isdir(os.path.join(ref, "domain", "include")):
        pytest.skip("reference sources not present")
    with tempfile.TemporaryDirectory() as d:
        f, obj = os.path.join(d, "t.cpp"), os.path.join(d, "t.o")
        open(f, "w").write(DROPIN_WITH_REFERENCE)
        cmd = ["g++", "-std=c++20", "-c", "-Wall", "-Wno-unused-variable", "-Wno-unknown-pragmas", "-D__HIP_PLATFORM_AMD__",
               "-I/opt/rocm/include", "-I", os.path.join(ROOT, "include"), "-I", f"{ref}/domain/include", "-I",
               f"{ref}/sph/include", "-I", f"{ref}/main/src", f, "-o", obj]
        out = subprocess.run(cmd, capture_output=True, text=True)
        assert out.returncode == 0, out.stderr[-3000:]
        undef = subprocess.run(["nm", "-u", obj], check=True, capture_output=True, text=True).stdout.split()
        defined = subprocess.run(["nm", "-C", "--defined-only", obj], check=True, capture_output=True,
                                 text=True).stdout
    missing = [c for c in DROPIN_CALLS if c not in undef]
    assert not missing, f"calls that did not reach the library: {missing}"
    assert not [s for s in undef if "Gpu" in s or s.startswith("MPI_")], undef
    for impl in ("findNeighborsSph", "computeDensityImpl", "computeIADImpl", "computeMomentumEnergyStdImpl",
                 "updateOctree(", "buildOctreeCpu", "computeXMassImpl"):
        assert impl not in defined, impl


def _syntax_check(src, extra_includes=()):
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "t.cpp")
        open(f, "w").write(src)
        cmd = ["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include")]
        for inc in extra_includes:
            cmd += ["-I", inc]
        out = subprocess.run(cmd + [f], capture_output=True, text=True)
        assert out.returncode == 0, out.stderr[-3000:]


def test_compat_header_compiles_with_mock_types():
    _syntax_check(MOCK)


def test_compat_header_compiles_with_reference_types():
    ref = "/root/reference/domain/include"
    if not os.path.isdir(ref):
        pytest.skip("reference sources not present")
    _syntax_check(REFERENCE_TYPES, [ref])


def test_target_groups_struct_points_into_bounds():
    import torch

    from sphexa_amd.hydro import TargetGroups

    b = torch.tensor([0, 64, 100], dtype=torch.int32)
    g = TargetGroups(0, 100, 2, b).struct()
    assert (g.firstBody, g.lastBody, g.numGroups) == (0, 100, 2)
    assert g.groupStart == b.data_ptr() and g.groupEnd == b.data_ptr() + 4


@pytest.mark.parametrize("struct,cname", [("Particles", "sphx_particles"), ("Timestep", "sphx_timestep"),
                                          ("VeFields", "sphx_ve_fields"),
                                          ("Box", "sphx_box"), ("Octree", "sphx_octree"), ("Groups", "sphx_groups")])
def test_ctypes_struct_layout_matches_header(struct, cname):
    """Every field of the ctypes mirror sits at the offset the C compiler gives it."""
    cls = getattr(_lib, struct)
    fields = [f[0] for f in cls._fields_]
    body = "".join(f'printf("%zu\\n", offsetof({cname}, {f}));' for f in fields)
    src = f'#include <stddef.h>\n#include <stdio.h>\n#include "sphx.h"\nint main(void){{{body}printf("%zu\\n", sizeof({cname}));return 0;}}\n'
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, "layout.c")
        exe = os.path.join(tmp, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", exe, c], check=True)
        out = [int(v) for v in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    want = [getattr(cls, f).offset for f in fields] + [C.sizeof(cls)]
    assert out == want
