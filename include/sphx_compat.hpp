/*
 * sphx_compat.hpp - C++ drop-in shims with the reference's entry-point names and
 * argument lists, forwarding to the C ABI in sphx.h.
 *
 * The templates are duck-typed on the caller's types, so they accept the reference's
 * own cstone::Box<double> (xmin()..zmax(), boundaryX()..), cstone::OctreeNsView /
 * OctreeView (public fields), cstone::GroupView and a sphexa::ParticlesData-like
 * dataset (fields with .data(), attributes K, ng0, ngmax, Kcour, gamma) without
 * including any reference header.  Field vectors must hold device memory.
 *
 * Failures throw sphx_compat::Error (the reference throws std::runtime_error or
 * exits; an exception keeps the caller in control).  One context per thread is
 * created on first use on the current device and the null stream; call
 * sphx_compat::setStream() to use another stream.
 */
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "sphx.h"

namespace sphx_compat
{

struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string& m)
        : std::runtime_error(m)
        , code(c)
    {
    }
};

inline sphx_ctx*& threadContext()
{
    thread_local sphx_ctx* ctx = nullptr;
    return ctx;
}

inline sphx_ctx* context()
{
    sphx_ctx*& c = threadContext();
    if (!c)
    {
        int st = sphx_ctx_create(0, nullptr, &c);
        if (st != SPHX_OK) throw Error(st, "sphx_ctx_create failed");
    }
    return c;
}

inline void check(int st)
{
    if (st != SPHX_OK) throw Error(st, sphx_last_error(threadContext()));
}

inline void setStream(void* stream) { check(sphx_ctx_set_stream(context(), stream)); }

template<class Box>
sphx_box toBox(const Box& b)
{
    auto bc = [](auto t) { return int32_t(static_cast<int>(t)); };
    return sphx_box{double(b.xmin()), double(b.xmax()),  double(b.ymin()),    double(b.ymax()), double(b.zmin()),
                    double(b.zmax()), bc(b.boundaryX()), bc(b.boundaryY()), bc(b.boundaryZ()), 0};
}

template<class V>
auto rawPtr(V& v) -> decltype(v.data())
{
    return v.data();
}

// ---- cstone::makeGlobalBox (sfc/box_mpi.hpp:68-107) on one rank: a Box of the caller's
// type with open dimensions fitted to the coordinates
template<class Box>
Box makeGlobalBox(const double* x, const double* y, const double* z, size_t n, const Box& previous)
{
    sphx_box p = toBox(previous), r{};
    check(sphx_make_global_box(context(), x, y, z, n, &p, &r));
    return Box(r.xmin, r.xmax, r.ymin, r.ymax, r.zmin, r.zmax, previous.boundaryX(), previous.boundaryY(),
               previous.boundaryZ());
}

// ---- cstone::computeSfcKeys<true> (domain/include/cstone/sfc/sfc_gpu.h:26-31)
template<class Box>
void computeSfcKeys(const double* x, const double* y, const double* z, uint64_t* keys, size_t n, const Box& box)
{
    sphx_box b = toBox(box);
    check(sphx_compute_sfc_keys(context(), x, y, z, keys, n, &b));
}

// ---- cstone::sortByKey<true> (primitives/primitives_acc.hpp:115-121); stable, in place
inline void sortByKey(uint64_t* keys, uint32_t* values, size_t n)
{
    check(sphx_sort_by_key(context(), keys, values, n));
}

// ---- cstone::gather (primitives/gather.hpp:90-98)
inline void gather(const uint32_t* order, const double* in, double* out, size_t n)
{
    check(sphx_gather_f64(context(), order, in, out, n));
}
inline void gather(const uint32_t* order, const float* in, float* out, size_t n)
{
    check(sphx_gather_f32(context(), order, in, out, n));
}

// ---- cstone::computeOctree (tree/csarray.hpp:430-440): device pointers owned by the
// context, valid until the next call on this thread
struct CsTree
{
    int32_t   numLeaves;
    uint64_t* leaves;
    uint32_t* counts;
    uint32_t* layout;
};
inline CsTree computeOctree(const uint64_t* keys, size_t n, unsigned bucketSize, unsigned maxCount = 0xFFFFFFFFu)
{
    CsTree t{};
    check(sphx_compute_octree(context(), keys, n, bucketSize, maxCount, &t.numLeaves, &t.leaves, &t.counts,
                              &t.layout));
    return t;
}

// ---- cstone::buildOctreeGpu(leaves, OctreeView) (tree/octree_gpu.h:34)
template<class OctreeView>
void buildOctreeGpu(const uint64_t* leaves, OctreeView o)
{
    sphx_octree t{};
    t.numLeafNodes   = o.numLeafNodes;
    t.prefixes       = reinterpret_cast<uint64_t*>(o.prefixes);
    t.childOffsets   = o.childOffsets;
    t.parents        = o.parents;
    t.levelRange     = o.d_levelRange;
    t.internalToLeaf = o.internalToLeaf;
    t.leafToInternal = o.leafToInternal;
    check(sphx_build_octree(context(), leaves, &t));
}

// ---- cstone::nodeFpCenters (focus/source_center.hpp:146-158), centers/sizes as Vec3<double>
template<class Vec3, class Box>
void nodeFpCenters(const uint64_t* prefixes, int32_t numNodes, Vec3* centers, Vec3* sizes, const Box& box)
{
    static_assert(sizeof(Vec3) == 3 * sizeof(double), "Vec3<double> expected");
    sphx_box b = toBox(box);
    check(sphx_node_centers(context(), prefixes, numNodes, &b, reinterpret_cast<double*>(centers),
                            reinterpret_cast<double*>(sizes)));
}

template<class NsView>
sphx_octree toOctree(const NsView& v, int32_t numNodes)
{
    sphx_octree t{};
    t.numLeafNodes   = v.numLeafNodes;
    t.numNodes       = numNodes;
    t.prefixes       = const_cast<uint64_t*>(reinterpret_cast<const uint64_t*>(v.prefixes));
    t.childOffsets   = const_cast<int32_t*>(v.childOffsets);
    t.parents        = const_cast<int32_t*>(v.parents);
    t.internalToLeaf = const_cast<int32_t*>(v.internalToLeaf);
    t.layout         = const_cast<uint32_t*>(v.layout);
    t.centers        = const_cast<double*>(reinterpret_cast<const double*>(v.centers));
    t.sizes          = const_cast<double*>(reinterpret_cast<const double*>(v.sizes));
    return t;
}

// ---- cstone::findNeighbors (findneighbors.hpp:149-170); OctreeNsView carries no node
// count, so the caller passes it
template<class Box, class NsView>
void findNeighbors(const double* x, const double* y, const double* z, const double* h, uint32_t firstId,
                   uint32_t lastId, const Box& box, const NsView& treeView, int32_t numNodes, unsigned ngmax,
                   uint32_t* neighbors, unsigned* neighborsCount)
{
    sphx_box    b = toBox(box);
    sphx_octree t = toOctree(treeView, numNodes);
    check(sphx_find_neighbors(context(), x, y, z, h, firstId, lastId, &b, &t, ngmax, neighbors, neighborsCount));
}

// ---- the std-SPH dataset (sphexa::ParticlesData field contract)
template<class Dataset>
sphx_particles toParticles(Dataset& d)
{
    using T = std::remove_cv_t<std::remove_pointer_t<decltype(rawPtr(d.h))>>;
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>, "hydro fields must be double or float");
    sphx_particles p{};
    p.n         = d.x.size();
    p.precision = std::is_same_v<T, double> ? SPHX_FP64 : SPHX_MIXED;
    p.ng0 = d.ng0, p.ngmax = d.ngmax, p.Kcour = float(d.Kcour), p.K = double(d.K), p.gamma = double(d.gamma);
    p.x = rawPtr(d.x), p.y = rawPtr(d.y), p.z = rawPtr(d.z);
    p.h = rawPtr(d.h), p.m = rawPtr(d.m), p.vx = rawPtr(d.vx), p.vy = rawPtr(d.vy), p.vz = rawPtr(d.vz);
    p.u = rawPtr(d.u), p.rho = rawPtr(d.rho), p.p = rawPtr(d.p), p.c = rawPtr(d.c);
    p.c11 = rawPtr(d.c11), p.c12 = rawPtr(d.c12), p.c13 = rawPtr(d.c13);
    p.c22 = rawPtr(d.c22), p.c23 = rawPtr(d.c23), p.c33 = rawPtr(d.c33);
    p.ax = rawPtr(d.ax), p.ay = rawPtr(d.ay), p.az = rawPtr(d.az), p.du = rawPtr(d.du);
    p.nc = rawPtr(d.nc), p.wh = rawPtr(d.wh);
    return p;
}

template<class GroupView>
sphx_groups toGroups(const GroupView& g)
{
    return sphx_groups{uint32_t(g.firstBody), uint32_t(g.lastBody), uint32_t(g.numGroups),
                       reinterpret_cast<const uint32_t*>(g.groupStart), reinterpret_cast<const uint32_t*>(g.groupEnd)};
}

// ---- cstone::computeGroupSplits with groupSize 64 (traversal/groups_gpu.cu:108-137), as
// sph::computeSpatialGroups calls it (sph/groups.hpp:42-55, tolFactor 2).  The returned
// view's arrays are owned by the context and valid until the next call on this thread.
struct GroupSplits
{
    uint32_t        firstBody, lastBody, numGroups;
    const uint32_t* groupStart;
    const uint32_t* groupEnd;
};
template<class Th, class Box>
GroupSplits computeGroupSplits(uint32_t first, uint32_t last, const double* x, const double* y, const double* z,
                               const Th* h, const uint64_t* leaves, int32_t numLeaves, const uint32_t* layout,
                               const Box& box, float tolFactor = 2.0f)
{
    static_assert(std::is_same_v<Th, double> || std::is_same_v<Th, float>, "h must be double or float");
    sphx_box    b = toBox(box);
    sphx_octree t{};
    t.numLeafNodes = numLeaves;
    t.leaves       = const_cast<uint64_t*>(leaves);
    t.layout       = const_cast<uint32_t*>(layout);
    sphx_groups g{};
    check(sphx_compute_groups(context(), first, last, x, y, z, h, std::is_same_v<Th, double> ? SPHX_FP64 : SPHX_MIXED,
                              &t, &b, tolFactor, &g));
    return GroupSplits{g.firstBody, g.lastBody, g.numGroups, g.groupStart, g.groupEnd};
}

// ---- sph::findNeighborsSfc (sph/include/sph/find_neighbors.hpp:44-53) + the GPU
// groups of computeGroups; the lists stay in the context for the loops below
template<class GroupView, class Dataset, class Box, class NsView>
void findNeighborsSfc(const GroupView& groups, Dataset& d, const Box& box, const NsView& treeView, int32_t numNodes)
{
    sphx_particles p = toParticles(d);
    sphx_groups    g = toGroups(groups);
    sphx_box       b = toBox(box);
    sphx_octree    t = toOctree(treeView, numNodes);
    check(sphx_find_neighbors_sph(context(), &g, &p, &b, &t));
}

// ---- sph::computeDensity (hydro_std/density.hpp:55-60)
template<class GroupView, class Dataset, class Box>
void computeDensity(const GroupView& groups, Dataset& d, const Box& box)
{
    sphx_particles p = toParticles(d);
    sphx_groups    g = toGroups(groups);
    sphx_box       b = toBox(box);
    check(sphx_compute_density(context(), &g, &p, &b));
}

// ---- sph::computeEOS_HydroStd, ideal gas (hydro_std/eos.hpp:144-150)
template<class Dataset>
void computeEOS_HydroStd(size_t startIndex, size_t endIndex, Dataset& d)
{
    sphx_particles p = toParticles(d);
    check(sphx_compute_eos(context(), uint32_t(startIndex), uint32_t(endIndex), &p));
}

// ---- sph::computeIAD (hydro_std/iad.hpp:74-79)
template<class GroupView, class Dataset, class Box>
void computeIAD(const GroupView& groups, Dataset& d, const Box& box)
{
    sphx_particles p = toParticles(d);
    sphx_groups    g = toGroups(groups);
    sphx_box       b = toBox(box);
    check(sphx_compute_iad(context(), &g, &p, &b));
}

// ---- sph::computeMomentumEnergySTD (hydro_std/momentum_energy.hpp:96-101); sets d.minDtCourant
template<class GroupView, class Dataset, class Box>
void computeMomentumEnergySTD(const GroupView& groups, Dataset& d, const Box& box)
{
    sphx_particles p = toParticles(d);
    sphx_groups    g = toGroups(groups);
    sphx_box       b = toBox(box);
    check(sphx_compute_momentum_energy(context(), &g, &p, &b));
    d.minDtCourant = p.minDtCourant;
}

// ---- sph::computeTimestep (sph/include/sph/ts_global.hpp:100-130) on one rank: updates
// d.minDt, d.minDt_m1 and d.ttot from d.minDtCourant, d.minDtRho, 1.1 d.minDt and, when
// d.g != 0, the acceleration limit (extra limits: fold them into d.minDtRho first)
template<class Dataset>
void computeTimestep(size_t first, size_t last, Dataset& d)
{
    sphx_particles p = toParticles(d);
    sphx_timestep  ts{};
    ts.minDt = double(d.minDt), ts.minDt_m1 = double(d.minDt_m1), ts.ttot = double(d.ttot);
    ts.minDtCourant = double(d.minDtCourant), ts.minDtRho = double(d.minDtRho);
    ts.g = double(d.g), ts.etaAcc = double(d.etaAcc);
    check(sphx_compute_timestep(context(), uint32_t(first), uint32_t(last), &p, &ts));
    d.minDt = ts.minDt, d.minDt_m1 = ts.minDt_m1, d.ttot = ts.ttot;
}

// ---- sph::computePositions (sph/include/sph/positions.hpp:207-230) with the CPU branch's
// semantics: the step is d.minDt, d.minDt_m1 (dt_forward and dt_m1 are the GPU branch's
// float copies of them); block time steps (rung != nullptr) are not supported
template<class GroupView, class Dataset, class Box, class DtM1>
void computePositions(const GroupView& groups, Dataset& d, const Box& box, float /*dt_forward*/,
                      const DtM1& /*dt_m1*/, const uint8_t* rung = nullptr)
{
    if (rung) throw std::runtime_error("sphx_compat::computePositions: rungs are not supported");
    sphx_particles p = toParticles(d);
    p.x_m1 = rawPtr(d.x_m1), p.y_m1 = rawPtr(d.y_m1), p.z_m1 = rawPtr(d.z_m1), p.du_m1 = rawPtr(d.du_m1);
    sphx_groups   g = toGroups(groups);
    sphx_box      b = toBox(box);
    sphx_timestep ts{};
    ts.minDt = double(d.minDt), ts.minDt_m1 = double(d.minDt_m1);
    check(sphx_compute_positions(context(), &g, &p, &b, &ts));
}

// ---- sph::updateSmoothingLength (sph/include/sph/update_h.hpp:35-49): returns whether a
// particle was left with nc <= 1 (its key becomes removeKey)
template<class GroupView, class Dataset>
bool updateSmoothingLength(const GroupView& groups, Dataset& d)
{
    sphx_particles p = toParticles(d);
    sphx_groups    g = toGroups(groups);
    int            removed = 0;
    check(sphx_update_smoothing_length(context(), &g, &p, reinterpret_cast<uint64_t*>(rawPtr(d.keys)), &removed));
    return removed != 0;
}

// ---- VE hydro (sph/include/sph/hydro_ve): the dataset also carries xm, kx, gradh, prho,
// divv, curlv, alpha, whd and the VE constants; dV11..dV33 are used when they hold n values
template<class Dataset>
sphx_ve_fields toVeFields(Dataset& d)
{
    sphx_ve_fields v{};
    v.xm = rawPtr(d.xm), v.kx = rawPtr(d.kx), v.gradh = rawPtr(d.gradh), v.prho = rawPtr(d.prho);
    v.divv = rawPtr(d.divv), v.alpha = rawPtr(d.alpha), v.whd = rawPtr(d.whd);
    v.curlv = d.curlv.size() == d.x.size() ? rawPtr(d.curlv) : nullptr;
    if (d.dV11.size() == d.x.size())
        v.dV11 = rawPtr(d.dV11), v.dV12 = rawPtr(d.dV12), v.dV13 = rawPtr(d.dV13), v.dV22 = rawPtr(d.dV22),
        v.dV23 = rawPtr(d.dV23), v.dV33 = rawPtr(d.dV33);
    v.alphamin = double(d.alphamin), v.alphamax = double(d.alphamax), v.decay = double(d.decay_constant);
    v.Atmin = double(d.Atmin), v.Atmax = double(d.Atmax), v.ramp = double(d.ramp);
    return v;
}

// sph::computeXMass (hydro_ve/xmass.hpp:69-75)
template<class GroupView, class Dataset, class Box>
void computeXMass(const GroupView& groups, Dataset& d, const Box& box)
{
    sphx_particles p = toParticles(d);
    sphx_ve_fields v = toVeFields(d);
    sphx_groups    g = toGroups(groups);
    sphx_box       b = toBox(box);
    check(sphx_compute_xmass(context(), &g, &p, &v, &b));
}

// sph::computeVeDefGradh (hydro_ve/ve_def_gradh.hpp:80-86)
template<class GroupView, class Dataset, class Box>
void computeVeDefGradh(const GroupView& groups, Dataset& d, const Box& box)
{
    sphx_particles p = toParticles(d);
    sphx_ve_fields v = toVeFields(d);
    sphx_groups    g = toGroups(groups);
    sphx_box       b = toBox(box);
    check(sphx_compute_ve_def_gradh(context(), &g, &p, &v, &b));
}

// sph::computeEOS, ideal gas with u (hydro_ve/eos.hpp:192-197)
template<class Dataset>
void computeEOS(size_t startIndex, size_t endIndex, Dataset& d)
{
    sphx_particles p = toParticles(d);
    sphx_ve_fields v = toVeFields(d);
    check(sphx_compute_eos_ve(context(), uint32_t(startIndex), uint32_t(endIndex), &p, &v));
}

// sph::computeIadDivvCurlv (hydro_ve/iad_divv_curlv.hpp:92-98)
template<class GroupView, class Dataset, class Box>
void computeIadDivvCurlv(const GroupView& groups, Dataset& d, const Box& box)
{
    sphx_particles p = toParticles(d);
    sphx_ve_fields v = toVeFields(d);
    sphx_groups    g = toGroups(groups);
    sphx_box       b = toBox(box);
    check(sphx_compute_iad_divv_curlv(context(), &g, &p, &v, &b));
}

// sph::computeAVswitches (hydro_ve/av_switches.hpp:81-87), over the step d.minDt
template<class GroupView, class Dataset, class Box>
void computeAVswitches(const GroupView& groups, Dataset& d, const Box& box)
{
    sphx_particles p = toParticles(d);
    sphx_ve_fields v = toVeFields(d);
    sphx_groups    g = toGroups(groups);
    sphx_box       b = toBox(box);
    check(sphx_compute_av_switches(context(), &g, &p, &v, &b, double(d.minDt)));
}

// sph::rhoTimestep (sph/include/sph/ts_global.hpp:73-98): d.Krho / |max divv|
template<class Dataset>
auto rhoTimestep(size_t first, size_t last, const Dataset& d)
{
    sphx_particles p = toParticles(const_cast<Dataset&>(d));
    sphx_ve_fields v = toVeFields(const_cast<Dataset&>(d));
    double         r = 0;
    check(sphx_rho_timestep(context(), uint32_t(first), uint32_t(last), &p, &v, double(d.Krho), &r));
    return r;
}

// sph::computeMomentumEnergy<avClean> (hydro_ve/momentum_energy.hpp:103-114); sets
// d.minDtCourant.  The per-group time steps (groupDt) are not supported.
template<bool avClean, class GroupView, class Dataset, class Box>
void computeMomentumEnergy(const GroupView& groups, float* groupDt, Dataset& d, const Box& box)
{
    if (groupDt) throw std::runtime_error("sphx_compat::computeMomentumEnergy: groupDt is not supported");
    sphx_particles p = toParticles(d);
    sphx_ve_fields v = toVeFields(d);
    v.avClean        = avClean ? 1 : 0;
    sphx_groups    g = toGroups(groups);
    sphx_box       b = toBox(box);
    check(sphx_compute_momentum_energy_ve(context(), &g, &p, &v, &b));
    d.minDtCourant = p.minDtCourant;
}

} // namespace sphx_compat
