/*
 * sphx_dropin.hpp - the reference's entry points of the single-rank hydro path under
 * their own names, namespaces and argument lists (cstone::..., sph::...), forwarding to
 * libsphx.  Opt-in and C++20.
 *
 * Each function here is an overload of the reference's function template with the same
 * template parameters and the same function parameters plus a requires-clause.  When both
 * are visible, overload resolution prefers the more constrained one (this one), so a
 * translation unit that includes this header after the reference's headers routes its
 * existing calls to the GPU library without editing a call site; without the reference's
 * headers the reference types are the caller's own (declared below, defined by the
 * caller).  The constraints select: 64-bit keys (uint64_t), FP64 coordinates, 32-bit
 * values and indices, and datasets whose AcceleratorType is cstone::GpuTag (fields in
 * device memory).  Other instantiations keep resolving to the reference's templates.
 *
 * The functions the reference runs on the CPU (computeOctree, updateInternalTree,
 * nodeFpCenters, findNeighbors) accept host or device memory: host arrays are copied to
 * the device and the results back.  Everything else expects device memory.
 *
 * Semantics that differ from the reference (all single-rank):
 * - cstone::updateOctreeGpu returns the converged cornerstone tree at once (and true):
 *   the rebalance fixed point is unique (a node is internal iff it holds more than
 *   bucketSize particles), so iterating from the caller's tree reaches the same one.
 * - sph::findNeighborsSfc runs the neighbor search with the smoothing-length iteration
 *   (the reference's CPU semantics; its GPU build returns at once and searches inside
 *   every loop) over the groups of the last sph::computeGroups on this thread when they
 *   cover [startIndex, endIndex), else over fixed groups of 64; the lists stay in the
 *   library for the loops, which must receive those same groups.
 * - sph::computeTimestep does not reduce across ranks; extra time steps are folded in.
 * - sortByKey<true> uses the library's scratch instead of keyBuf / valueBuf.
 * Failures throw sphx_compat::Error.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <span>
#include <tuple>
#include <type_traits>
#include <vector>

#include "sphx_compat.hpp"

// the reference's types (domain/include/cstone/...), declared only: their definitions come
// from the reference's headers, or from the caller
namespace util
{
template<class T, std::size_t N>
struct array; // cstone/util/array.hpp (cstone::Vec3<T> = util::array<T, 3>)
} // namespace util

namespace cstone
{
template<class T>
class Box; // sfc/box.hpp
template<class KeyType>
struct OctreeView; // tree/octree.hpp
template<class T, class KeyType>
struct OctreeNsView; // tree/octree.hpp
struct GroupView;    // traversal/groups.hpp
template<class Accelerator>
class GroupData; // traversal/groups.hpp
struct CpuTag;   // primitives/primitives_acc.hpp
struct GpuTag;
} // namespace cstone

namespace sphx_dropin
{

template<class K>
concept Key64 = std::is_same_v<std::remove_cv_t<K>, uint64_t>;

template<class D>
concept DeviceDataset = std::is_same_v<typename D::AcceleratorType, cstone::GpuTag>;

inline void hipCheck(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw sphx_compat::Error(SPHX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// whether p is device (or managed) memory; host memory, pinned or not, is staged
inline bool onDevice(const void* p)
{
    if (!p) return true;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess)
    {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// n elements at p as device memory: p itself, or a device copy of host memory (copied in
// when `in`, written back by copyOut())
template<class T>
class Staged
{
public:
    Staged(T* p, size_t n, bool in)
        : host_(p)
        , n_(n)
    {
        if (onDevice(p) || n == 0)
        {
            dev_ = p;
            return;
        }
        using U = std::remove_const_t<T>;
        U* d    = nullptr;
        hipCheck(hipMalloc(&d, n * sizeof(U)), "hipMalloc");
        dev_ = d;
        own_ = true;
        if (in) hipCheck(hipMemcpy(d, p, n * sizeof(U), hipMemcpyHostToDevice), "hipMemcpy to device");
    }
    Staged(const Staged&)            = delete;
    Staged& operator=(const Staged&) = delete;
    ~Staged()
    {
        if (own_) (void)hipFree(const_cast<std::remove_const_t<T>*>(dev_));
    }
    T* get() const { return dev_; }
    void copyOut()
    {
        if (own_)
            hipCheck(hipMemcpy(const_cast<std::remove_const_t<T>*>(host_), dev_, n_ * sizeof(T), hipMemcpyDeviceToHost),
                     "hipMemcpy to host");
    }

private:
    T*     host_;
    T*     dev_ = nullptr;
    size_t n_;
    bool   own_ = false;
};

// every queued library launch has finished (before plain hipMemcpy of its results)
inline void sync() { sphx_compat::check(sphx_ctx_synchronize(sphx_compat::context())); }

// the octree node count of a cornerstone tree with numLeaves leaves
inline int32_t numNodesOf(int32_t numLeaves) { return numLeaves + (numLeaves - 1) / 7; }

// groups of the last sph::computeGroups on this thread (pointers into the caller's GroupData)
inline sphx_groups& lastGroups()
{
    thread_local sphx_groups g{0, 0, 0, nullptr, nullptr};
    return g;
}

} // namespace sphx_dropin

namespace cstone
{

// cstone::computeSfcKeys<true> (sfc/sfc_gpu.h:26-31): Hilbert keys of device coordinates
template<bool useGpu, class KeyType, class T>
    requires(useGpu && sphx_dropin::Key64<KeyType> && std::is_same_v<T, double>)
void computeSfcKeys(const T* x, const T* y, const T* z, KeyType* keys, size_t numKeys, const Box<T>& box)
{
    sphx_compat::computeSfcKeys(x, y, z, keys, numKeys, box);
}

// cstone::sortByKey<true> (primitives/primitives_acc.hpp:115-121): stable, in place, device
// spans; the library's scratch replaces keyBuf / valueBuf
template<bool useGpu, class KeyType, class ValueType, class KeyBuf, class ValueBuf>
    requires(useGpu && sphx_dropin::Key64<KeyType> && std::is_same_v<ValueType, unsigned>)
void sortByKey(std::span<KeyType> keys, std::span<ValueType> values, KeyBuf& /*keyBuf*/, ValueBuf& /*valueBuf*/,
               double /*growth*/)
{
    if (keys.size() != values.size()) throw sphx_compat::Error(SPHX_ERR_ARG, "sortByKey: keys and values differ in size");
    sphx_compat::check(sphx_sort_by_key(sphx_compat::context(), keys.data(), values.data(), keys.size()));
}

// cstone::gatherAcc<true> (primitives/primitives_acc.hpp:66-71): destination[i] =
// source[ordering[i]] on the device, FP64, FP32 or 32-bit fields
template<bool useGpu, class IndexType, class ValueType>
    requires(useGpu && std::is_same_v<std::remove_cv_t<IndexType>, unsigned> &&
             (std::is_same_v<ValueType, double> || std::is_same_v<ValueType, float> ||
              std::is_same_v<ValueType, unsigned>))
void gatherAcc(std::span<const IndexType> ordering, const ValueType* source, ValueType* destination)
{
    sphx_ctx* c = sphx_compat::context();
    if constexpr (std::is_same_v<ValueType, double>)
        sphx_compat::check(sphx_gather_f64(c, ordering.data(), source, destination, ordering.size()));
    else if constexpr (std::is_same_v<ValueType, float>)
        sphx_compat::check(sphx_gather_f32(c, ordering.data(), source, destination, ordering.size()));
    else
        sphx_compat::check(sphx_gather_u32(c, ordering.data(), source, destination, ordering.size()));
}

// cstone::computeOctree (tree/csarray.hpp:430-440): the converged cornerstone leaves and
// their counts as host vectors, from keys in host or device memory
template<class KeyType>
    requires sphx_dropin::Key64<KeyType>
std::tuple<std::vector<KeyType>, std::vector<unsigned>>
computeOctree(std::span<const KeyType> keys, unsigned bucketSize, unsigned maxCount = std::numeric_limits<unsigned>::max())
{
    if (keys.empty()) return {std::vector<KeyType>{0, KeyType(1) << 63}, std::vector<unsigned>{0}};
    sphx_dropin::Staged<const KeyType> k(keys.data(), keys.size(), true);
    auto t = sphx_compat::computeOctree(reinterpret_cast<const uint64_t*>(k.get()), keys.size(), bucketSize, maxCount);
    std::vector<KeyType>  leaves(size_t(t.numLeaves) + 1);
    std::vector<unsigned> counts(t.numLeaves);
    sphx_dropin::sync();
    sphx_dropin::hipCheck(hipMemcpy(leaves.data(), t.leaves, leaves.size() * sizeof(KeyType), hipMemcpyDeviceToHost),
                          "computeOctree: leaves");
    sphx_dropin::hipCheck(hipMemcpy(counts.data(), t.counts, counts.size() * sizeof(unsigned), hipMemcpyDeviceToHost),
                          "computeOctree: counts");
    return {std::move(leaves), std::move(counts)};
}

// cstone::updateOctreeGpu (tree/update_gpu.cuh:41-64): tree and counts (device vectors with
// resize() and data()) become the converged cornerstone tree of the device keys; returns true
template<class KeyType, class DevKeyVec, class DevCountVec, class DevIdxVec>
    requires sphx_dropin::Key64<KeyType>
bool updateOctreeGpu(std::span<const KeyType> keys, unsigned bucketSize, DevKeyVec& tree, DevCountVec& counts,
                     DevKeyVec& /*tmpTree*/, DevIdxVec& /*workArray*/,
                     unsigned maxCount = std::numeric_limits<unsigned>::max())
{
    if (keys.empty())
    {
        const KeyType root[2] = {0, KeyType(1) << 63};
        const unsigned zero    = 0;
        tree.resize(2);
        counts.resize(1);
        sphx_dropin::hipCheck(hipMemcpy(tree.data(), root, sizeof(root), hipMemcpyHostToDevice), "updateOctreeGpu");
        sphx_dropin::hipCheck(hipMemcpy(counts.data(), &zero, sizeof(zero), hipMemcpyHostToDevice), "updateOctreeGpu");
        return true;
    }
    auto t = sphx_compat::computeOctree(reinterpret_cast<const uint64_t*>(keys.data()), keys.size(), bucketSize,
                                        maxCount);
    tree.resize(size_t(t.numLeaves) + 1);
    counts.resize(size_t(t.numLeaves));
    sphx_dropin::sync();
    sphx_dropin::hipCheck(hipMemcpy(tree.data(), t.leaves, (size_t(t.numLeaves) + 1) * sizeof(KeyType),
                                    hipMemcpyDeviceToDevice), "updateOctreeGpu: leaves");
    sphx_dropin::hipCheck(hipMemcpy(counts.data(), t.counts, size_t(t.numLeaves) * sizeof(unsigned),
                                    hipMemcpyDeviceToDevice), "updateOctreeGpu: counts");
    return true;
}

// cstone::buildOctreeGpu (tree/octree_gpu.h:34): the linked internal octree of device leaves
// into the device arrays of the view
template<class KeyType>
    requires sphx_dropin::Key64<KeyType>
void buildOctreeGpu(const KeyType* cstoneTree, OctreeView<KeyType> d)
{
    sphx_compat::buildOctreeGpu(reinterpret_cast<const uint64_t*>(cstoneTree), d);
}

// cstone::updateInternalTree (tree/octree.hpp:361-367): buildOctreeCpu into the view's arrays
// (host or device), levelRange copied to d_levelRange
template<class KeyType>
    requires sphx_dropin::Key64<KeyType>
void updateInternalTree(std::span<const KeyType> leaves, OctreeView<KeyType> o)
{
    using sphx_dropin::Staged, sphx_dropin::hipCheck, sphx_dropin::sync, sphx_dropin::numNodesOf;
    const int32_t L = o.numLeafNodes;
    if (size_t(L) + 1 != leaves.size())
        throw sphx_compat::Error(SPHX_ERR_ARG, "updateInternalTree: numLeafNodes differs from the leaves");
    int32_t ni = 0, nn = 0, np = 0;
    sphx_compat::check(sphx_octree_sizes(L, &ni, &nn, &np));
    Staged<const KeyType> lv(leaves.data(), leaves.size(), true);
    Staged<KeyType>       prefixes(o.prefixes, nn, false);
    Staged<int32_t>       childOffsets(o.childOffsets, size_t(nn), false), parents(o.parents, size_t(np), false),
        levelRange(o.levelRange, 23, false), internalToLeaf(o.internalToLeaf, size_t(nn), false),
        leafToInternal(o.leafToInternal, size_t(nn), false);
    sphx_octree t{};
    t.numLeafNodes   = L;
    t.prefixes       = reinterpret_cast<uint64_t*>(prefixes.get());
    t.childOffsets   = childOffsets.get();
    t.parents        = parents.get();
    t.levelRange     = levelRange.get();
    t.internalToLeaf = internalToLeaf.get();
    t.leafToInternal = leafToInternal.get();
    sphx_compat::check(sphx_build_octree(sphx_compat::context(), reinterpret_cast<const uint64_t*>(lv.get()), &t));
    sync();
    prefixes.copyOut(), childOffsets.copyOut(), parents.copyOut(), levelRange.copyOut(), internalToLeaf.copyOut();
    leafToInternal.copyOut();
    if (o.d_levelRange && o.d_levelRange != o.levelRange)
        hipCheck(hipMemcpy(o.d_levelRange, levelRange.get(), 23 * sizeof(int32_t), hipMemcpyDefault),
                 "updateInternalTree: d_levelRange");
}

// cstone::nodeFpCenters (focus/source_center.hpp:146-158): node centres and half-sizes
// (Vec3<double> = util::array<double, 3>) in host or device memory
template<class KeyType, class T>
    requires(sphx_dropin::Key64<KeyType> && std::is_same_v<T, double>)
void nodeFpCenters(std::span<const KeyType> prefixes, util::array<T, 3>* centers, util::array<T, 3>* sizes,
                   const Box<T>& box)
{
    using sphx_dropin::Staged, sphx_dropin::hipCheck, sphx_dropin::sync, sphx_dropin::numNodesOf;
    const size_t n = prefixes.size();
    Staged<const KeyType> p(prefixes.data(), n, true);
    Staged<T>             c(reinterpret_cast<T*>(centers), 3 * n, false), s(reinterpret_cast<T*>(sizes), 3 * n, false);
    sphx_box              b = sphx_compat::toBox(box);
    sphx_compat::check(sphx_node_centers(sphx_compat::context(), reinterpret_cast<const uint64_t*>(p.get()),
                                         int32_t(n), &b, c.get(), s.get()));
    sync();
    c.copyOut(), s.copyOut();
}

// cstone::findNeighbors (findneighbors.hpp:149-170): counts (without self) and lists in the
// CPU layout neighbors[(i - firstId) * ngmax + k], arrays in host or device memory
template<class T, class KeyType>
    requires(std::is_same_v<T, double> && sphx_dropin::Key64<KeyType>)
void findNeighbors(const T* x, const T* y, const T* z, const T* h, unsigned firstId, unsigned lastId,
                   const Box<T>& box, const OctreeNsView<T, KeyType>& treeView, unsigned ngmax, unsigned* neighbors,
                   unsigned* neighborsCount)
{
    using sphx_dropin::Staged, sphx_dropin::hipCheck, sphx_dropin::sync, sphx_dropin::numNodesOf;
    if (lastId <= firstId) return;
    const int32_t L  = treeView.numLeafNodes;
    const int32_t nn = numNodesOf(L);
    const size_t  np = lastId; // particles the search may read: [0, lastId) and its tree's layout
    size_t        numParticles = np;
    {
        // the tree's particle count: layout[L] (host or device)
        unsigned last = 0;
        hipCheck(hipMemcpy(&last, treeView.layout + L, sizeof(unsigned), hipMemcpyDefault), "findNeighbors: layout");
        numParticles = std::max(np, size_t(last));
    }
    Staged<const T>        xs(x, numParticles, true), ys(y, numParticles, true), zs(z, numParticles, true),
        hs(h, numParticles, true);
    Staged<const int32_t>  co(treeView.childOffsets, size_t(nn), true), i2l(treeView.internalToLeaf, size_t(nn), true),
        par(treeView.parents, size_t(std::max(1, (nn - 1) / 8)), true);
    Staged<const uint32_t> lay(reinterpret_cast<const uint32_t*>(treeView.layout), size_t(L) + 1, true);
    Staged<const T>        ce(reinterpret_cast<const T*>(treeView.centers), 3 * size_t(nn), true),
        si(reinterpret_cast<const T*>(treeView.sizes), 3 * size_t(nn), true);
    const size_t      nw = lastId - firstId;
    Staged<uint32_t>  nb(neighbors, neighbors ? nw * ngmax : 0, false), nc(neighborsCount, nw, false);
    sphx_box          b = sphx_compat::toBox(box);
    sphx_octree       t{};
    t.numLeafNodes   = L;
    t.numNodes       = nn;
    t.childOffsets   = const_cast<int32_t*>(co.get());
    t.parents        = const_cast<int32_t*>(par.get());
    t.internalToLeaf = const_cast<int32_t*>(i2l.get());
    t.layout         = const_cast<uint32_t*>(lay.get());
    t.centers        = const_cast<double*>(ce.get());
    t.sizes          = const_cast<double*>(si.get());
    sphx_compat::check(sphx_find_neighbors(sphx_compat::context(), xs.get(), ys.get(), zs.get(), hs.get(), firstId,
                                           lastId, &b, &t, ngmax, nb.get(), nc.get()));
    sync();
    nb.copyOut(), nc.copyOut();
}

} // namespace cstone

namespace sph
{

// sph::computeGroups (sph/groups.hpp:58-74), GPU branch (computeSpatialGroups :41-55):
// target groups of up to 64 particles split at spatial gaps (tolFactor 2) into the caller's
// GroupData (data resized to numGroups + 1 boundaries); also remembered for findNeighborsSfc
template<typename Tc, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeGroups(size_t startIndex, size_t endIndex, Dataset& d, const cstone::Box<Tc>& box,
                   cstone::GroupData<typename Dataset::AcceleratorType>& groups)
{
    const auto& tv = d.treeView;
    using sphx_compat::rawPtr;
    auto g = sphx_compat::computeGroupSplits(uint32_t(startIndex), uint32_t(endIndex), rawPtr(d.x), rawPtr(d.y),
                                             rawPtr(d.z), rawPtr(d.h), reinterpret_cast<const uint64_t*>(tv.leaves),
                                             int32_t(tv.numLeafNodes), reinterpret_cast<const uint32_t*>(tv.layout), box);
    groups.data.resize(size_t(g.numGroups) + 1);
    sphx_dropin::sync();
    if (g.numGroups > 0)
        sphx_dropin::hipCheck(hipMemcpy(rawPtr(groups.data), g.groupStart, (size_t(g.numGroups) + 1) * sizeof(unsigned),
                                        hipMemcpyDeviceToDevice),
                              "computeGroups");
    else
    {
        const unsigned end = unsigned(endIndex); // no groups: one boundary at lastBody
        sphx_dropin::hipCheck(hipMemcpy(rawPtr(groups.data), &end, sizeof(end), hipMemcpyHostToDevice), "computeGroups");
    }
    groups.firstBody  = g.firstBody;
    groups.lastBody   = g.lastBody;
    groups.numGroups  = g.numGroups;
    groups.groupStart = rawPtr(groups.data);
    groups.groupEnd   = rawPtr(groups.data) + 1;
    sphx_dropin::lastGroups() = sphx_groups{g.firstBody, g.lastBody, g.numGroups, groups.groupStart, groups.groupEnd};
}

// sph::findNeighborsSfc (sph/find_neighbors.hpp:44-53): the neighbor search with the
// smoothing-length iteration over d.treeView; h and nc are updated in d and the lists kept
// in the library for the loops
template<class T, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void findNeighborsSfc(size_t startIndex, size_t endIndex, Dataset& d, const cstone::Box<T>& box)
{
    sphx_groups g = sphx_dropin::lastGroups();
    if (!(g.groupStart && g.firstBody == startIndex && g.lastBody == endIndex))
        g = sphx_groups{uint32_t(startIndex), uint32_t(endIndex), uint32_t((endIndex - startIndex + 63) / 64), nullptr,
                        nullptr};
    const auto& tv = d.treeView;
    sphx_octree t{};
    t.numLeafNodes   = tv.numLeafNodes;
    t.numNodes       = sphx_dropin::numNodesOf(tv.numLeafNodes);
    t.childOffsets   = const_cast<int32_t*>(tv.childOffsets);
    t.parents        = const_cast<int32_t*>(tv.parents);
    t.internalToLeaf = const_cast<int32_t*>(tv.internalToLeaf);
    t.layout         = const_cast<uint32_t*>(reinterpret_cast<const uint32_t*>(tv.layout));
    t.centers        = const_cast<double*>(reinterpret_cast<const double*>(tv.centers));
    t.sizes          = const_cast<double*>(reinterpret_cast<const double*>(tv.sizes));
    sphx_particles p = sphx_compat::toParticles(d);
    sphx_box       b = sphx_compat::toBox(box);
    sphx_compat::check(sphx_find_neighbors_sph(sphx_compat::context(), &g, &p, &b, &t));
}

// the std-hydro loops (hydro_std/density.hpp:55-60, eos.hpp:144-150, iad.hpp:74-79,
// momentum_energy.hpp:96-101)
template<class T, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeDensity(const cstone::GroupView& groups, Dataset& d, const cstone::Box<T>& box)
{
    sphx_compat::computeDensity(groups, d, box);
}

template<class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeEOS_HydroStd(size_t startIndex, size_t endIndex, Dataset& d)
{
    sphx_compat::computeEOS_HydroStd(startIndex, endIndex, d);
}

template<class T, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeIAD(const cstone::GroupView& groups, Dataset& d, const cstone::Box<T>& box)
{
    sphx_compat::computeIAD(groups, d, box);
}

template<class T, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeMomentumEnergySTD(const cstone::GroupView& groups, Dataset& d, const cstone::Box<T>& box)
{
    sphx_compat::computeMomentumEnergySTD(groups, d, box);
}

// sph::computeTimestep (ts_global.hpp:100-130) on one rank; extra limits join the minimum
template<class Dataset, class... Ts>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeTimestep(size_t first, size_t last, Dataset& d, Ts... extraTimesteps)
{
    auto saved = d.minDtRho;
    ((d.minDtRho = std::min<decltype(saved)>(d.minDtRho, decltype(saved)(extraTimesteps))), ...);
    sphx_compat::computeTimestep(first, last, d);
    d.minDtRho = saved;
}

// sph::computePositions (positions.hpp:207-230), the CPU branch's semantics: the step is
// d.minDt, d.minDt_m1 (dt_forward and dt_m1 are the GPU branch's float copies of them;
// util::array<float, 4> is util::array<float, Timestep::maxNumRungs>, sph/timestep.h:19);
// block time steps (rung != nullptr) are not supported
template<class T, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computePositions(const cstone::GroupView& grp, Dataset& d, const cstone::Box<T>& box, float dt_forward,
                      util::array<float, 4> dt_m1, const uint8_t* rung = nullptr)
{
    sphx_compat::computePositions(grp, d, box, dt_forward, dt_m1, rung);
}

// sph::updateSmoothingLength (update_h.hpp:35-49)
template<class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
bool updateSmoothingLength(const cstone::GroupView& grp, Dataset& d)
{
    return sphx_compat::updateSmoothingLength(grp, d);
}

// the VE loops (hydro_ve/xmass.hpp:69-75, ve_def_gradh.hpp:80-86, eos.hpp:191-197,
// iad_divv_curlv.hpp:93-98, av_switches.hpp:81-87, momentum_energy.hpp:107-114) and
// rhoTimestep (ts_global.hpp:75-98)
template<typename Tc, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeXMass(const cstone::GroupView& grp, Dataset& d, const cstone::Box<Tc>& box)
{
    sphx_compat::computeXMass(grp, d, box);
}

template<typename Tc, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeVeDefGradh(const cstone::GroupView& grp, Dataset& d, const cstone::Box<Tc>& box)
{
    sphx_compat::computeVeDefGradh(grp, d, box);
}

template<class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeEOS(size_t startIndex, size_t endIndex, Dataset& d)
{
    sphx_compat::computeEOS(startIndex, endIndex, d);
}

template<class Tc, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeIadDivvCurlv(const cstone::GroupView& grp, Dataset& d, const cstone::Box<Tc>& box)
{
    sphx_compat::computeIadDivvCurlv(grp, d, box);
}

template<class T, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeAVswitches(const cstone::GroupView& grp, Dataset& d, const cstone::Box<T>& box)
{
    sphx_compat::computeAVswitches(grp, d, box);
}

template<bool avClean, class T, class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
void computeMomentumEnergy(const cstone::GroupView& grp, float* groupDt, Dataset& d, const cstone::Box<T>& box)
{
    sphx_compat::computeMomentumEnergy<avClean>(grp, groupDt, d, box);
}

template<class Dataset>
    requires sphx_dropin::DeviceDataset<Dataset>
auto rhoTimestep(size_t first, size_t last, const Dataset& d)
{
    return sphx_compat::rhoTimestep(first, last, d);
}

} // namespace sph
