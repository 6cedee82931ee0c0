// compat_driver.cpp - TEST INFRASTRUCTURE: one std-hydro time step driven through
// include/sphx_dropin.hpp with the reference's own entry-point names, namespaces and
// argument lists, the way Domain::sync and HydroProp::computeForces / ::integrate
// (main/src/propagator/std_hydro.hpp:103-186) call them.  The reference's types are given
// minimal definitions here (the caller's own), in the reference's namespaces.
//
// usage: compat_driver <input.bin> <output.bin>
//   input : uint64 n, then x, y, z, h, m, vx, vy, vz, u (n doubles each, unsorted), then
//           box xmin, xmax (double) and the boundary kind (int32) for all three dimensions
//   output: the SFC-sorted x, y, z, h, rho, p, c, ax, ay, az, du (after computeForces),
//           nc (uint32), then minDtCourant and minDt (double), then x, y, z, vx, u, h after
//           integrate (n doubles each); then the checks of the CPU-signature functions on
//           host memory: int32 1 if computeOctree on host keys gave updateOctreeGpu's leaves
//           and counts, the plain findNeighbors counts (n uint32, without self) over host
//           copies of the tree built by updateInternalTree and nodeFpCenters on host arrays,
//           and the first kRows rows of its lists (kRows x ngmax uint32)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <span>
#include <vector>

namespace
{

void hipCheck(hipError_t e, const char* what)
{
    if (e != hipSuccess)
    {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

// a device vector with the DeviceVector interface the entry points use
template<class T>
struct DevVec
{
    T*     p = nullptr;
    size_t n = 0;
    DevVec() = default;
    explicit DevVec(size_t count) { resize(count); }
    DevVec(const DevVec&)            = delete;
    DevVec& operator=(const DevVec&) = delete;
    DevVec& operator=(DevVec&& o) noexcept
    {
        std::swap(p, o.p), std::swap(n, o.n);
        return *this;
    }
    ~DevVec()
    {
        if (p) (void)hipFree(p);
    }
    void resize(size_t count)
    {
        if (count <= n && p)
        {
            n = count;
            return;
        }
        T* q = nullptr;
        hipCheck(hipMalloc(&q, (count ? count : 1) * sizeof(T)), "hipMalloc");
        hipCheck(hipMemset(q, 0, (count ? count : 1) * sizeof(T)), "hipMemset");
        if (p)
        {
            hipCheck(hipMemcpy(q, p, n * sizeof(T), hipMemcpyDeviceToDevice), "resize copy");
            (void)hipFree(p);
        }
        p = q, n = count;
    }
    T*       data() { return p; }
    const T* data() const { return p; }
    size_t   size() const { return n; }
    void     upload(const T* h) { hipCheck(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice), "upload"); }
    std::vector<T> download() const
    {
        std::vector<T> h(n);
        hipCheck(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "download");
        return h;
    }
};

} // namespace

// ---- minimal definitions of the reference's types (domain/include/cstone/...)
namespace util
{
template<class T, std::size_t N>
struct array
{
    T        v[N];
    T&       operator[](std::size_t i) { return v[i]; }
    const T& operator[](std::size_t i) const { return v[i]; }
};
} // namespace util

namespace cstone
{
using LocalIndex    = unsigned;
using TreeNodeIndex = int;
enum class BoundaryType : char
{
    open     = 0,
    periodic = 1,
    fixed    = 2
};
template<class T>
class Box
{
public:
    Box(T x0, T x1, T y0, T y1, T z0, T z1, BoundaryType bx, BoundaryType by, BoundaryType bz)
        : lo{x0, y0, z0}
        , hi{x1, y1, z1}
        , bc{bx, by, bz}
    {
    }
    T            xmin() const { return lo[0]; }
    T            xmax() const { return hi[0]; }
    T            ymin() const { return lo[1]; }
    T            ymax() const { return hi[1]; }
    T            zmin() const { return lo[2]; }
    T            zmax() const { return hi[2]; }
    BoundaryType boundaryX() const { return bc[0]; }
    BoundaryType boundaryY() const { return bc[1]; }
    BoundaryType boundaryZ() const { return bc[2]; }

private:
    T            lo[3], hi[3];
    BoundaryType bc[3];
};
template<class T>
using Vec3 = util::array<T, 3>;
struct CpuTag
{
};
struct GpuTag
{
};
template<class KeyType>
struct OctreeView
{
    TreeNodeIndex  numLeafNodes, numInternalNodes, numNodes, numParents;
    KeyType*       prefixes;
    TreeNodeIndex *childOffsets, *parents, *levelRange, *d_levelRange, *internalToLeaf, *leafToInternal;
    KeyType*       leaves;
};
template<class T, class KeyType>
struct OctreeNsView
{
    TreeNodeIndex        numLeafNodes;
    const KeyType*       prefixes;
    const TreeNodeIndex *childOffsets, *parents, *internalToLeaf, *levelRange;
    const KeyType*       leaves;
    const LocalIndex*    layout;
    const Vec3<T>*       centers;
    const Vec3<T>*       sizes;
    float                searchExtFactor{1.0};
};
struct GroupView
{
    LocalIndex        firstBody, lastBody, numGroups;
    const LocalIndex* groupStart;
    const LocalIndex* groupEnd;
};
template<class Accelerator>
class GroupData
{
public:
    GroupView          view() const { return {firstBody, lastBody, numGroups, groupStart, groupEnd}; }
    DevVec<LocalIndex> data;
    LocalIndex         firstBody = 0, lastBody = 0, numGroups = 0;
    LocalIndex*        groupStart = nullptr;
    LocalIndex*        groupEnd   = nullptr;
};
} // namespace cstone

#include "sphx_dropin.hpp"

namespace
{

// the fields of sphexa::ParticlesData<GpuTag> that the std-hydro path reads and writes
struct Dataset
{
    using AcceleratorType = cstone::GpuTag;
    using RealType        = double;
    DevVec<double>   x, y, z, h, m, vx, vy, vz, u, rho, p, c, c11, c12, c13, c22, c23, c33, ax, ay, az, du, wh;
    DevVec<double>   x_m1, y_m1, z_m1, du_m1;
    DevVec<unsigned> nc;
    DevVec<uint64_t> keys;
    unsigned         ng0 = 100, ngmax = 150;
    double           K = 0, gamma = 5.0 / 3.0, minDtCourant = INFINITY, minDt = 1e-6, minDt_m1 = 1e-6, ttot = 0;
    double           minDtRho = INFINITY, g = 0, etaAcc = 0.2;
    float            Kcour = 0.2f;
    cstone::OctreeNsView<double, uint64_t> treeView{};
};

constexpr unsigned kRows = 64;

} // namespace

int main(int argc, char** argv)
{
    using cstone::BoundaryType;
    if (argc != 3)
    {
        std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 1;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 1;
    uint64_t n = 0;
    if (std::fread(&n, 8, 1, in) != 1) return 1;
    std::vector<std::vector<double>> host(9, std::vector<double>(n));
    for (auto& f : host)
        if (std::fread(f.data(), 8, n, in) != n) return 1;
    double       lo[3], hi[3];
    BoundaryType bc[3];
    for (int d = 0; d < 3; ++d)
    {
        int32_t b = 0;
        if (std::fread(&lo[d], 8, 1, in) != 1 || std::fread(&hi[d], 8, 1, in) != 1 || std::fread(&b, 4, 1, in) != 1)
            return 1;
        bc[d] = BoundaryType(b);
    }
    std::fclose(in);
    cstone::Box<double> box(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], bc[0], bc[1], bc[2]);

    try
    {
        // unsorted inputs
        std::vector<DevVec<double>> raw(9);
        for (int f = 0; f < 9; ++f)
        {
            raw[f].resize(n);
            raw[f].upload(host[f].data());
        }
        Dataset d;
        for (auto* f : {&d.x, &d.y, &d.z, &d.h, &d.m, &d.vx, &d.vy, &d.vz, &d.u, &d.rho, &d.p, &d.c, &d.c11, &d.c12,
                        &d.c13, &d.c22, &d.c23, &d.c33, &d.ax, &d.ay, &d.az, &d.du, &d.x_m1, &d.y_m1, &d.z_m1,
                        &d.du_m1})
            f->resize(n);
        d.nc.resize(n);
        d.keys.resize(n);
        std::vector<double> wh(SPHX_TABLE_SIZE);
        sphx_compat::check(sphx_kernel_tables(6.0, wh.data(), nullptr, &d.K));
        d.wh.resize(SPHX_TABLE_SIZE);
        d.wh.upload(wh.data());

        // Domain::sync on one rank: keys, sort, permutation (domain/assignment.hpp:108-111)
        cstone::computeSfcKeys<true>(raw[0].data(), raw[1].data(), raw[2].data(), d.keys.data(), n, box);
        std::vector<unsigned> iota(n);
        std::iota(iota.begin(), iota.end(), 0u);
        DevVec<unsigned> order(n);
        order.upload(iota.data());
        DevVec<uint64_t> keyBuf;
        DevVec<unsigned> valBuf;
        cstone::sortByKey<true>(std::span<uint64_t>(d.keys.data(), n), std::span<unsigned>(order.data(), n), keyBuf,
                                valBuf, 1.05);
        DevVec<double>* dst[9] = {&d.x, &d.y, &d.z, &d.h, &d.m, &d.vx, &d.vy, &d.vz, &d.u};
        for (int f = 0; f < 9; ++f)
            cstone::gatherAcc<true>(std::span<const unsigned>(order.data(), n), raw[f].data(), dst[f]->data());

        // cornerstone leaves, internal octree, layout, centres
        DevVec<uint64_t> tree(2), tmpTree;
        DevVec<unsigned> counts(1), work;
        std::span<const uint64_t> keySpan(d.keys.data(), n);
        while (!cstone::updateOctreeGpu(keySpan, 64, tree, counts, tmpTree, work))
            ;
        int32_t L  = int32_t(counts.size());
        int32_t ni = 0, nn = 0, np = 0;
        sphx_compat::check(sphx_octree_sizes(L, &ni, &nn, &np));
        DevVec<uint64_t>      prefixes(nn);
        DevVec<int32_t>       childOffsets(nn + 1), parents(np), levelRange(23), i2l(nn), l2i(nn);
        DevVec<cstone::Vec3<double>> centers(nn), sizes(nn);
        cstone::OctreeView<uint64_t> o{L, ni, nn, np, prefixes.data(), childOffsets.data(), parents.data(),
                                       levelRange.data(), levelRange.data(), i2l.data(), l2i.data(), tree.data()};
        cstone::buildOctreeGpu(tree.data(), o);
        cstone::nodeFpCenters(std::span<const uint64_t>(prefixes.data(), nn), centers.data(), sizes.data(), box);
        auto                  hc = counts.download();
        std::vector<unsigned> hl(size_t(L) + 1, 0);
        std::partial_sum(hc.begin(), hc.end(), hl.begin() + 1);
        DevVec<unsigned> layout(size_t(L) + 1);
        layout.upload(hl.data());
        d.treeView = {L,           prefixes.data(), childOffsets.data(), parents.data(), i2l.data(), levelRange.data(),
                      tree.data(), layout.data(),   centers.data(),      sizes.data()};

        // HydroProp::computeForces (std_hydro.hpp:119-168)
        cstone::GroupData<cstone::GpuTag> groups;
        sph::computeGroups(0, n, d, box, groups);
        sph::findNeighborsSfc(0, n, d, box);
        sph::computeDensity(groups.view(), d, box);
        sph::computeEOS_HydroStd(0, n, d);
        sph::computeIAD(groups.view(), d, box);
        sph::computeMomentumEnergySTD(groups.view(), d, box);

        FILE* out = std::fopen(argv[2], "wb");
        if (!out) return 1;
        for (auto* f : {&d.x, &d.y, &d.z, &d.h, &d.rho, &d.p, &d.c, &d.ax, &d.ay, &d.az, &d.du})
        {
            auto v = f->download();
            std::fwrite(v.data(), 8, n, out);
        }
        auto nc = d.nc.download();
        std::fwrite(nc.data(), 4, n, out);
        std::fwrite(&d.minDtCourant, 8, 1, out);

        // the CPU-signature functions on host memory, before integrate moves the particles
        std::vector<uint64_t> hkeys(n);
        hipCheck(hipMemcpy(hkeys.data(), d.keys.data(), n * 8, hipMemcpyDeviceToHost), "keys");
        auto [cpuTree, cpuCounts] = cstone::computeOctree(std::span<const uint64_t>(hkeys), 64);
        int32_t sameTree = cpuTree == tree.download() && cpuCounts == hc ? 1 : 0;
        std::vector<uint64_t> hPrefixes(nn);
        std::vector<int32_t>  hChild(nn + 1), hParents(np), hLevel(23), hi2l(nn), hl2i(nn);
        cstone::OctreeView<uint64_t> ho{L, ni, nn, np, hPrefixes.data(), hChild.data(), hParents.data(), hLevel.data(),
                                        hLevel.data(), hi2l.data(), hl2i.data(), cpuTree.data()};
        cstone::updateInternalTree(std::span<const uint64_t>(cpuTree), ho);
        std::vector<cstone::Vec3<double>> hCenters(nn), hSizes(nn);
        cstone::nodeFpCenters(std::span<const uint64_t>(hPrefixes), hCenters.data(), hSizes.data(), box);
        std::vector<double> hx = d.x.download(), hy = d.y.download(), hz = d.z.download(), hh = d.h.download();
        cstone::OctreeNsView<double, uint64_t> hView{L,           hPrefixes.data(), hChild.data(), hParents.data(),
                                                     hi2l.data(), hLevel.data(),    cpuTree.data(), hl.data(),
                                                     hCenters.data(), hSizes.data()};
        const unsigned        ngmax = d.ngmax;
        std::vector<unsigned> plainNb(n * ngmax), plainNc(n);
        cstone::findNeighbors(hx.data(), hy.data(), hz.data(), hh.data(), 0u, unsigned(n), box, hView, ngmax,
                              plainNb.data(), plainNc.data());

        // HydroProp::integrate (std_hydro.hpp:170-186)
        sph::computeTimestep(0, n, d);
        sph::computePositions(groups.view(), d, box, float(d.minDt), util::array<float, 4>{float(d.minDt_m1)});
        bool unconverged = sph::updateSmoothingLength(groups.view(), d);
        if (unconverged) std::fprintf(stderr, "note: particles with nc <= 1\n");
        std::fwrite(&d.minDt, 8, 1, out);
        for (auto* f : {&d.x, &d.y, &d.z, &d.vx, &d.u, &d.h})
        {
            auto v = f->download();
            std::fwrite(v.data(), 8, n, out);
        }
        std::fwrite(&sameTree, 4, 1, out);
        std::fwrite(plainNc.data(), 4, n, out);
        std::fwrite(plainNb.data(), 4, size_t(std::min<uint64_t>(kRows, n)) * ngmax, out);
        std::fclose(out);
    }
    catch (const sphx_compat::Error& e)
    {
        std::fprintf(stderr, "sphx_compat error %d: %s\n", e.code, e.what());
        return 3;
    }
    std::printf("compat driver ok: n=%llu\n", (unsigned long long)n);
    return 0;
}
