// compat_driver.cpp - TEST INFRASTRUCTURE: one std-hydro time step driven through the C++
// drop-in shims of include/sphx_compat.hpp, the way HydroProp::computeForces and ::integrate
// (main/src/propagator/std_hydro.hpp:119-186) call the reference's entry points, with the
// caller's own device vectors, box, octree view and group view types.
//
// usage: compat_driver <input.bin> <output.bin>
//   input : uint64 n, then x, y, z, h, m, vx, vy, vz, u (n doubles each, unsorted), then
//           box xmin, xmax (double) and the boundary kind (int32) for all three dimensions
//   output: the SFC-sorted x, y, z, h, rho, p, c, ax, ay, az, du (after computeForces),
//           nc (uint32), then minDtCourant and minDt (double), then x, y, z, vx, u, h after
//           integrate (n doubles each)
#include <hip/hip_runtime.h>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "sphx_compat.hpp"

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

template<class T>
struct DevVec
{
    T*     p = nullptr;
    size_t n = 0;
    DevVec() = default;
    explicit DevVec(size_t count)
        : n(count)
    {
        hipCheck(hipMalloc(&p, (count ? count : 1) * sizeof(T)), "hipMalloc");
        hipCheck(hipMemset(p, 0, (count ? count : 1) * sizeof(T)), "hipMemset");
    }
    T*     data() { return p; }
    size_t size() const { return n; }
    void   upload(const T* h) { hipCheck(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice), "upload"); }
    std::vector<T> download() const
    {
        std::vector<T> h(n);
        hipCheck(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "download");
        return h;
    }
};

enum class BoundaryType : char
{
    open     = 0,
    periodic = 1,
    fixed    = 2
};

// the caller's box type (cstone::Box<double> interface)
struct Box
{
    double       lo[3], hi[3];
    BoundaryType bc[3];
    double       xmin() const { return lo[0]; }
    double       xmax() const { return hi[0]; }
    double       ymin() const { return lo[1]; }
    double       ymax() const { return hi[1]; }
    double       zmin() const { return lo[2]; }
    double       zmax() const { return hi[2]; }
    BoundaryType boundaryX() const { return bc[0]; }
    BoundaryType boundaryY() const { return bc[1]; }
    BoundaryType boundaryZ() const { return bc[2]; }
};

struct Vec3
{
    double x, y, z;
};

// the caller's octree (cstone::OctreeView / OctreeNsView fields)
struct Octree
{
    int32_t   numLeafNodes, numInternalNodes, numNodes;
    uint64_t* prefixes;
    int32_t * childOffsets, *parents, *d_levelRange, *internalToLeaf, *leafToInternal;
    uint32_t* layout;
    Vec3 *    centers, *sizes;
};

// the fields of sphexa::ParticlesData that the std-hydro path reads and writes
struct Dataset
{
    DevVec<double>   x, y, z, h, m, vx, vy, vz, u, rho, p, c, c11, c12, c13, c22, c23, c33, ax, ay, az, du, wh;
    DevVec<double>   x_m1, y_m1, z_m1, du_m1;
    DevVec<unsigned> nc;
    DevVec<uint64_t> keys;
    unsigned         ng0 = 100, ngmax = 150;
    double           K = 0, gamma = 5.0 / 3.0, minDtCourant = INFINITY, minDt = 1e-6, minDt_m1 = 1e-6, ttot = 0;
    double           minDtRho = INFINITY, g = 0, etaAcc = 0.2;
    float            Kcour = 0.2f;
};

} // namespace

int main(int argc, char** argv)
{
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
    Box box{};
    for (int d = 0; d < 3; ++d)
    {
        int32_t bc = 0;
        if (std::fread(&box.lo[d], 8, 1, in) != 1 || std::fread(&box.hi[d], 8, 1, in) != 1 ||
            std::fread(&bc, 4, 1, in) != 1)
            return 1;
        box.bc[d] = BoundaryType(bc);
    }
    std::fclose(in);

    try
    {
        // unsorted inputs
        std::vector<DevVec<double>> raw;
        for (auto& f : host)
        {
            raw.emplace_back(n);
            raw.back().upload(f.data());
        }
        Dataset d;
        for (auto* f : {&d.x, &d.y, &d.z, &d.h, &d.m, &d.vx, &d.vy, &d.vz, &d.u, &d.rho, &d.p, &d.c, &d.c11, &d.c12,
                        &d.c13, &d.c22, &d.c23, &d.c33, &d.ax, &d.ay, &d.az, &d.du, &d.x_m1, &d.y_m1, &d.z_m1,
                        &d.du_m1})
            *f = DevVec<double>(n);
        d.nc   = DevVec<unsigned>(n);
        d.keys = DevVec<uint64_t>(n);
        std::vector<double> wh(SPHX_TABLE_SIZE);
        sphx_compat::check(sphx_kernel_tables(6.0, wh.data(), nullptr, &d.K));
        d.wh = DevVec<double>(SPHX_TABLE_SIZE);
        d.wh.upload(wh.data());

        // Domain::sync on one rank: keys, sort, permutation
        sphx_compat::computeSfcKeys(raw[0].data(), raw[1].data(), raw[2].data(), d.keys.data(), n, box);
        std::vector<uint32_t> iota(n);
        std::iota(iota.begin(), iota.end(), 0u);
        DevVec<uint32_t> order(n);
        order.upload(iota.data());
        sphx_compat::sortByKey(d.keys.data(), order.data(), n);
        DevVec<double>* dst[9] = {&d.x, &d.y, &d.z, &d.h, &d.m, &d.vx, &d.vy, &d.vz, &d.u};
        for (int f = 0; f < 9; ++f)
            sphx_compat::gather(order.data(), raw[f].data(), dst[f]->data(), n);

        // cornerstone leaves, internal octree, centres
        auto    cs = sphx_compat::computeOctree(d.keys.data(), n, 64);
        int32_t ni = 0, nn = 0, np = 0;
        sphx_compat::check(sphx_octree_sizes(cs.numLeaves, &ni, &nn, &np));
        DevVec<uint64_t> prefixes(nn);
        DevVec<int32_t>  childOffsets(nn + 1), parents(np), levelRange(23), i2l(nn), l2i(nn);
        DevVec<Vec3>     centers(nn), sizes(nn);
        Octree o{cs.numLeaves, ni, nn, prefixes.data(), childOffsets.data(), parents.data(), levelRange.data(),
                 i2l.data(), l2i.data(), cs.layout, centers.data(), sizes.data()};
        sphx_compat::buildOctreeGpu(cs.leaves, o);
        sphx_compat::nodeFpCenters(prefixes.data(), nn, centers.data(), sizes.data(), box);

        // HydroProp::computeForces
        auto g = sphx_compat::computeGroupSplits(0u, uint32_t(n), d.x.data(), d.y.data(), d.z.data(), d.h.data(),
                                                 cs.leaves, cs.numLeaves, cs.layout, box);
        sphx_compat::findNeighborsSfc(g, d, box, o, nn);
        sphx_compat::computeDensity(g, d, box);
        sphx_compat::computeEOS_HydroStd(0, n, d);
        sphx_compat::computeIAD(g, d, box);
        sphx_compat::computeMomentumEnergySTD(g, d, box);

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

        // HydroProp::integrate
        sphx_compat::computeTimestep(0, n, d);
        sphx_compat::computePositions(g, d, box, float(d.minDt), std::array<float, 1>{float(d.minDt_m1)});
        bool unconverged = sphx_compat::updateSmoothingLength(g, d);
        if (unconverged) std::fprintf(stderr, "note: particles with nc <= 1\n");
        std::fwrite(&d.minDt, 8, 1, out);
        for (auto* f : {&d.x, &d.y, &d.z, &d.vx, &d.u, &d.h})
        {
            auto v = f->download();
            std::fwrite(v.data(), 8, n, out);
        }
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
