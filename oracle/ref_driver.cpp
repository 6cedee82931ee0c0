/*
 * ref_driver.cpp - extern "C" shims around the reference SPH-EXA CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the reference
 * headers where they lie (SPHEXA_SRC/domain/include, SPHEXA_SRC/sph/include) into
 * oracle/_ref/libsphexa_ref.so.  Nothing of the reference is copied: this file only
 * calls its templates (computeSfcKeys, sort_by_key, computeOctree,
 * updateInternalTree, nodeFpCenters, findNeighbors, findNeighborsSph,
 * computeDensityImpl, computeIdealGasEOS_HydroStd_Impl, computeIADImpl,
 * computeMomentumEnergyStdImpl, createTables pieces) with T = double.
 * It pins the oracle restatement (tests/test_oracle_vs_ref.py) and serves as the
 * "reference" CPU baseline in bench.py.
 */
#include <utility> // the reference's util/strong_type.hpp uses std::move without including it

#include <chrono>
#include <cstring>
#include <numeric>
#include <span>
#include <vector>

#include "cstone/findneighbors.hpp"
#include "cstone/focus/source_center.hpp"
#include "cstone/primitives/gather.hpp"
#include "cstone/sfc/sfc.hpp"
#include "cstone/traversal/groups.hpp"
#include "cstone/tree/csarray.hpp"
#include "cstone/tree/octree.hpp"

#include "sph/kernels.hpp" // updateH must be declared before find_neighbors.hpp's template uses it
#include "sph/find_neighbors.hpp"
#include "sph/hydro_std/density.hpp"
#include "sph/hydro_std/eos.hpp"
#include "sph/hydro_std/iad.hpp"
#include "sph/hydro_std/momentum_energy.hpp"
#include "sph/hydro_ve/av_switches.hpp"
#include "sph/hydro_ve/eos.hpp"
#include "sph/hydro_ve/iad_divv_curlv.hpp"
#include "sph/hydro_ve/momentum_energy.hpp"
#include "sph/hydro_ve/ve_def_gradh.hpp"
#include "sph/hydro_ve/xmass.hpp"
#include "sph/positions.hpp"
#include "sph/sph_kernel_tables.hpp"
#include "sph/update_h.hpp"

#include "../include/sphx.h"

namespace
{
using Key = uint64_t;

cstone::Box<double> toBox(const sphx_box* b)
{
    auto bt = [](int v) { return cstone::BoundaryType(char(v)); };
    return cstone::Box<double>(b->xmin, b->xmax, b->ymin, b->ymax, b->zmin, b->zmax, bt(b->bx), bt(b->by),
                               bt(b->bz));
}

// non-owning field with the .data()/.size() interface the *Impl templates use
template<class T>
struct Field
{
    T*     p = nullptr;
    size_t n = 0;
    T*     data() const { return p; }
    size_t size() const { return n; }
    T&     operator[](size_t i) const { return p[i]; }
    bool   empty() const { return n == 0; }
    friend void swap(Field& a, Field& b)
    {
        std::swap(a.p, b.p);
        std::swap(a.n, b.n);
    }
};

// the subset of sphexa::ParticlesData that the std-hydro *Impl templates read
struct MiniData
{
    using AcceleratorType = cstone::CpuTag;
    using HydroType       = double;
    using RealType        = double;
    using KeyType         = uint64_t;

    Field<double> x, y, z, h, m, vx, vy, vz, u, temp, rho, xm, p, c;
    Field<double> c11, c12, c13, c22, c23, c33, ax, ay, az, du;
    Field<cstone::LocalIndex> neighbors;
    Field<unsigned> nc;
    Field<double> wh, whd;
    double K{0}, gamma{5.0 / 3.0}, muiConst{10.0}, minDtCourant{INFINITY};
    double Kcour{0.2};
    unsigned ngmax{150}, ng0{100};
};

// the subset of sphexa::ParticlesData that updatePositionsHost / updateIntEnergyHost read
struct IntData
{
    using RealType = double;
    Field<double> x, y, z, x_m1, y_m1, z_m1, vx, vy, vz, ax, ay, az, h, u, du, du_m1;
    double        minDt{0}, minDt_m1{0};
};

// the fields of the VE loops, in the order of the oracle's OrcVe
struct OrcVe
{
    const double *x, *y, *z, *vx, *vy, *vz, *h, *m, *c, *xm, *kx, *gradh, *prho, *divv, *alpha;
    const double *c11, *c12, *c13, *c22, *c23, *c33;
    const double *dV11, *dV12, *dV13, *dV22, *dV23, *dV33;
    const double *wh, *whd;
    double        K;
};

// the subset of sphexa::ParticlesData that the VE *Impl templates read
struct VeData
{
    using AcceleratorType = cstone::CpuTag;
    using HydroType       = double;
    using RealType        = double;

    Field<double> x, y, z, h, m, vx, vy, vz, u, temp, c, xm, kx, gradh, prho, tdpdTrho, rho, p, divv, curlv, alpha;
    Field<double> c11, c12, c13, c22, c23, c33, dV11, dV12, dV13, dV22, dV23, dV33, ax, ay, az, du;
    Field<cstone::LocalIndex> neighbors;
    Field<unsigned> nc;
    Field<double> wh, whd;
    double K{0}, sincIndex{6.0}, gamma{5.0 / 3.0}, muiConst{10.0}, minDtCourant{INFINITY}, minDt{0};
    double Kcour{0.2}, alphamin{0.05}, alphamax{1.0}, decay_constant{0.2};
    double Atmin{0.1}, Atmax{0.2}, ramp{10.0};
    unsigned ngmax{150};
};

template<class T>
Field<T> F(const T* p, size_t n)
{
    return Field<T>{const_cast<T*>(p), n};
}

} // namespace

extern "C"
{

void ref_kernel_tables(double sincIndex, double* wh64, float* wh32, double* K)
{
    auto k = sph::getSphKernel(sph::SphKernelType::sinc_n, sincIndex);
    if (K) *K = sph::kernel_3D_k(k, 2.0);
    if (wh64)
    {
        auto t = sph::tabulateFunction<double, sph::lt::kTableSize>(k, 0, 2);
        std::copy(t.begin(), t.end(), wh64);
    }
    if (wh32)
    {
        auto t = sph::tabulateFunction<float, sph::lt::kTableSize>(k, 0, 2);
        std::copy(t.begin(), t.end(), wh32);
    }
}

double ref_sphynx_k(double n) { return sph::sphynx_3D_k(n); }

uint64_t ref_ihilbert(unsigned x, unsigned y, unsigned z) { return cstone::iHilbert<Key>(x, y, z); }

void ref_decode_hilbert(uint64_t key, unsigned* xyz)
{
    auto [a, b, c] = cstone::decodeHilbert<Key>(key);
    xyz[0] = a, xyz[1] = b, xyz[2] = c;
}

void ref_compute_sfc_keys(const double* x, const double* y, const double* z, uint64_t* keys, uint64_t n,
                          const sphx_box* b)
{
    cstone::computeSfcKeys(x, y, z, cstone::sfcKindPointer(keys), n, toBox(b));
}

void ref_sort_by_key(uint64_t* keys, uint32_t* vals, uint64_t n) { cstone::sort_by_key(keys, keys + n, vals); }

int ref_compute_octree(const uint64_t* keys, uint64_t n, unsigned bucket, unsigned maxCount, uint64_t* leaves,
                       unsigned* counts, int cap)
{
    auto [tree, cnt] = cstone::computeOctree<Key>(std::span<const Key>(keys, n), bucket, maxCount);
    int L = int(cstone::nNodes(tree));
    if (cap < L + 1) return -L;
    std::copy(tree.begin(), tree.end(), leaves);
    std::copy(cnt.begin(), cnt.end(), counts);
    return L;
}

void ref_build_octree(const uint64_t* leaves, int numLeaves, uint64_t* prefixes, int* childOffsets, int* parents,
                      int* levelRange, int* internalToLeaf, int* leafToInternal)
{
    cstone::OctreeData<Key, cstone::CpuTag> o;
    o.resize(numLeaves);
    cstone::updateInternalTree<Key>(std::span<const Key>(leaves, numLeaves + 1), o.data());
    std::copy(o.prefixes.begin(), o.prefixes.end(), prefixes);
    std::copy(o.childOffsets.begin(), o.childOffsets.begin() + o.numNodes, childOffsets);
    std::copy(o.parents.begin(), o.parents.end(), parents);
    std::copy(o.levelRange.begin(), o.levelRange.end(), levelRange);
    std::copy(o.internalToLeaf.begin(), o.internalToLeaf.end(), internalToLeaf);
    std::copy(o.leafToInternal.begin(), o.leafToInternal.end(), leafToInternal);
}

void ref_node_centers(const uint64_t* prefixes, int numNodes, const sphx_box* b, double* centers, double* sizes)
{
    cstone::nodeFpCenters<Key>(std::span<const Key>(prefixes, numNodes), reinterpret_cast<cstone::Vec3<double>*>(centers),
                               reinterpret_cast<cstone::Vec3<double>*>(sizes), toBox(b));
}

static cstone::OctreeNsView<double, Key> nsView(int numLeaves, const uint64_t* prefixes, const int* childOffsets,
                                                const int* parents, const int* internalToLeaf, const int* levelRange,
                                                const uint32_t* layout, const double* centers, const double* sizes)
{
    return cstone::OctreeNsView<double, Key>{numLeaves,
                                             prefixes,
                                             childOffsets,
                                             parents,
                                             internalToLeaf,
                                             levelRange,
                                             nullptr,
                                             layout,
                                             reinterpret_cast<const cstone::Vec3<double>*>(centers),
                                             reinterpret_cast<const cstone::Vec3<double>*>(sizes)};
}

void ref_find_neighbors(const double* x, const double* y, const double* z, const double* h, uint32_t first,
                        uint32_t last, const sphx_box* b, int numLeaves, const uint64_t* prefixes,
                        const int* childOffsets, const int* parents, const int* internalToLeaf,
                        const int* levelRange, const uint32_t* layout, const double* centers, const double* sizes,
                        unsigned ngmax, uint32_t* nbrs, unsigned* nc)
{
    auto view = nsView(numLeaves, prefixes, childOffsets, parents, internalToLeaf, levelRange, layout, centers, sizes);
    cstone::findNeighbors(x, y, z, h, first, last, toBox(b), view, ngmax, nbrs, nc);
}

void ref_find_neighbors_sph(const double* x, const double* y, const double* z, double* h, uint32_t first,
                            uint32_t last, const sphx_box* b, int numLeaves, const uint64_t* prefixes,
                            const int* childOffsets, const int* parents, const int* internalToLeaf,
                            const int* levelRange, const uint32_t* layout, const double* centers,
                            const double* sizes, unsigned ng0, unsigned ngmax, uint32_t* nbrs, unsigned* nc)
{
    auto view = nsView(numLeaves, prefixes, childOffsets, parents, internalToLeaf, levelRange, layout, centers, sizes);
    sph::findNeighborsSph(x, y, z, h, first, last, toBox(b), view, ng0, ngmax, nbrs, nc);
}

static MiniData mini(uint64_t n, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax, const double* x,
                     const double* y, const double* z, const double* h, const double* m, const double* wh, double K)
{
    MiniData d;
    d.x = F(x, n), d.y = F(y, n), d.z = F(z, n), d.h = F(h, n), d.m = F(m, n);
    d.neighbors = F(nbrs, n * ngmax);
    d.nc = F(nc, n);
    d.wh = F(wh, size_t(sph::lt::kTableSize));
    d.whd = d.wh;
    d.K = K;
    d.ngmax = ngmax;
    return d;
}

// computeDensityImpl on [first, last) of an n-particle set; rho has n entries
void ref_density(uint64_t n, uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax,
                 const double* x, const double* y, const double* z, const double* h, const double* m,
                 const double* wh, double K, const sphx_box* b, double* rho)
{
    MiniData d = mini(n, nbrs, nc, ngmax, x, y, z, h, m, wh, K);
    std::vector<double> xm(n);
    d.rho = F(rho, n);
    d.xm = F(xm.data(), n);
    cstone::GroupView g{first, last, 1, nullptr, nullptr};
    sph::computeDensityImpl(g, d, toBox(b));
}

void ref_eos(uint64_t n, uint32_t first, uint32_t last, const double* u, const double* rho, double gamma, double* p,
             double* c)
{
    MiniData d;
    d.u = F(u, n), d.rho = F(rho, n), d.p = F(p, n), d.c = F(c, n);
    d.gamma = gamma;
    sph::computeIdealGasEOS_HydroStd_Impl(first, last, d);
}

void ref_iad(uint64_t n, uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax,
             const double* x, const double* y, const double* z, const double* h, const double* m, const double* rho,
             const double* wh, double K, const sphx_box* b, double* c11, double* c12, double* c13, double* c22,
             double* c23, double* c33)
{
    MiniData d = mini(n, nbrs, nc, ngmax, x, y, z, h, m, wh, K);
    d.rho = F(rho, n);
    d.c11 = F(c11, n), d.c12 = F(c12, n), d.c13 = F(c13, n), d.c22 = F(c22, n), d.c23 = F(c23, n),
    d.c33 = F(c33, n);
    sph::computeIADImpl(first, last, d, toBox(b));
}

double ref_momentum_energy(uint64_t n, uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc,
                           unsigned ngmax, const double* x, const double* y, const double* z, const double* vx,
                           const double* vy, const double* vz, const double* h, const double* m, const double* rho,
                           const double* p, const double* c, const double* c11, const double* c12, const double* c13,
                           const double* c22, const double* c23, const double* c33, const double* wh, double K,
                           float Kcour, const sphx_box* b, double* ax, double* ay, double* az, double* du)
{
    MiniData d = mini(n, nbrs, nc, ngmax, x, y, z, h, m, wh, K);
    d.vx = F(vx, n), d.vy = F(vy, n), d.vz = F(vz, n), d.rho = F(rho, n), d.p = F(p, n), d.c = F(c, n);
    d.c11 = F(c11, n), d.c12 = F(c12, n), d.c13 = F(c13, n), d.c22 = F(c22, n), d.c23 = F(c23, n),
    d.c33 = F(c33, n);
    d.ax = F(ax, n), d.ay = F(ay, n), d.az = F(az, n), d.du = F(du, n);
    d.Kcour = Kcour;
    sph::computeMomentumEnergyStdImpl(first, last, d, toBox(b));
    return d.minDtCourant;
}

double ref_xmass_one(uint32_t i, const uint32_t* nb, unsigned cnt, const double* x, const double* y, const double* z,
                     const double* h, const double* m, const double* wh, double K, const sphx_box* b)
{
    return sph::xmassJLoop(i, K, toBox(b), nb, cnt, x, y, z, h, m, wh, wh);
}

/* The reference CPU hydro step on one rank, in the stage order of
 * HydroProp::computeForces (main/src/propagator/std_hydro.hpp:119-168) and the
 * single-rank harness (domain/test/unit/neighbors/findneighbors.cpp:43-91):
 * keys -> stable sort -> gather -> computeOctree -> internal tree -> layout ->
 * node centres -> findNeighborsSph -> density -> EOS -> IAD -> momentum-energy.
 * Inputs are unsorted host arrays of n particles; outputs are in SFC order.
 * stage_s (11 entries) receives the wall time of each stage.  Returns
 * minDtCourant. */
double ref_hydro_step(uint64_t n, const double* x0, const double* y0, const double* z0, const double* h0,
                      const double* m0, const double* vx0, const double* vy0, const double* vz0, const double* u0,
                      const sphx_box* b, unsigned bucket, unsigned ng0, unsigned ngmax, double sincIndex, double gamma,
                      float Kcour, double* rho_out, double* ax_out, double* du_out, unsigned* nc_out,
                      double* stage_s)
{
    using clk = std::chrono::steady_clock;
    auto box = toBox(b);
    auto t0 = clk::now();
    auto lap = [&](int k) {
        auto t = clk::now();
        stage_s[k] = std::chrono::duration<double>(t - t0).count();
        t0 = t;
    };
    std::vector<double> wh(sph::lt::kTableSize);
    double K;
    ref_kernel_tables(sincIndex, wh.data(), nullptr, &K);

    std::vector<Key> keys(n, 0);
    cstone::computeSfcKeys(x0, y0, z0, cstone::sfcKindPointer(keys.data()), n, box);
    lap(0);
    std::vector<cstone::LocalIndex> order(n);
    std::iota(order.begin(), order.end(), 0u);
    cstone::sort_by_key(keys.begin(), keys.end(), order.begin());
    std::vector<double> x(n), y(n), z(n), h(n), m(n), vx(n), vy(n), vz(n), u(n);
    std::span<const cstone::LocalIndex> ord(order);
    cstone::gather(ord, x0, x.data());
    cstone::gather(ord, y0, y.data());
    cstone::gather(ord, z0, z.data());
    cstone::gather(ord, h0, h.data());
    cstone::gather(ord, m0, m.data());
    cstone::gather(ord, vx0, vx.data());
    cstone::gather(ord, vy0, vy.data());
    cstone::gather(ord, vz0, vz.data());
    cstone::gather(ord, u0, u.data());
    lap(1);
    auto [leaves, counts] = cstone::computeOctree<Key>(std::span<const Key>(keys), bucket);
    lap(2);
    cstone::OctreeData<Key, cstone::CpuTag> o;
    o.resize(cstone::nNodes(leaves));
    cstone::updateInternalTree<Key>(leaves, o.data());
    std::vector<cstone::LocalIndex> layout(cstone::nNodes(leaves) + 1, 0);
    std::inclusive_scan(counts.begin(), counts.end(), layout.begin() + 1);
    std::vector<cstone::Vec3<double>> centers(o.numNodes), sizes(o.numNodes);
    cstone::nodeFpCenters<Key>(std::span<const Key>(o.prefixes.data(), o.numNodes), centers.data(), sizes.data(),
                               box);
    lap(3);
    cstone::OctreeNsView<double, Key> view{o.numLeafNodes,
                                          o.prefixes.data(),
                                          o.childOffsets.data(),
                                          o.parents.data(),
                                          o.internalToLeaf.data(),
                                          o.levelRange.data(),
                                          nullptr,
                                          layout.data(),
                                          centers.data(),
                                          sizes.data()};
    std::vector<cstone::LocalIndex> nbrs(size_t(n) * ngmax);
    std::vector<unsigned> nc(n);
    sph::findNeighborsSph(x.data(), y.data(), z.data(), h.data(), 0, n, box, view, ng0, ngmax, nbrs.data(),
                          nc.data());
    lap(4);
    std::vector<double> rho(n), xm(n), p(n), c(n), c11(n), c12(n), c13(n), c22(n), c23(n), c33(n), ax(n), ay(n),
        az(n), du(n);
    MiniData d = mini(n, nbrs.data(), nc.data(), ngmax, x.data(), y.data(), z.data(), h.data(), m.data(), wh.data(),
                      K);
    d.vx = F(vx.data(), n), d.vy = F(vy.data(), n), d.vz = F(vz.data(), n), d.u = F(u.data(), n);
    d.rho = F(rho.data(), n), d.xm = F(xm.data(), n), d.p = F(p.data(), n), d.c = F(c.data(), n);
    d.c11 = F(c11.data(), n), d.c12 = F(c12.data(), n), d.c13 = F(c13.data(), n);
    d.c22 = F(c22.data(), n), d.c23 = F(c23.data(), n), d.c33 = F(c33.data(), n);
    d.ax = F(ax.data(), n), d.ay = F(ay.data(), n), d.az = F(az.data(), n), d.du = F(du.data(), n);
    d.gamma = gamma;
    d.Kcour = Kcour;
    cstone::GroupView g{0, cstone::LocalIndex(n), 1, nullptr, nullptr};
    sph::computeDensityImpl(g, d, box);
    lap(5);
    sph::computeIdealGasEOS_HydroStd_Impl(0, n, d);
    lap(6);
    sph::computeIADImpl(0, n, d, box);
    lap(7);
    sph::computeMomentumEnergyStdImpl(0, n, d, box);
    lap(8);
    if (rho_out) std::copy(rho.begin(), rho.end(), rho_out);
    if (ax_out) std::copy(ax.begin(), ax.end(), ax_out);
    if (du_out) std::copy(du.begin(), du.end(), du_out);
    if (nc_out) std::copy(nc.begin(), nc.end(), nc_out);
    stage_s[9] = 0;
    stage_s[10] = 0;
    return d.minDtCourant;
}


// sph::updatePositionsHost + updateIntEnergyHost (sph/include/sph/positions.hpp:113-179),
// the CPU branch of computePositions (:207-230)
void ref_compute_positions(uint64_t n, uint32_t first, uint32_t last, double* x, double* y, double* z, double* x_m1,
                           double* y_m1, double* z_m1, double* vx, double* vy, double* vz, const double* ax,
                           const double* ay, const double* az, const double* h, double* u, const double* du,
                           double* du_m1, double dt, double dt_m1, const sphx_box* b)
{
    IntData d;
    d.x = F(x, n), d.y = F(y, n), d.z = F(z, n), d.x_m1 = F(x_m1, n), d.y_m1 = F(y_m1, n), d.z_m1 = F(z_m1, n);
    d.vx = F(vx, n), d.vy = F(vy, n), d.vz = F(vz, n), d.ax = F(ax, n), d.ay = F(ay, n), d.az = F(az, n);
    d.h = F(h, n), d.u = F(u, n), d.du = F(du, n), d.du_m1 = F(du_m1, n);
    d.minDt = dt, d.minDt_m1 = dt_m1;
    auto box = toBox(b);
    sph::updatePositionsHost(first, last, d, box);
    sph::updateIntEnergyHost(first, last, d, box);
}

// sph::updateSmoothingLengthCpu (sph/include/sph/update_h.hpp:13-32)
int ref_update_smoothing_length(uint32_t first, uint32_t last, unsigned ng0, const unsigned* nc, double* h,
                                uint64_t* keys)
{
    return sph::updateSmoothingLengthCpu(first, last, ng0, nc, h, keys) ? 1 : 0;
}

// the same with the reference app's FP32 h (HydroType = float, sph/types.hpp:39-46)
int ref_update_smoothing_length_f32(uint32_t first, uint32_t last, unsigned ng0, const unsigned* nc, float* h,
                                    uint64_t* keys)
{
    return sph::updateSmoothingLengthCpu(first, last, ng0, nc, h, keys) ? 1 : 0;
}


// ---- VE loops: the reference's computeXMassImpl, computeVeDefGradhImpl,
// computeIdealGasEOS_Impl, computeIadDivvCurlvImpl, computeAVswitchesImpl and
// computeMomentumEnergyImpl<avClean> (sph/include/sph/hydro_ve/*.hpp), T = double
static VeData veData(uint64_t n, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax, uint32_t first,
                     uint32_t last, const OrcVe* f)
{
    VeData d;
    auto   c = [n](const double* p) { return p ? F(p, n) : Field<double>{}; };
    d.x = c(f->x), d.y = c(f->y), d.z = c(f->z), d.vx = c(f->vx), d.vy = c(f->vy), d.vz = c(f->vz);
    d.h = c(f->h), d.m = c(f->m), d.c = c(f->c), d.xm = c(f->xm), d.kx = c(f->kx), d.gradh = c(f->gradh);
    d.prho = c(f->prho), d.divv = c(f->divv), d.alpha = c(f->alpha);
    d.c11 = c(f->c11), d.c12 = c(f->c12), d.c13 = c(f->c13), d.c22 = c(f->c22), d.c23 = c(f->c23), d.c33 = c(f->c33);
    d.dV11 = c(f->dV11), d.dV12 = c(f->dV12), d.dV13 = c(f->dV13), d.dV22 = c(f->dV22), d.dV23 = c(f->dV23);
    d.dV33 = c(f->dV33);
    d.wh = F(f->wh, size_t(sph::lt::kTableSize)), d.whd = F(f->whd, size_t(sph::lt::kTableSize));
    d.K = f->K, d.ngmax = ngmax;
    d.neighbors = F(nbrs, size_t(ngmax) * (last - first));
    d.nc        = F(nc, n);
    return d;
}

void ref_ve_xmass(uint64_t n, uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc,
                  unsigned ngmax, const OrcVe* f, const sphx_box* b, double* xm)
{
    VeData d = veData(n, nbrs, nc, ngmax, first, last, f);
    d.xm     = F(xm, n);
    sph::computeXMassImpl(first, last, d, toBox(b));
}

void ref_ve_def_gradh(uint64_t n, uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc,
                      unsigned ngmax, const OrcVe* f, const sphx_box* b, double* kx, double* gradh)
{
    VeData d = veData(n, nbrs, nc, ngmax, first, last, f);
    d.kx = F(kx, n), d.gradh = F(gradh, n);
    sph::computeVeDefGradhImpl(first, last, d, toBox(b));
}

void ref_ve_eos(uint64_t n, uint32_t first, uint32_t last, const double* u, const double* m, const double* kx,
                const double* xm, const double* gradh, double gamma, double* prho, double* c, double* rho, double* p)
{
    VeData d;
    d.u = F(u, n), d.m = F(m, n), d.kx = F(kx, n), d.xm = F(xm, n), d.gradh = F(gradh, n);
    d.prho = F(prho, n), d.c = F(c, n);
    if (rho) d.rho = F(rho, n);
    if (p) d.p = F(p, n);
    d.gamma = gamma;
    sph::computeIdealGasEOS_Impl(first, last, d);
}

void ref_ve_iad_divv_curlv(uint64_t n, uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc,
                           unsigned ngmax, const OrcVe* f, const sphx_box* b, int /*givenIad: not supported*/,
                           double* c11, double* c12, double* c13,
                           double* c22, double* c23, double* c33, double* divv, double* curlv, double* dV11,
                           double* dV12, double* dV13, double* dV22, double* dV23, double* dV33)
{
    VeData d = veData(n, nbrs, nc, ngmax, first, last, f);
    d.c11 = F(c11, n), d.c12 = F(c12, n), d.c13 = F(c13, n), d.c22 = F(c22, n), d.c23 = F(c23, n);
    d.c33 = F(c33, n), d.divv = F(divv, n);
    d.curlv = curlv ? F(curlv, n) : Field<double>{};
    if (dV11)
        d.dV11 = F(dV11, n), d.dV12 = F(dV12, n), d.dV13 = F(dV13, n), d.dV22 = F(dV22, n), d.dV23 = F(dV23, n),
        d.dV33 = F(dV33, n);
    else
        d.dV11 = d.dV12 = d.dV13 = d.dV22 = d.dV23 = d.dV33 = Field<double>{};
    sph::computeIadDivvCurlvImpl(first, last, d, toBox(b));
}

void ref_ve_av_switches(uint64_t n, uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc,
                        unsigned ngmax, const OrcVe* f, const sphx_box* b, double dt, double alphamin,
                        double alphamax, double decay_constant, double* alpha)
{
    VeData d = veData(n, nbrs, nc, ngmax, first, last, f);
    d.alpha = F(alpha, n);
    d.minDt = dt, d.alphamin = alphamin, d.alphamax = alphamax, d.decay_constant = decay_constant;
    sph::computeAVswitchesImpl(first, last, d, toBox(b));
}

double ref_ve_momentum_energy(uint64_t n, uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc,
                              unsigned ngmax, const OrcVe* f, const sphx_box* b, int avClean, double Atmin,
                              double Atmax, double ramp, float Kcour, double* gpx, double* gpy, double* gpz,
                              double* du)
{
    VeData d = veData(n, nbrs, nc, ngmax, first, last, f);
    d.ax = F(gpx, n), d.ay = F(gpy, n), d.az = F(gpz, n), d.du = F(du, n);
    d.Atmin = Atmin, d.Atmax = Atmax, d.ramp = ramp, d.Kcour = Kcour;
    if (avClean) sph::computeMomentumEnergyImpl<true>(first, last, d, toBox(b));
    else sph::computeMomentumEnergyImpl<false>(first, last, d, toBox(b));
    return d.minDtCourant;
}

// cstone::limitBoxShrinking (domain/include/cstone/sfc/box.hpp:397-414)
void ref_limit_box_shrinking(const sphx_box* fit, const sphx_box* prev, double shrinkLimit, sphx_box* out)
{
    auto r    = cstone::limitBoxShrinking(toBox(fit), toBox(prev), shrinkLimit);
    *out      = *prev;
    out->xmin = r.xmin(), out->xmax = r.xmax();
    out->ymin = r.ymin(), out->ymax = r.ymax();
    out->zmin = r.zmin(), out->zmax = r.zmax();
    out->bx = int32_t(r.boundaryX()), out->by = int32_t(r.boundaryY()), out->bz = int32_t(r.boundaryZ());
}

void ref_kernel_derivative_table(double sincIndex, double* whd)
{
    auto t = sph::tabulateFunction<double, sph::lt::kTableSize>(
        sph::getSphKernelDerivative(sph::SphKernelType::sinc_n, sincIndex), 0, 2);
    std::copy(t.begin(), t.end(), whd);
}

} // extern "C"
