/*
 * sphx_oracle.cpp - CPU restatement of SPH-EXA's single-rank std-hydro path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke() entry
 * point and bench.py's cpu_baseline leg load it; the product library
 * (sphexa_amd/csrc) never links or calls it.
 *
 * Every function restates the reference algorithm it names (file:line relative to
 * the SPH-EXA source tree) with the same operation order, in double precision,
 * parallelised with OpenMP like the reference CPU path.  Built with
 * -ffp-contract=off so that every product and sum rounds separately (the
 * reference Release build would otherwise contract to FMA per compiler whim).
 *
 * How each part is pinned (see tests/test_oracle.py):
 *   keys        decode(encode) = id, curve continuity, first-level octant map,
 *               hilbertIBox known answers (domain/test/unit/sfc/hilbert.cpp:83-296),
 *               and equality with the compiled reference (oracle/_ref).
 *   sort        brute-force std::sort on (key, index) pairs + stability check.
 *   octree      csarray known answers (domain/test/unit/tree/csarray.cpp:102-206),
 *               cornerstone invariants, equality with oracle/_ref.
 *   link        octree.cpp known shapes (4x4x4 -> 73 nodes, irregular 15 leaves),
 *               connectivity invariants, equality with oracle/_ref.
 *   neighbors   brute-force all pairs (domain/test/unit/neighbors/all_to_all.hpp)
 *               and equality with oracle/_ref.
 *   density     ve.cpp XMass known answer (rho0 = 34.515038498081417),
 *               brute-force sums, equality with oracle/_ref.
 *   IAD, ME     std.cpp known answers (sph/test/std.cpp:98-143), oracle/_ref.
 *   tables, K   Simpson known answers (sph/test/table_creation.cpp:17-40),
 *               closed-form integral.
 *   EOS         closed form.
 *   integration positionUpdate time reversal and constant-acceleration kinematics,
 *               energyUpdate closed forms, putInBox folds, the timestep minimum
 *               rule, and equality with oracle/_ref (positions, energy, h update).
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "../include/sphx.h"

namespace
{

using Key = uint64_t;

constexpr unsigned kMaxLevel   = 21;                 // tree/definitions.h:55
constexpr Key      kRemoveKey  = Key(1) << 63;       // tree/definitions.h:70-75
constexpr Key      kRootRange  = Key(1) << 63;       // nodeRange(0), sfc/common.hpp:108-115
constexpr int      kTableSize  = SPHX_TABLE_SIZE;    // sph/table_lookup.hpp:10

// ---------------------------------------------------------------------------
// box (sfc/box.hpp:94-174)
struct Box
{
    double lo[3], hi[3], len[3], ilen[3];
    int    bc[3];
};

Box makeBox(const sphx_box* b)
{
    Box r;
    r.lo[0] = b->xmin, r.hi[0] = b->xmax;
    r.lo[1] = b->ymin, r.hi[1] = b->ymax;
    r.lo[2] = b->zmin, r.hi[2] = b->zmax;
    r.bc[0] = b->bx, r.bc[1] = b->by, r.bc[2] = b->bz;
    for (int d = 0; d < 3; ++d)
    {
        r.len[d]  = r.hi[d] - r.lo[d];
        r.ilen[d] = 1.0 / (r.hi[d] - r.lo[d]);
    }
    return r;
}

// ---------------------------------------------------------------------------
// key bit algebra (sfc/common.hpp)
inline unsigned clz64(Key k) { return k == 0 ? 64u : unsigned(__builtin_clzll(k)); }

inline Key nodeRange(unsigned level) { return Key(1) << (3u * (kMaxLevel - level)); } // common.hpp:108-115

inline unsigned treeLevel(Key range) { return (clz64(range - 1) - 1) / 3; } // common.hpp:156-161

inline int commonPrefix(Key a, Key b) { return int(clz64(a ^ b)) - 1; } // common.hpp:144-148

inline Key encodePlaceholderBit(Key code, int prefixLength) // common.hpp:172-180
{
    int nShifts = 3 * int(kMaxLevel) - prefixLength;
    return (Key(1) << prefixLength) | (code >> nShifts);
}

inline unsigned decodePrefixLength(Key code) { return 63u - clz64(code); } // common.hpp:190-194

inline Key decodePlaceholderBit(Key code) // common.hpp:205-213
{
    unsigned len = decodePrefixLength(code);
    return (code ^ (Key(1) << len)) << (3 * kMaxLevel - len);
}

inline unsigned octalDigit(Key code, unsigned position) // common.hpp:298-302
{
    return unsigned(code >> (3u * (kMaxLevel - position))) & 7u;
}

inline int digitWeight(int digit) // common.hpp:318-322
{
    int fourGeqMask = -int(digit >= 4);
    return ((7 - digit) & fourGeqMask) - (digit & ~fourGeqMask);
}

inline unsigned log8ceil(Key n) // common.hpp:118-124
{
    if (n == 0) return 0;
    return kMaxLevel - (clz64(n - 1) - 1) / 3;
}

// ---------------------------------------------------------------------------
// Hilbert curve (sfc/hilbert.hpp:43-94, 130-173, 259-275)
Key iHilbert(unsigned px, unsigned py, unsigned pz)
{
    static const unsigned mortonToHilbert[8] = {0, 1, 3, 2, 7, 6, 4, 5};
    Key key = 0;
    for (int level = int(kMaxLevel) - 1; level >= 0; --level)
    {
        unsigned xi = (px >> level) & 1u;
        unsigned yi = (py >> level) & 1u;
        unsigned zi = (pz >> level) & 1u;
        unsigned octant = (xi << 2) | (yi << 1) | zi;
        key = (key << 3) + mortonToHilbert[octant];

        px ^= -(xi & ((!yi) | zi));
        py ^= -((xi & (yi | zi)) | (yi & (!zi)));
        pz ^= -((xi & (!yi) & (!zi)) | (yi & (!zi)));

        if (zi)
        {
            unsigned t = px;
            px = py, py = pz, pz = t;
        }
        else if (!yi)
        {
            unsigned t = px;
            px = pz, pz = t;
        }
    }
    return key;
}

void decodeHilbert(Key key, unsigned& ox, unsigned& oy, unsigned& oz)
{
    unsigned px = 0, py = 0, pz = 0;
    for (unsigned level = 0; level < kMaxLevel; ++level)
    {
        unsigned octant = unsigned(key >> (3 * level)) & 7u;
        unsigned xi = octant >> 2u, yi = (octant >> 1u) & 1u, zi = octant & 1u;
        if (yi ^ zi)
        {
            unsigned t = px;
            px = pz, pz = py, py = t;
        }
        else if ((!xi & !yi & !zi) || (xi & yi & zi))
        {
            unsigned t = px;
            px = pz, pz = t;
        }
        unsigned mask = (1u << level) - 1;
        px ^= mask & (-(xi & (yi | zi)));
        py ^= mask & (-((xi & ((!yi) | (!zi))) | ((!xi) & yi & zi)));
        pz ^= mask & (-((xi & (!yi) & (!zi)) | (yi & zi)));
        px |= (xi << level);
        py |= ((xi ^ yi) << level);
        pz |= ((yi ^ zi) << level);
    }
    ox = px, oy = py, oz = pz;
}

// sfc3D, sfc/sfc.hpp:141-178: ix = int(floor(x*mx) - xmin*mx), clamped above
Key sfc3D(double x, double y, double z, const Box& box)
{
    const int    mcoord     = (1 << kMaxLevel) - 1;
    const double cubeLength = double(1u << kMaxLevel);
    double mx = cubeLength * box.ilen[0], my = cubeLength * box.ilen[1], mz = cubeLength * box.ilen[2];
    int ix = int(std::floor(x * mx) - box.lo[0] * mx);
    int iy = int(std::floor(y * my) - box.lo[1] * my);
    int iz = int(std::floor(z * mz) - box.lo[2] * mz);
    ix = std::min(ix, mcoord);
    iy = std::min(iy, mcoord);
    iz = std::min(iz, mcoord);
    return iHilbert(unsigned(ix), unsigned(iy), unsigned(iz));
}

// ---------------------------------------------------------------------------
// cornerstone leaves (tree/csarray.hpp)
unsigned calculateNodeCount(Key a, Key b, const Key* k0, const Key* k1, size_t maxCount) // :77-86
{
    size_t c = std::lower_bound(k0, k1, b) - std::lower_bound(k0, k1, a);
    return unsigned(std::min(c, maxCount));
}

// computeNodeCounts without guesses (:182-235); the guessed variant returns the same counts
void computeNodeCounts(const Key* tree, unsigned* counts, int nNodes, const Key* keys, size_t n, unsigned maxCount)
{
    int firstNode = nNodes, lastNode = nNodes;
    if (n > 0)
    {
        firstNode = int(std::upper_bound(tree, tree + nNodes, keys[0]) - tree) - 1;
        lastNode  = int(std::upper_bound(tree, tree + nNodes, keys[n - 1]) - tree);
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nNodes; ++i)
    {
        if (i < firstNode || i >= lastNode) { counts[i] = 0; }
        else { counts[i] = calculateNodeCount(tree[i], tree[i + 1], keys, keys + n, maxCount); }
    }
}

// siblingAndLevel (:251-264)
void siblingAndLevel(const Key* t, int i, int& sib, unsigned& level)
{
    Key thisNode = t[i];
    level        = treeLevel(t[i + 1] - thisNode);
    if (level == 0)
    {
        sib = -1;
        return;
    }
    sib = int(octalDigit(thisNode, level));
    bool siblings = (t[i - sib + 8] == t[i - sib] + nodeRange(level - 1));
    if (!siblings) sib = -1;
}

// calculateNodeOp (:268-290)
int calculateNodeOp(const Key* t, int i, const unsigned* counts, unsigned bucket)
{
    int sib;
    unsigned level;
    siblingAndLevel(t, i, sib, level);
    if (sib > 0)
    {
        const unsigned* g = counts + i - sib;
        size_t parentCount = 0;
        for (int k = 0; k < 8; ++k)
            parentCount += size_t(g[k]);
        if (parentCount <= size_t(bucket)) return 0;
    }
    if (counts[i] > bucket * 512 && level + 3 < kMaxLevel) return 4096;
    if (counts[i] > bucket * 64 && level + 2 < kMaxLevel) return 512;
    if (counts[i] > bucket * 8 && level + 1 < kMaxLevel) return 64;
    if (counts[i] > bucket && level < kMaxLevel) return 8;
    return 1;
}

// updateOctree (:409-426) = rebalanceDecision (:310-329) + rebalanceTree (:376-389) + counts
bool updateOctree(const Key* keys, size_t n, unsigned bucket, std::vector<Key>& tree, std::vector<unsigned>& counts,
                  unsigned maxCount)
{
    int nNodes = int(tree.size()) - 1;
    std::vector<int> ops(nNodes + 1);
    bool converged = true;
#pragma omp parallel for schedule(static) reduction(&& : converged)
    for (int i = 0; i < nNodes; ++i)
    {
        ops[i] = calculateNodeOp(tree.data(), i, counts.data(), bucket);
        converged = converged && (ops[i] == 1);
    }
    ops[nNodes] = 0;
    // exclusive scan
    int run = 0;
    for (int i = 0; i <= nNodes; ++i)
    {
        int v = ops[i];
        ops[i] = run;
        run += v;
    }
    std::vector<Key> newTree(ops[nNodes] + 1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nNodes; ++i) // processNode (:340-366)
    {
        Key thisNode = tree[i];
        unsigned level = treeLevel(tree[i + 1] - thisNode);
        int opCode = ops[i + 1] - ops[i];
        int out = ops[i];
        if (opCode == 1) { newTree[out] = thisNode; }
        else if (opCode == 8)
        {
            for (int s = 0; s < 8; ++s)
                newTree[out + s] = thisNode + Key(s) * nodeRange(level + 1);
        }
        else if (opCode > 8)
        {
            unsigned levelDiff = log8ceil(Key(opCode));
            for (int s = 0; s < opCode; ++s)
                newTree[out + s] = thisNode + Key(s) * nodeRange(level + levelDiff);
        }
    }
    newTree.back() = tree.back();
    tree.swap(newTree);
    counts.resize(tree.size() - 1);
    computeNodeCounts(tree.data(), counts.data(), int(tree.size()) - 1, keys, n, maxCount);
    return converged;
}

// ---------------------------------------------------------------------------
// internal octree (tree/octree.hpp:56-196)
int binaryKeyWeight(Key key, unsigned level) // :56-65
{
    int ret = 0;
    for (unsigned l = 1; l <= level + 1; ++l)
        ret += digitWeight(int(octalDigit(key, l)));
    return ret;
}

// ---------------------------------------------------------------------------
// SPH helpers
inline double lookup(const double* table, double v) // sph/table_lookup.hpp:13-26
{
    constexpr int    numIntervals = kTableSize - 1;
    constexpr double support      = 2.0;
    constexpr double dx           = support / numIntervals;
    constexpr double invDx        = 1.0 / dx;
    int    idx        = int(v * invDx);
    double derivative = (idx >= numIntervals) ? 0.0 : (table[idx + 1] - table[idx]) * invDx;
    return (idx >= numIntervals) ? 0.0 : table[idx] + derivative * (v - double(idx) * dx);
}

// applyPBC legacy threshold fold (sfc/box.hpp:217-238)
inline void applyPBC(const Box& box, double r, double& xx, double& yy, double& zz)
{
    double* c[3] = {&xx, &yy, &zz};
    for (int d = 0; d < 3; ++d)
    {
        if (box.bc[d] != SPHX_BC_PERIODIC) continue;
        if (*c[d] > r) *c[d] -= box.len[d];
        else if (*c[d] < -r) *c[d] += box.len[d];
    }
}

inline double wharmonic(double v) // sph/kernels.hpp:36-43
{
    if (v == 0.0) return 1.0;
    const double Pv = M_PI_2 * v;
    return std::sin(Pv) / Pv;
}

double simpson(double a, double b, uint64_t n, double sincIndex) // sph_kernel_tables.hpp:28-57
{
    auto f = [sincIndex](double x) { return 4.0 * M_PI * x * x * std::pow(wharmonic(x), sincIndex); };
    uint64_t numOdd  = n / 2;
    uint64_t numEven = (numOdd >= 1) ? numOdd - 1 : 0;
    double   h       = (b - a) / double(n);
    std::vector<double> odd(numOdd), even(numEven);
    for (uint64_t i = 0; i < numOdd; ++i)
        odd[i] = f(a + double(2 * (i + 1) - 1) * h);
    for (uint64_t i = 0; i < numEven; ++i)
        even[i] = f(a + double(2 * (i + 1)) * h);
    std::sort(odd.begin(), odd.end());
    std::sort(even.begin(), even.end());
    double so = 0.0, se = 0.0;
    for (double v : odd)
        so += v;
    for (double v : even)
        se += v;
    return h / 3.0 * (f(a) + f(b) + 4.0 * so + 2.0 * se);
}

} // namespace

extern "C"
{

// --------------------------------------------------------------------------- tables
// ParticlesData::createTables (particles_data.hpp:377-387) with HydroType = double and float
void orc_kernel_tables(double sincIndex, double* wh64, float* wh32, double* K)
{
    if (K) *K = 1.0 / simpson(0.0, 2.0, 2000, sincIndex); // kernel_3D_k, sph_kernel_tables.hpp:78-85
    const int numIntervals = kTableSize - 1;              // tabulateFunction, :88-102
    if (wh64)
    {
        const double dx = (2.0 - 0.0) / numIntervals;
        for (int i = 0; i < kTableSize; ++i)
        {
            double v = 0.0 + double(size_t(i)) * dx;
            wh64[i]  = std::pow(wharmonic(v), sincIndex);
        }
    }
    if (wh32)
    {
        const float dx = float((2.0 - 0.0) / numIntervals);
        for (int i = 0; i < kTableSize; ++i)
        {
            float v = float(0.0 + double(float(size_t(i)) * dx));
            wh32[i] = float(std::pow(wharmonic(double(v)), sincIndex));
        }
    }
}

// sphynx_3D_k (sph_kernel_tables.hpp:64-75)
double orc_sphynx_k(double n)
{
    double b0 = 2.7012593e-2, b1 = 2.0410827e-2, b2 = 3.7451957e-3, b3 = 4.7013839e-2;
    return b0 + b1 * std::sqrt(n) + b2 * n + b3 * std::sqrt(n * n * n);
}

double orc_simpson_sin(double a, double b, uint64_t n) // util::simpson on sin, for the Simpson known answer
{
    double h = (b - a) / double(n);
    uint64_t numOdd = n / 2, numEven = numOdd >= 1 ? numOdd - 1 : 0;
    std::vector<double> odd(numOdd), even(numEven);
    for (uint64_t i = 0; i < numOdd; ++i)
        odd[i] = std::sin(a + double(2 * (i + 1) - 1) * h);
    for (uint64_t i = 0; i < numEven; ++i)
        even[i] = std::sin(a + double(2 * (i + 1)) * h);
    std::sort(odd.begin(), odd.end());
    std::sort(even.begin(), even.end());
    double so = 0, se = 0;
    for (double v : odd)
        so += v;
    for (double v : even)
        se += v;
    return h / 3.0 * (std::sin(a) + std::sin(b) + 4.0 * so + 2.0 * se);
}

double orc_lookup(const double* table, double v) { return lookup(table, v); }

// --------------------------------------------------------------------------- keys
uint64_t orc_ihilbert(unsigned x, unsigned y, unsigned z) { return iHilbert(x, y, z); }

void orc_decode_hilbert(uint64_t key, unsigned* xyz) { decodeHilbert(key, xyz[0], xyz[1], xyz[2]); }

// computeSfcKeys (sfc/sfc.hpp:267-275)
void orc_compute_sfc_keys(const double* x, const double* y, const double* z, uint64_t* keys, uint64_t n,
                          const sphx_box* b)
{
    Box box = makeBox(b);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i)
        if (keys[i] != kRemoveKey) keys[i] = sfc3D(x[i], y[i], z[i], box);
}

// sort_by_key (primitives/gather.hpp:42-67): std::stable_sort of (key, value) pairs
void orc_sort_by_key(uint64_t* keys, uint32_t* vals, uint64_t n)
{
    std::vector<std::pair<uint64_t, uint32_t>> kv(n);
    for (uint64_t i = 0; i < n; ++i)
        kv[i] = {keys[i], vals[i]};
    std::stable_sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (uint64_t i = 0; i < n; ++i)
        keys[i] = kv[i].first, vals[i] = kv[i].second;
}

// --------------------------------------------------------------------------- octree
// computeOctree (tree/csarray.hpp:430-440).  Returns the number of leaves L, or -(L)
// if cap < L + 1 (nothing written then).
int orc_compute_octree(const uint64_t* keys, uint64_t n, unsigned bucket, unsigned maxCount, uint64_t* leaves,
                       unsigned* counts, int cap)
{
    std::vector<Key> tree{0, kRootRange};
    std::vector<unsigned> cnt{unsigned(n)};
    while (!updateOctree(keys, n, bucket, tree, cnt, maxCount))
        ;
    int L = int(tree.size()) - 1;
    if (cap < L + 1) return -L;
    std::copy(tree.begin(), tree.end(), leaves);
    std::copy(cnt.begin(), cnt.end(), counts);
    return L;
}

// single rebalance step on a given tree (updateOctree, csarray.hpp:409-426); returns
// the converged flag; tree/counts capacities must be >= the new sizes (cap).
int orc_update_octree(const uint64_t* keys, uint64_t n, unsigned bucket, uint64_t* tree, unsigned* counts,
                      int* numNodes, int cap)
{
    std::vector<Key> t(tree, tree + *numNodes + 1);
    std::vector<unsigned> c(counts, counts + *numNodes);
    bool conv = updateOctree(keys, n, bucket, t, c, ~0u);
    if (int(t.size()) > cap) return -1;
    std::copy(t.begin(), t.end(), tree);
    std::copy(c.begin(), c.end(), counts);
    *numNodes = int(t.size()) - 1;
    return conv ? 1 : 0;
}

// rebalanceDecision (csarray.hpp:310-329) on its own, for the known-answer tests
int orc_rebalance_decision(const uint64_t* tree, const unsigned* counts, int nNodes, unsigned bucket, int* ops)
{
    int conv = 1;
    for (int i = 0; i < nNodes; ++i)
    {
        ops[i] = calculateNodeOp(tree, i, counts, bucket);
        if (ops[i] != 1) conv = 0;
    }
    return conv;
}

// buildOctreeCpu (tree/octree.hpp:168-196)
void orc_build_octree(const uint64_t* leaves, int numLeaves, uint64_t* prefixes, int* childOffsets, int* parents,
                      int* levelRange, int* internalToLeaf, int* leafToInternal)
{
    int numInternal = (numLeaves - 1) / 7;
    int numNodes    = numLeaves + numInternal;
    // createUnsortedLayoutCpu (:78-100)
#pragma omp parallel for schedule(static)
    for (int tid = 0; tid < numLeaves; ++tid)
    {
        Key key = leaves[tid];
        unsigned level = treeLevel(leaves[tid + 1] - key);
        prefixes[tid + numInternal]       = encodePlaceholderBit(key, 3 * int(level));
        internalToLeaf[tid + numInternal] = tid + numInternal;
        unsigned prefixLength = unsigned(commonPrefix(key, leaves[tid + 1]));
        if (prefixLength % 3 == 0 && tid < numLeaves - 1)
        {
            int octIndex = (tid + binaryKeyWeight(key, prefixLength / 3)) / 7;
            prefixes[octIndex]       = encodePlaceholderBit(key, int(prefixLength));
            internalToLeaf[octIndex] = octIndex;
        }
    }
    // sort_by_key(prefixes, internalToLeaf)
    orc_sort_by_key(prefixes, reinterpret_cast<uint32_t*>(internalToLeaf), uint64_t(numNodes));
    for (int i = 0; i < numNodes; ++i)
        leafToInternal[internalToLeaf[i]] = i;
    for (int i = 0; i < numNodes; ++i)
        internalToLeaf[i] -= numInternal;
    // getLevelRangeCpu (:152-160)
    for (unsigned level = 0; level <= kMaxLevel; ++level)
        levelRange[level] =
            int(std::lower_bound(prefixes, prefixes + numNodes, encodePlaceholderBit(0, 3 * int(level))) - prefixes);
    levelRange[kMaxLevel + 1] = numNodes;
    std::fill(childOffsets, childOffsets + numNodes, 0);
    // linkTreeCpu (:115-148)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < numInternal; ++i)
    {
        int idxA = leafToInternal[i];
        Key prefix = prefixes[idxA];
        Key nodeKey = decodePlaceholderBit(prefix);
        unsigned prefixLength = decodePrefixLength(prefix);
        unsigned level = prefixLength / 3;
        Key childPrefix = encodePlaceholderBit(nodeKey, int(prefixLength) + 3);
        int s = levelRange[level + 1], e = levelRange[level + 2];
        int childIdx = int(std::lower_bound(prefixes + s, prefixes + e, childPrefix) - prefixes);
        if (childIdx != e && childPrefix == prefixes[childIdx])
        {
            childOffsets[idxA] = childIdx;
            parents[(childIdx - 1) / 8] = idxA;
        }
    }
}

// nodeFpCenters (focus/source_center.hpp:146-158) with hilbertIBox (hilbert.hpp:259-275)
// and centerAndSize (box.hpp:317-334); centers/sizes are xyz-interleaved
void orc_node_centers(const uint64_t* prefixes, int numNodes, const sphx_box* b, double* centers, double* sizes)
{
    Box box = makeBox(b);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < numNodes; ++i)
    {
        Key prefix = prefixes[i];
        Key startKey = decodePlaceholderBit(prefix);
        unsigned level = decodePrefixLength(prefix) / 3;
        unsigned cubeLength = (1u << kMaxLevel) >> level;
        unsigned mask = ~(cubeLength - 1);
        unsigned ix, iy, iz;
        decodeHilbert(startKey, ix, iy, iz);
        int imin[3] = {int(ix & mask), int(iy & mask), int(iz & mask)};
        for (int d = 0; d < 3; ++d)
        {
            int imax = imin[d] + int(cubeLength);
            const double uL = 1.0 / double(1u << kMaxLevel);
            double halfUnit = 0.5 * uL * box.len[d];
            centers[3 * i + d] = box.lo[d] + double(imax + imin[d]) * halfUnit;
            sizes[3 * i + d]   = double(imax - imin[d]) * halfUnit;
        }
    }
}

// --------------------------------------------------------------------------- neighbors
// findNeighbors of one particle (findneighbors.hpp:77-147) with singleTraversal
// (traversal/traversal.hpp:26-69) and minDistance (traversal/boxoverlap.hpp:230-250)
static unsigned findNeighborsOne(uint32_t i, const double* x, const double* y, const double* z, const double* h,
                                 const int* childOffsets, const int* parents, const int* internalToLeaf,
                                 const uint32_t* layout, const double* centers, const double* sizes,
                                 const Box& box, unsigned ngmax, uint32_t* nbrs)
{
    double xi = x[i], yi = y[i], zi = z[i], hi = h[i];
    double radiusSq = 4.0 * hi * hi;
    double cellRadiusSq = radiusSq * 1.0f * 1.0f; // searchExtFactor = 1.0
    unsigned nn = 0;
    bool anyPbc = box.bc[0] == SPHX_BC_PERIODIC || box.bc[1] == SPHX_BC_PERIODIC || box.bc[2] == SPHX_BC_PERIODIC;
    double p[3] = {xi, yi, zi};
    bool inside = true; // insideBox (boxoverlap.hpp:218-227)
    for (int d = 0; d < 3; ++d)
    {
        double lo = p[d] - 2.0 * hi, hi2 = p[d] + 2.0 * hi;
        inside = inside && lo >= box.lo[d] && hi2 <= box.hi[d];
    }
    bool usePbc = anyPbc && !inside;

    auto overlaps = [&](int idx)
    {
        double dist[3];
        for (int d = 0; d < 3; ++d)
        {
            double dX = centers[3 * idx + d] - p[d];
            if (usePbc) // applyPbc (box.hpp:177-189)
                dX -= double(box.bc[d] == SPHX_BC_PERIODIC) * box.len[d] * std::rint(dX * box.ilen[d]);
            dX = std::fabs(dX);
            dX -= sizes[3 * idx + d];
            dX += std::fabs(dX);
            dX *= 0.5;
            dist[d] = dX;
        }
        return dist[0] * dist[0] + dist[1] * dist[1] + dist[2] * dist[2] < cellRadiusSq;
    };
    auto search = [&](int idx)
    {
        int leafIdx = internalToLeaf[idx];
        for (uint32_t j = layout[leafIdx]; j < layout[leafIdx + 1]; ++j)
        {
            if (j == i) continue;
            double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
            if (usePbc) // distanceSq<true> (findneighbors.hpp:33-48)
            {
                dx -= double(box.bc[0] == SPHX_BC_PERIODIC) * box.len[0] * std::rint(dx * box.ilen[0]);
                dy -= double(box.bc[1] == SPHX_BC_PERIODIC) * box.len[1] * std::rint(dy * box.ilen[1]);
                dz -= double(box.bc[2] == SPHX_BC_PERIODIC) * box.len[2] * std::rint(dz * box.ilen[2]);
            }
            if (dx * dx + dy * dy + dz * dz < radiusSq)
            {
                if (nn < ngmax && nbrs) nbrs[nn] = j;
                nn++;
            }
        }
    };
    // singleTraversal
    if (!overlaps(0)) return 0;
    if (childOffsets[0] == 0)
    {
        search(0);
        return nn;
    }
    int node = childOffsets[0];
    bool backtrack = false;
    while (node != 0)
    {
        bool isLeaf = childOffsets[node] == 0;
        bool descend = !backtrack && overlaps(node);
        if (isLeaf && descend) search(node);
        int siblingIdx = (node - 1) % 8;
        if (!isLeaf && descend)
        {
            node = childOffsets[node];
            backtrack = false;
        }
        else if (siblingIdx < 7)
        {
            node++;
            backtrack = false;
        }
        else
        {
            node = parents[(node - 1) / 8];
            backtrack = true;
        }
    }
    return nn;
}

// findNeighbors batch (findneighbors.hpp:149-170); nbrs may be NULL (counts only)
void orc_find_neighbors(const double* x, const double* y, const double* z, const double* h, uint32_t first,
                        uint32_t last, const sphx_box* b, const int* childOffsets, const int* parents,
                        const int* internalToLeaf, const uint32_t* layout, const double* centers,
                        const double* sizes, unsigned ngmax, uint32_t* nbrs, unsigned* nc)
{
    Box box = makeBox(b);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(last - first); ++i)
    {
        uint32_t id = uint32_t(i) + first;
        nc[i] = findNeighborsOne(id, x, y, z, h, childOffsets, parents, internalToLeaf, layout, centers, sizes, box,
                                 ngmax, nbrs ? nbrs + size_t(i) * ngmax : nullptr);
    }
}

// updateH (sph/kernels.hpp:27-33)
double orc_update_h(unsigned ng0, unsigned nc, double h)
{
    const double c0 = 1023.0, ex = 1.0 / 10.0;
    return h * 0.5 * std::pow(1.0 + c0 * ng0 / double(nc), ex);
}

// findNeighborsSph (sph/include/sph/find_neighbors.hpp:10-41); h is updated in place
void orc_find_neighbors_sph(const double* x, const double* y, const double* z, double* h, uint32_t first,
                            uint32_t last, const sphx_box* b, const int* childOffsets, const int* parents,
                            const int* internalToLeaf, const uint32_t* layout, const double* centers,
                            const double* sizes, unsigned ng0, unsigned ngmax, uint32_t* nbrs, unsigned* nc)
{
    Box box = makeBox(b);
    unsigned ngmin = ng0 / 4;
    const int maxIteration = 10;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(last - first); ++i)
    {
        uint32_t id = uint32_t(i) + first;
        uint32_t* nb = nbrs + size_t(i) * ngmax;
        unsigned ncSph = 1 + findNeighborsOne(id, x, y, z, h, childOffsets, parents, internalToLeaf, layout, centers,
                                              sizes, box, ngmax, nb);
        int iteration = 0;
        while ((ngmin > ncSph || (ncSph - 1) > ngmax) && iteration++ < maxIteration)
        {
            h[id] = orc_update_h(ng0, ncSph, h[id]);
            ncSph = 1 + findNeighborsOne(id, x, y, z, h, childOffsets, parents, internalToLeaf, layout, centers,
                                         sizes, box, ngmax, nb);
        }
        nc[i] = ncSph;
        if (iteration == maxIteration && (ngmin > ncSph || (ncSph - 1) > ngmax))
        {
            h[i]  = 0.0; // the reference indexes with the local i here (find_neighbors.hpp:37)
            nc[i] = 1;
        }
    }
}

// --------------------------------------------------------------------------- SPH loops (T = double)
// computeDensityImpl -> computeXMassImpl -> xmassJLoop (hydro_std/density.hpp:41-53,
// hydro_ve/xmass.hpp:39-67, hydro_ve/xmass_kern.hpp:50-79); rho = m / xm
void orc_density(uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax,
                 const double* x, const double* y, const double* z, const double* h, const double* m,
                 const double* wh, double K, const sphx_box* b, double* rho)
{
    Box box = makeBox(b);
#pragma omp parallel for schedule(static)
    for (int64_t ii = int64_t(first); ii < int64_t(last); ++ii)
    {
        size_t i = size_t(ii), ni = i - first;
        unsigned cnt = std::min(nc[i] - 1, ngmax);
        const uint32_t* nb = nbrs + ngmax * ni;
        double xi = x[i], yi = y[i], zi = z[i], hi = h[i], mi = m[i];
        double hInv = 1.0 / hi;
        double h3Inv = hInv * hInv * hInv;
        double rho0i = mi;
        for (unsigned pj = 0; pj < cnt; ++pj)
        {
            uint32_t j = nb[pj];
            double xx = xi - x[j], yy = yi - y[j], zz = zi - z[j]; // distancePBC (box.hpp:240-250)
            applyPBC(box, 2.0 * hi, xx, yy, zz);
            double dist = std::sqrt(xx * xx + yy * yy + zz * zz);
            double vloc = dist * hInv;
            double w = lookup(wh, vloc);
            rho0i += w * m[j];
        }
        double xm = mi / (rho0i * K * h3Inv);
        rho[i] = m[i] / xm;
    }
}

// idealGasEOS_u over a range (hydro_std/eos.hpp:52-78, sph/eos.hpp:34-44)
void orc_eos(uint32_t first, uint32_t last, const double* u, const double* rho, double gamma, double* p, double* c)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = first; i < int64_t(last); ++i)
    {
        double tmp = u[i] * (gamma - 1.0);
        p[i] = rho[i] * tmp;
        c[i] = std::sqrt(gamma * tmp);
    }
}

// IADJLoopSTD (hydro_std/iad_kern.hpp:12-78) over [first,last) with capped counts (iad.hpp:64-70)
static void iadOne(size_t i, unsigned cnt, const uint32_t* nb, const double* x, const double* y, const double* z,
                   const double* h, const double* m, const double* rho, const double* wh, double K, const Box& box,
                   double* c11, double* c12, double* c13, double* c22, double* c23, double* c33)
{
    double t11 = 0, t12 = 0, t13 = 0, t22 = 0, t23 = 0, t33 = 0;
    double xi = x[i], yi = y[i], zi = z[i], hi = h[i], hiInv = 1.0 / hi;
    for (unsigned pj = 0; pj < cnt; ++pj)
    {
        uint32_t j = nb[pj];
        double rx = xi - x[j], ry = yi - y[j], rz = zi - z[j];
        applyPBC(box, 2.0 * hi, rx, ry, rz);
        double dist = std::sqrt(rx * rx + ry * ry + rz * rz);
        double vloc = dist * hiInv;
        double w = lookup(wh, vloc);
        double mj_roj_w = m[j] / rho[j] * w;
        t11 += rx * rx * mj_roj_w;
        t12 += rx * ry * mj_roj_w;
        t13 += rx * rz * mj_roj_w;
        t22 += ry * ry * mj_roj_w;
        t23 += ry * rz * mj_roj_w;
        t33 += rz * rz * mj_roj_w;
    }
    auto getExp = [](double v) { return v == 0.0 ? 0 : std::ilogb(v); };
    int expSum = getExp(t11) + getExp(t12) + getExp(t13) + getExp(t22) + getExp(t23) + getExp(t33);
    double norm = std::ldexp(1.0, -expSum / 6);
    t11 *= norm, t12 *= norm, t13 *= norm, t22 *= norm, t23 *= norm, t33 *= norm;
    double det = t11 * t22 * t33 + 2.0 * t12 * t23 * t13 - t11 * t23 * t23 - t22 * t13 * t13 - t33 * t12 * t12;
    double factor = norm * (hi * hi * hi) / (det * K);
    if (std::isnan(factor) && cnt == 0) factor = 0.0;
    c11[i] = (t22 * t33 - t23 * t23) * factor;
    c12[i] = (t13 * t23 - t33 * t12) * factor;
    c13[i] = (t12 * t23 - t22 * t13) * factor;
    c22[i] = (t11 * t33 - t13 * t13) * factor;
    c23[i] = (t13 * t12 - t11 * t23) * factor;
    c33[i] = (t11 * t22 - t12 * t12) * factor;
}

void orc_iad(uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax, const double* x,
             const double* y, const double* z, const double* h, const double* m, const double* rho, const double* wh,
             double K, const sphx_box* b, double* c11, double* c12, double* c13, double* c22, double* c23,
             double* c33)
{
    Box box = makeBox(b);
#pragma omp parallel for schedule(static)
    for (int64_t ii = int64_t(first); ii < int64_t(last); ++ii)
    {
        size_t i = size_t(ii), ni = i - first;
        unsigned cnt = std::min(nc[i] - 1, ngmax);
        iadOne(i, cnt, nbrs + ngmax * ni, x, y, z, h, m, rho, wh, K, box, c11, c12, c13, c22, c23, c33);
    }
}

// single-particle IAD with an explicit neighbor list (the std.cpp known-answer form)
void orc_iad_one(uint32_t i, const uint32_t* nb, unsigned cnt, const double* x, const double* y, const double* z,
                 const double* h, const double* m, const double* rho, const double* wh, double K, const sphx_box* b,
                 double* out6)
{
    Box box = makeBox(b);
    std::vector<double> t(6 * (size_t(i) + 1));
    size_t s = size_t(i) + 1;
    iadOne(i, cnt, nb, x, y, z, h, m, rho, wh, K, box, &t[0], &t[s], &t[2 * s], &t[3 * s], &t[4 * s], &t[5 * s]);
    for (int k = 0; k < 6; ++k)
        out6[k] = t[k * s + i];
}

// momentumAndEnergyJLoop (hydro_std/momentum_energy_kern.hpp:12-134)
static void meOne(size_t i, unsigned cnt, const uint32_t* nb, const double* x, const double* y, const double* z,
                  const double* vx, const double* vy, const double* vz, const double* h, const double* m,
                  const double* rho, const double* p, const double* c, const double* c11, const double* c12,
                  const double* c13, const double* c22, const double* c23, const double* c33, const double* wh,
                  double K, const Box& box, double* ax, double* ay, double* az, double* du, double* maxvsignal)
{
    const double gradh_i = 1.0, gradh_j = 1.0;
    double xi = x[i], yi = y[i], zi = z[i], vxi = vx[i], vyi = vy[i], vzi = vz[i];
    double hi = h[i], roi = rho[i], pri = p[i], ci = c[i];
    double mi_roi = m[i] / rho[i];
    double hiInv = 1.0 / hi, hiInv3 = hiInv * hiInv * hiInv;
    double maxv = 0.0, mx = 0.0, my = 0.0, mz = 0.0, energy = 0.0;
    double c11i = c11[i], c12i = c12[i], c13i = c13[i], c22i = c22[i], c23i = c23[i], c33i = c33[i];
    for (unsigned pj = 0; pj < cnt; ++pj)
    {
        uint32_t j = nb[pj];
        double rx = xi - x[j], ry = yi - y[j], rz = zi - z[j];
        applyPBC(box, 2.0 * hi, rx, ry, rz);
        double r2 = rx * rx + ry * ry + rz * rz;
        double dist = std::sqrt(r2);
        double vx_ij = vxi - vx[j], vy_ij = vyi - vy[j], vz_ij = vzi - vz[j];
        double hj = h[j], hjInv = 1.0 / hj;
        double v1 = dist * hiInv, v2 = dist * hjInv;
        double rv = rx * vx_ij + ry * vy_ij + rz * vz_ij;
        double hjInv3 = hjInv * hjInv * hjInv;
        double Wi = hiInv3 * lookup(wh, v1);
        double Wj = hjInv3 * lookup(wh, v2);
        double tA1i = c11i * rx + c12i * ry + c13i * rz;
        double tA2i = c12i * rx + c22i * ry + c23i * rz;
        double tA3i = c13i * rx + c23i * ry + c33i * rz;
        double tA1j = c11[j] * rx + c12[j] * ry + c13[j] * rz;
        double tA2j = c12[j] * rx + c22[j] * ry + c23[j] * rz;
        double tA3j = c13[j] * rx + c23[j] * ry + c33[j] * rz;
        double roj = rho[j], cj = c[j];
        double wij = rv / dist;
        // artificial_viscosity (kernels.hpp:71-85) with alpha_i = alpha_j = 1, beta = 2
        double visc = 0.0;
        if (wij < 0.0)
        {
            double vij_signal = (1.0 + 1.0) * 0.25 * (ci + cj) - 2.0 * wij;
            visc = -vij_signal * wij;
        }
        double viscosity_ij = 0.5 * visc;
        double vijsignal = ci + cj - 3.0 * wij;
        maxv = (vijsignal > maxv) ? vijsignal : maxv;
        double mj = m[j];
        double mj_roj_Wj = mj / roj * Wj;
        double mj_pro_i = mj * pri / (gradh_i * roi * roi);
        {
            double a = Wi * (mj_pro_i + viscosity_ij * mi_roi);
            double bb = mj_roj_Wj * (p[j] / (roj * gradh_j) + viscosity_ij);
            mx += a * tA1i + bb * tA1j;
            my += a * tA2i + bb * tA2j;
            mz += a * tA3i + bb * tA3j;
        }
        {
            double a = Wi * (2.0 * mj_pro_i + viscosity_ij * mi_roi);
            double bb = viscosity_ij * mj_roj_Wj;
            energy += vx_ij * (a * tA1i + bb * tA1j) + vy_ij * (a * tA2i + bb * tA2j) + vz_ij * (a * tA3i + bb * tA3j);
        }
    }
    du[i] = -K * 0.5 * energy;
    ax[i] = K * mx;
    ay[i] = K * my;
    az[i] = K * mz;
    *maxvsignal = maxv;
}

// computeMomentumEnergyStdImpl (hydro_std/momentum_energy.hpp:41-94); returns minDtCourant
double orc_momentum_energy(uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax,
                           const double* x, const double* y, const double* z, const double* vx, const double* vy,
                           const double* vz, const double* h, const double* m, const double* rho, const double* p,
                           const double* c, const double* c11, const double* c12, const double* c13,
                           const double* c22, const double* c23, const double* c33, const double* wh, double K,
                           float Kcour, const sphx_box* b, double* ax, double* ay, double* az, double* du)
{
    Box box = makeBox(b);
    double minDt = INFINITY;
#pragma omp parallel for schedule(static) reduction(min : minDt)
    for (int64_t ii = int64_t(first); ii < int64_t(last); ++ii)
    {
        size_t i = size_t(ii), ni = i - first;
        unsigned cnt = std::min(nc[i] - 1, ngmax);
        double maxv = 0;
        meOne(i, cnt, nbrs + ngmax * ni, x, y, z, vx, vy, vz, h, m, rho, p, c, c11, c12, c13, c22, c23, c33, wh, K,
              box, ax, ay, az, du, &maxv);
        // tsKCourant (kernels.hpp:10-17)
        double v = maxv > 0.0 ? maxv : c[i];
        double dt = h[i] > 0.0 ? double(Kcour * h[i] / v) : INFINITY;
        minDt = std::min(minDt, dt);
    }
    return minDt;
}

// single-particle momentum-energy with an explicit list (std.cpp known-answer form);
// out5 = {grad_Px, grad_Py, grad_Pz, du, maxvsignal}
void orc_momentum_energy_one(uint32_t i, const uint32_t* nb, unsigned cnt, const double* x, const double* y,
                             const double* z, const double* vx, const double* vy, const double* vz, const double* h,
                             const double* m, const double* rho, const double* p, const double* c,
                             const double* c11, const double* c12, const double* c13, const double* c22,
                             const double* c23, const double* c33, const double* wh, double K, const sphx_box* b,
                             double* out5)
{
    Box box = makeBox(b);
    size_t s = size_t(i) + 1;
    std::vector<double> t(4 * s);
    meOne(i, cnt, nb, x, y, z, vx, vy, vz, h, m, rho, p, c, c11, c12, c13, c22, c23, c33, wh, K, box, &t[0], &t[s],
          &t[2 * s], &t[3 * s], &out5[4]);
    for (int k = 0; k < 4; ++k)
        out5[k] = t[k * s + i];
}

// single-particle xmass with an explicit list (ve.cpp XMass known-answer form)
double orc_xmass_one(uint32_t i, const uint32_t* nb, unsigned cnt, const double* x, const double* y, const double* z,
                     const double* h, const double* m, const double* wh, double K, const sphx_box* b)
{
    Box box = makeBox(b);
    double xi = x[i], yi = y[i], zi = z[i], hi = h[i], mi = m[i];
    double hInv = 1.0 / hi, h3Inv = hInv * hInv * hInv;
    double rho0i = mi;
    for (unsigned pj = 0; pj < cnt; ++pj)
    {
        uint32_t j = nb[pj];
        double xx = xi - x[j], yy = yi - y[j], zz = zi - z[j];
        applyPBC(box, 2.0 * hi, xx, yy, zz);
        double dist = std::sqrt(xx * xx + yy * yy + zz * zz);
        rho0i += lookup(wh, dist * hInv) * m[j];
    }
    return mi / (rho0i * K * h3Inv);
}

// computeGroupSplits semantics (domain/include/cstone/traversal/groups_gpu.cuh:37-77,
// 91-114, 148-216; groups_gpu.cu:108-137) with groupSize 64: writes the group
// boundaries (numGroups + 1 entries, cap checked) and returns numGroups (-1 if cap too small)
int orc_compute_groups(uint32_t first, uint32_t last, const double* x, const double* y, const double* z,
                       const double* h, const uint64_t* leaves, int numLeaves, const uint32_t* layout,
                       const sphx_box* b, float tolFactor, uint32_t* bounds, int cap)
{
    Box box = makeBox(b);
    double minExt = std::min(std::min(box.len[0], box.len[1]), box.len[2]);
    uint32_t numFixed = (last - first + 63) / 64;
    std::vector<uint32_t> out;
    for (uint32_t g = 0; g < numFixed; ++g)
    {
        double pos[64][4];
        double nodeVolume = 1.0;
        for (int lane = 0; lane < 64; ++lane)
        {
            uint32_t bi = std::min(first + 64 * g + lane, last - 1);
            int leaf = int(std::upper_bound(layout, layout + numLeaves, bi) - layout) - 1;
            unsigned level = treeLevel(leaves[leaf + 1] - leaves[leaf]);
            double s = std::ldexp(0.5, -int(level)); // centerAndSize half-size of the leaf in the unit box
            nodeVolume = std::min(nodeVolume, 8 * s * s * s);
            pos[lane][0] = x[bi] * box.ilen[0];
            pos[lane][1] = y[bi] * box.ilen[1];
            pos[lane][2] = z[bi] * box.ilen[2];
            pos[lane][3] = h ? 2.0 * h[bi] / minExt : 1.0;
        }
        double distCrit = std::cbrt(nodeVolume) * double(tolFactor);
        out.push_back(first + 64 * g);
        for (int lane = 0; lane < 63; ++lane)
        {
            double dx = pos[lane + 1][0] - pos[lane][0], dy = pos[lane + 1][1] - pos[lane][1],
                   dz = pos[lane + 1][2] - pos[lane][2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > std::min(distCrit * distCrit, pos[lane][3] * pos[lane][3])) out.push_back(first + 64 * g + lane + 1);
        }
    }
    out.push_back(last);
    int numGroups = int(out.size()) - 1;
    if (numFixed == 0) numGroups = 0;
    if (cap < numGroups + 1) return -1;
    std::copy(out.begin(), out.end(), bounds);
    return numGroups;
}

// makeGlobalBox (domain/include/cstone/sfc/box_mpi.hpp:33-107) on one rank: periodic and
// fixed dimensions keep the previous limits, open ones take min / max of the coordinates
void orc_make_global_box(const double* x, const double* y, const double* z, uint64_t n, const sphx_box* prev,
                         sphx_box* out)
{
    *out = *prev;
    if (n == 0) return;
    const double* c[3] = {x, y, z};
    int32_t bc[3] = {prev->bx, prev->by, prev->bz};
    double* lo[3] = {&out->xmin, &out->ymin, &out->zmin};
    double* hi[3] = {&out->xmax, &out->ymax, &out->zmax};
    for (int d = 0; d < 3; ++d)
    {
        if (bc[d] == SPHX_BC_PERIODIC || bc[d] == SPHX_BC_FIXED) continue;
        double mn = INFINITY, mx = -INFINITY;
        for (uint64_t i = 0; i < n; ++i)
        {
            mn = std::min(mn, c[d][i]);
            mx = std::max(mx, c[d][i]);
        }
        *lo[d] = mn, *hi[d] = mx;
    }
}

// limitBoxShrinking (domain/include/cstone/sfc/box.hpp:397-414): each side of the fitted box
// moves inward by at most shrinkLimit of the previous box's length; the boundary types are
// the previous box's.  Domain::sync applies it from the second call on
// (domain/assignment.hpp:104-106).
void orc_limit_box_shrinking(const sphx_box* fit, const sphx_box* prev, double shrinkLimit, sphx_box* out)
{
    const double lx = prev->xmax - prev->xmin, ly = prev->ymax - prev->ymin, lz = prev->zmax - prev->zmin;
    *out      = *prev;
    out->xmin = std::min(fit->xmin, prev->xmin + shrinkLimit * lx);
    out->xmax = std::max(fit->xmax, prev->xmax - shrinkLimit * lx);
    out->ymin = std::min(fit->ymin, prev->ymin + shrinkLimit * ly);
    out->ymax = std::max(fit->ymax, prev->ymax - shrinkLimit * ly);
    out->zmin = std::min(fit->zmin, prev->zmin + shrinkLimit * lz);
    out->zmax = std::max(fit->zmax, prev->zmax - shrinkLimit * lz);
}


// --------------------------------------------------------------------------- integration
// accelerationTimestep, CPU branch (sph/include/sph/ts_global.hpp:48-71) without the
// final etaAcc * pow(., 0.25): min over [first, last) of h_i^2 / |a_i|^2; a NaN term
// leaves the minimum unchanged (std::min(m, NaN) returns m); INFINITY if empty
double orc_min_h2_a2(uint32_t first, uint32_t last, const double* h, const double* ax, const double* ay,
                     const double* az)
{
    double m = INFINITY;
    for (uint32_t i = first; i < last; ++i)
    {
        // norm2 = dot(a, a), a right fold a0 a0 + (a1 a1 + a2 a2) (util/array.hpp:236-240,311-320)
        double a2 = ax[i] * ax[i] + (ay[i] * ay[i] + az[i] * az[i]);
        m         = std::min(m, h[i] * h[i] / a2);
    }
    return m;
}

// computeTimestep (ts_global.hpp:100-130) on one rank (the MPI_Allreduce of one value
// is the identity); ts = {minDt, minDt_m1, ttot} updated in place; returns minDtAcc
double orc_compute_timestep(uint32_t first, uint32_t last, const double* h, const double* ax, const double* ay,
                            const double* az, double g, double etaAcc, double minDtCourant, double minDtRho,
                            double* ts)
{
    const double maxDtIncrease = 1.1; // particles_data.hpp:138
    double minDtAcc = INFINITY;
    if (g != 0.0)
    {
        minDtAcc = last > first ? etaAcc * std::pow(orc_min_h2_a2(first, last, h, ax, ay, az), 0.25) : INFINITY;
    }
    double minDt = std::min({minDtAcc, minDtCourant, minDtRho, maxDtIncrease * ts[0]});
    ts[2] += minDt;
    ts[1] = ts[0];
    ts[0] = minDt;
    return minDtAcc;
}

// fbcAdjustFactors (sph/include/sph/positions.hpp:47-74)
static void fbcFactors(const double X[3], const Box& box, double hi, double f[3])
{
    const double width = 2.0, invWidth = 1.0 / width, offset = 1.0;
    for (int j = 0; j < 3; ++j)
    {
        f[j] = 1.0;
        double md = std::min(std::abs(box.hi[j] - X[j]), std::abs(box.lo[j] - X[j])) * (1.0 / hi);
        if (box.bc[j] == SPHX_BC_FIXED && md < 2 * width + offset)
        {
            f[j] = 0.5 * (std::tanh(4 * (md - offset) * invWidth - 4) + 1);
        }
    }
}

// energyUpdate (positions.hpp:77-84)
static double energyUpdate(double u_old, double dt, double dt_m1, double du, double du_m1)
{
    double u_new = u_old + du * dt + 0.5 * (du - du_m1) / dt_m1 * std::abs(dt) * dt;
    if (u_new < 0.) { u_new = u_old * std::exp(u_new * dt / u_old); }
    return u_new;
}

// updatePositionsHost (positions.hpp:113-137) with positionUpdate (:100-111) and putInBox
// (cstone/sfc/box.hpp:193-214), then updateIntEnergyHost (:162-179)
void orc_compute_positions(uint32_t first, uint32_t last, double* x, double* y, double* z, double* x_m1,
                           double* y_m1, double* z_m1, double* vx, double* vy, double* vz, const double* ax,
                           const double* ay, const double* az, const double* h, double* u, const double* du,
                           double* du_m1, double dt, double dt_m1, const sphx_box* b)
{
    Box box = makeBox(b);
    bool anyFBC = box.bc[0] == SPHX_BC_FIXED || box.bc[1] == SPHX_BC_FIXED || box.bc[2] == SPHX_BC_FIXED;
#pragma omp parallel for schedule(static)
    for (uint32_t i = first; i < last; ++i)
    {
        double X[3] = {x[i], y[i], z[i]};
        double f[3] = {1.0, 1.0, 1.0};
        if (anyFBC) fbcFactors(X, box, h[i], f);
        double A[3]   = {ax[i] * f[0], ay[i] * f[1], az[i] * f[2]};
        double dXn[3] = {x_m1[i] * f[0], y_m1[i] * f[1], z_m1[i] * f[2]};
        double V[3], dX[3];
        for (int k = 0; k < 3; ++k)
        {
            double Vnmhalf = dXn[k] * (1.0 / dt_m1);
            double Vn      = Vnmhalf + 0.5 * dt_m1 * A[k];
            V[k]           = Vn + A[k] * dt;
            dX[k]          = (Vn + 0.5 * A[k] * std::abs(dt)) * dt;
            X[k]           = X[k] + dX[k];
            if (box.bc[k] == SPHX_BC_PERIODIC && X[k] > box.hi[k]) { X[k] -= box.len[k]; }
            else if (box.bc[k] == SPHX_BC_PERIODIC && X[k] < box.lo[k]) { X[k] += box.len[k]; }
        }
        x[i] = X[0], y[i] = X[1], z[i] = X[2];
        x_m1[i] = dX[0], y_m1[i] = dX[1], z_m1[i] = dX[2];
        vx[i] = V[0], vy[i] = V[1], vz[i] = V[2];
        // updateIntEnergyHost runs after every position is updated: its factor is that
        // of the new position (the reference GPU kernel uses the old one)
        double fe = 1.0;
        if (anyFBC)
        {
            fbcFactors(X, box, h[i], f);
            fe = std::min({f[0], f[1], f[2]});
        }
        u[i]      = energyUpdate(u[i], dt * fe, dt_m1 * fe, du[i], du_m1[i]);
        du_m1[i]  = du[i];
    }
}

// updateSmoothingLengthCpu (sph/include/sph/update_h.hpp:13-32); keys may be NULL;
// returns keysRemoved
int orc_update_smoothing_length(uint32_t first, uint32_t last, unsigned ng0, const unsigned* nc, double* h,
                                uint64_t* keys)
{
    int removed = 0;
    for (uint32_t i = first; i < last; ++i)
    {
        if (nc[i] <= 1)
        {
            if (keys) keys[i] = kRemoveKey;
            removed = 1;
        }
        h[i] = orc_update_h(ng0, nc[i], h[i]);
    }
    return removed;
}

// updateH<float> (sph/kernels.hpp:27-33): the reference app's h is HydroType = float
// (sph/types.hpp:39-46), so its update runs in float arithmetic with powf
float orc_update_h_f32(unsigned ng0, unsigned nc, float h)
{
    const float c0 = 1023.0f, ex = 1.0 / 10.0;
    return h * 0.5f * std::pow(1.0f + c0 * ng0 / float(nc), ex);
}

// updateSmoothingLengthCpu<float> (sph/include/sph/update_h.hpp:13-32)
int orc_update_smoothing_length_f32(uint32_t first, uint32_t last, unsigned ng0, const unsigned* nc, float* h,
                                    uint64_t* keys)
{
    int removed = 0;
    for (uint32_t i = first; i < last; ++i)
    {
        if (nc[i] <= 1)
        {
            if (keys) keys[i] = kRemoveKey;
            removed = 1;
        }
        h[i] = orc_update_h_f32(ng0, nc[i], h[i]);
    }
    return removed;
}


// --------------------------------------------------------------------------- VE hydro (T = double)
// The fields the VE loops read; unused pointers may be NULL.  Layout mirrored by the
// ctypes structure in oracle/__init__.py.
struct OrcVe
{
    const double *x, *y, *z, *vx, *vy, *vz, *h, *m, *c, *xm, *kx, *gradh, *prho, *divv, *alpha;
    const double *c11, *c12, *c13, *c22, *c23, *c33;
    const double *dV11, *dV12, *dV13, *dV22, *dV23, *dV33;
    const double *wh, *whd;
    double        K;
};

// xmassJLoop (sph/include/sph/hydro_ve/xmass_kern.hpp:50-79) over a range
// (computeXMassImpl, hydro_ve/xmass.hpp:39-67); distancePBC = applyPBC with 2 h
void orc_ve_xmass(uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax,
                  const OrcVe* f, const sphx_box* b, double* xm)
{
    Box box = makeBox(b);
#pragma omp parallel for schedule(static)
    for (int64_t ii = int64_t(first); ii < int64_t(last); ++ii)
    {
        size_t i = size_t(ii), ni = i - first;
        unsigned cnt = std::min(nc[i] - 1, ngmax);
        const uint32_t* nb = nbrs + size_t(ngmax) * ni;
        double xi = f->x[i], yi = f->y[i], zi = f->z[i], hi = f->h[i], mi = f->m[i];
        double hInv = 1.0 / hi, h3Inv = hInv * hInv * hInv;
        double rho0i = mi;
        for (unsigned pj = 0; pj < cnt; ++pj)
        {
            uint32_t j = nb[pj];
            double xx = xi - f->x[j], yy = yi - f->y[j], zz = zi - f->z[j];
            applyPBC(box, 2.0 * hi, xx, yy, zz);
            double dist = std::sqrt(xx * xx + yy * yy + zz * zz);
            rho0i += lookup(f->wh, dist * hInv) * f->m[j];
        }
        xm[i] = mi / (rho0i * f->K * h3Inv); // veDefinition (xmass_kern.hpp:45-48)
    }
}

// veDefGradhJLoop (hydro_ve/ve_def_gradh_kern.hpp:43-89) over a range
// (computeVeDefGradhImpl, hydro_ve/ve_def_gradh.hpp:40-79)
void orc_ve_def_gradh(uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax,
                      const OrcVe* f, const sphx_box* b, double* kx, double* gradh)
{
    Box box = makeBox(b);
    const double K = f->K;
#pragma omp parallel for schedule(static)
    for (int64_t ii = int64_t(first); ii < int64_t(last); ++ii)
    {
        size_t i = size_t(ii), ni = i - first;
        unsigned cnt = std::min(nc[i] - 1, ngmax);
        const uint32_t* nb = nbrs + size_t(ngmax) * ni;
        double xi = f->x[i], yi = f->y[i], zi = f->z[i], hi = f->h[i], mi = f->m[i], xmassi = f->xm[i];
        double hInv = 1.0 / hi, h3Inv = hInv * hInv * hInv;
        double kxi = xmassi, whomegai = -3.0 * xmassi, wrho0i = -3.0 * mi;
        for (unsigned pj = 0; pj < cnt; ++pj)
        {
            uint32_t j = nb[pj];
            double xx = xi - f->x[j], yy = yi - f->y[j], zz = zi - f->z[j];
            applyPBC(box, 2.0 * hi, xx, yy, zz);
            double dist  = std::sqrt(xx * xx + yy * yy + zz * zz);
            double vloc  = dist * hInv;
            double w     = lookup(f->wh, vloc);
            double dw    = lookup(f->whd, vloc);
            double dterh = -(3.0 * w + vloc * dw);
            double xmj   = f->xm[j];
            kxi += w * xmj;
            whomegai += dterh * xmj;
            wrho0i += dterh * f->m[j];
        }
        kxi *= K * h3Inv;
        whomegai *= K * h3Inv * hInv;
        wrho0i *= K * h3Inv * hInv;
        whomegai      = whomegai * mi / xmassi + (kxi - K * xmassi * h3Inv) * wrho0i;
        double rhoi   = kxi * mi / xmassi;
        double dhdrho = -hi / (rhoi * 3.0);
        kx[i]         = kxi;
        gradh[i]      = 1.0 - dhdrho * whomegai;
    }
}

// computeIdealGasEOS_Impl, u branch (hydro_ve/eos.hpp:52-91) with idealGasEOS_u
// (sph/eos.hpp:34-44); rho and p may be NULL
void orc_ve_eos(uint32_t first, uint32_t last, const double* u, const double* m, const double* kx,
                const double* xm, const double* gradh, double gamma, double* prho, double* c, double* rho, double* p)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = first; i < int64_t(last); ++i)
    {
        double r   = kx[i] * m[i] / xm[i];
        double tmp = u[i] * (gamma - 1.0);
        double pi  = r * tmp;
        prho[i]    = pi / (kx[i] * m[i] * m[i] * gradh[i]);
        c[i]       = std::sqrt(gamma * tmp);
        if (rho) rho[i] = r;
        if (p) p[i] = pi;
    }
}

// computeIadDivvCurlvImpl (hydro_ve/iad_divv_curlv.hpp:40-91): IADJLoop
// (hydro_ve/iad_kern.hpp:43-110, the std loop with volume xm/kx) then divV_curlVJLoop
// (hydro_ve/divv_curlv_kern.hpp:43-122) with the fresh c_i; curlv and the dV may be NULL.
// givenIad != 0 skips the IAD part and reads c11..c33 from f (divV_curlVJLoop on its own,
// the form of the reference's known-answer test); the outputs c11..c33 are then untouched
void orc_ve_iad_divv_curlv(uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc,
                           unsigned ngmax, const OrcVe* f, const sphx_box* b, int givenIad, double* c11, double* c12,
                           double* c13, double* c22, double* c23, double* c33, double* divv, double* curlv,
                           double* dV11, double* dV12, double* dV13, double* dV22, double* dV23, double* dV33)
{
    if (givenIad)
    {
        c11 = const_cast<double*>(f->c11), c12 = const_cast<double*>(f->c12), c13 = const_cast<double*>(f->c13);
        c22 = const_cast<double*>(f->c22), c23 = const_cast<double*>(f->c23), c33 = const_cast<double*>(f->c33);
    }
    else
        orc_iad(first, last, nbrs, nc, ngmax, f->x, f->y, f->z, f->h, f->xm, f->kx, f->wh, f->K, b, c11, c12, c13,
                c22, c23, c33);
    Box box = makeBox(b);
    const double K = f->K;
#pragma omp parallel for schedule(static)
    for (int64_t ii = int64_t(first); ii < int64_t(last); ++ii)
    {
        size_t i = size_t(ii), ni = i - first;
        unsigned cnt = std::min(nc[i] - 1, ngmax);
        const uint32_t* nb = nbrs + size_t(ngmax) * ni;
        double xi = f->x[i], yi = f->y[i], zi = f->z[i], vxi = f->vx[i], vyi = f->vy[i], vzi = f->vz[i];
        double hi = f->h[i], kxi = f->kx[i];
        double hiInv = 1.0 / hi, hiInv3 = hiInv * hiInv * hiInv;
        double dVx[3] = {0, 0, 0}, dVy[3] = {0, 0, 0}, dVz[3] = {0, 0, 0};
        double a11 = c11[i], a12 = c12[i], a13 = c13[i], a22 = c22[i], a23 = c23[i], a33 = c33[i];
        for (unsigned pj = 0; pj < cnt; ++pj)
        {
            uint32_t j = nb[pj];
            double rx = xi - f->x[j], ry = yi - f->y[j], rz = zi - f->z[j];
            applyPBC(box, 2.0 * hi, rx, ry, rz);
            double r2   = rx * rx + ry * ry + rz * rz;
            double dist = std::sqrt(r2);
            double vxji = f->vx[j] - vxi, vyji = f->vy[j] - vyi, vzji = f->vz[j] - vzi;
            double Wi   = lookup(f->wh, dist * hiInv);
            double tA[3];
            tA[0] = -(a11 * rx + a12 * ry + a13 * rz) * Wi;
            tA[1] = -(a12 * rx + a22 * ry + a23 * rz) * Wi;
            tA[2] = -(a13 * rx + a23 * ry + a33 * rz) * Wi;
            double xmj = f->xm[j];
            double sx = vxji * xmj, sy = vyji * xmj, sz = vzji * xmj;
            for (int k = 0; k < 3; ++k)
            {
                dVx[k] += sx * tA[k];
                dVy[k] += sy * tA[k];
                dVz[k] += sz * tA[k];
            }
        }
        double nk = K * hiInv3 / kxi;
        divv[i]   = nk * (dVx[0] + dVy[1] + dVz[2]);
        if (curlv)
        {
            double cv0 = dVz[1] - dVy[2], cv1 = dVx[2] - dVz[0], cv2 = dVy[0] - dVx[1];
            curlv[i]   = nk * std::sqrt(cv0 * cv0 + (cv1 * cv1 + cv2 * cv2)); // norm2: right fold
        }
        if (dV11)
        {
            dV11[i] = nk * dVx[0];
            dV12[i] = nk * (dVx[1] + dVy[0]);
            dV13[i] = nk * (dVx[2] + dVz[0]);
            dV22[i] = nk * dVy[1];
            dV23[i] = nk * (dVy[2] + dVz[1]);
            dV33[i] = nk * dVz[2];
        }
    }
}

// AVswitchesJLoop (hydro_ve/av_switches_kern.hpp:43-136) over a range
// (computeAVswitchesImpl, hydro_ve/av_switches.hpp:40-78); alpha updated in place
void orc_ve_av_switches(uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax,
                        const OrcVe* f, const sphx_box* b, double dt, double alphamin, double alphamax,
                        double decay_constant, double* alpha)
{
    Box box = makeBox(b);
    const double K = f->K;
#pragma omp parallel for schedule(static)
    for (int64_t ii = int64_t(first); ii < int64_t(last); ++ii)
    {
        size_t i = size_t(ii), ni = i - first;
        unsigned cnt = std::min(nc[i] - 1, ngmax);
        const uint32_t* nb = nbrs + size_t(ngmax) * ni;
        double xi = f->x[i], yi = f->y[i], zi = f->z[i], vxi = f->vx[i], vyi = f->vy[i], vzi = f->vz[i];
        double hi = f->h[i], ci = f->c[i];
        double a11 = f->c11[i], a12 = f->c12[i], a13 = f->c13[i], a22 = f->c22[i], a23 = f->c23[i], a33 = f->c33[i];
        double vsig   = 1.e-40 * ci;
        double hiInv  = 1.0 / hi, hiInv3 = hiInv * hiInv * hiInv;
        double divv_i = f->divv[i];
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (unsigned pj = 0; pj < cnt; ++pj)
        {
            uint32_t j = nb[pj];
            double rx = xi - f->x[j], ry = yi - f->y[j], rz = zi - f->z[j];
            applyPBC(box, 2.0 * hi, rx, ry, rz);
            double r2   = rx * rx + ry * ry + rz * rz;
            double dist = std::sqrt(r2);
            double vxij = vxi - f->vx[j], vyij = vyi - f->vy[j], vzij = vzi - f->vz[j];
            double rv   = rx * vxij + ry * vyij + rz * vzij;
            double vs   = 0.0;
            if (rv < 0.0) { vs = ci + f->c[j] - 3.0 * rv / dist; }
            vsig = (vsig < vs) ? vs : vsig; // stl::max
            double Wi     = K * hiInv3 * lookup(f->wh, dist * hiInv);
            double tA1    = -(a11 * rx + a12 * ry + a13 * rz) * Wi;
            double tA2    = -(a12 * rx + a22 * ry + a23 * rz) * Wi;
            double tA3    = -(a13 * rx + a23 * ry + a33 * rz) * Wi;
            double volj   = f->xm[j] / f->kx[j];
            double factor = volj * (divv_i - f->divv[j]);
            gx += factor * tA1;
            gy += factor * tA2;
            gz += factor * tA3;
        }
        double graddivv = std::sqrt(gx * gx + gy * gy + gz * gz);
        double alphaloc = 0.0;
        if (divv_i < 0.0)
        {
            double a_const = hi * hi * graddivv;
            alphaloc       = alphamax * a_const / (a_const + hi * std::abs(divv_i) + 0.05 * ci);
        }
        double alpha_i = alpha[i];
        if (alphaloc >= alpha_i) { alpha_i = alphaloc; }
        else
        {
            double decay    = hi / (decay_constant * vsig);
            double alphadot = 0.0;
            if (alphaloc >= alphamin) { alphadot = (alphaloc - alpha_i) / decay; }
            else { alphadot = (alphamin - alpha_i) / decay; }
            alpha_i += alphadot * dt;
        }
        alpha[i] = alpha_i;
    }
}

// avRvCorrection (hydro_ve/momentum_energy_kern.hpp:43-63) with symv (sph/kernels.hpp)
static double avRvCorrection(const double R[3], double eta_ab, double eta_crit, const double gi[6], const double gj[6])
{
    auto dotSymv = [&](const double g[6]) {
        double s0 = g[0] * R[0] + g[1] * R[1] + g[2] * R[2];
        double s1 = g[3] * R[1] + g[4] * R[2];
        double s2 = g[5] * R[2];
        return R[0] * s0 + (R[1] * s1 + R[2] * s2); // dot: right fold (util/array.hpp:236-240)
    };
    double dmy1 = dotSymv(gi), dmy2 = dotSymv(gj), dmy3 = 1.0;
    if (eta_ab < eta_crit)
    {
        double etaDiff = 5.0 * (eta_ab - eta_crit);
        dmy3           = std::exp(-etaDiff * etaDiff);
    }
    double A_ab   = (dmy2 != 0.0) ? dmy1 / dmy2 : 0.0;
    double A_abp1 = 1.0 + A_ab;
    double mn     = 4.0 * A_ab / (A_abp1 * A_abp1);
    mn            = (mn < 1.0) ? mn : 1.0; // stl::min(1, .)
    double mx     = (0.0 < mn) ? mn : 0.0; // stl::max(0, .)
    double phi_ab = 0.5 * dmy3 * mx;
    return -phi_ab * (dmy1 + dmy2);
}

// momentumAndEnergyJLoop (hydro_ve/momentum_energy_kern.hpp:65-221, tdpdTrho absent) over a
// range (computeMomentumEnergyImpl, hydro_ve/momentum_energy.hpp:40-100); returns minDtCourant
double orc_ve_momentum_energy(uint32_t first, uint32_t last, const uint32_t* nbrs, const unsigned* nc, unsigned ngmax,
                              const OrcVe* f, const sphx_box* b, int avClean, double Atmin, double Atmax,
                              double ramp, float Kcour, double* gpx, double* gpy, double* gpz, double* du)
{
    Box box = makeBox(b);
    const double K = f->K;
    double minDt = INFINITY;
#pragma omp parallel for schedule(static) reduction(min : minDt)
    for (int64_t ii = int64_t(first); ii < int64_t(last); ++ii)
    {
        size_t i = size_t(ii), ni = i - first;
        unsigned cnt = std::min(nc[i] - 1, ngmax);
        const uint32_t* nb = nbrs + size_t(ngmax) * ni;
        double xi = f->x[i], yi = f->y[i], zi = f->z[i], vxi = f->vx[i], vyi = f->vy[i], vzi = f->vz[i];
        double hi = f->h[i], mi = f->m[i], ci = f->c[i], kxi = f->kx[i], alpha_i = f->alpha[i];
        double xmassi = f->xm[i];
        double rhoi   = kxi * mi / xmassi;
        double prhoi  = f->prho[i];
        double hiInv  = 1.0 / hi, hiInv3 = hiInv * hiInv * hiInv;
        double maxv = 0.0, mx = 0.0, my = 0.0, mz = 0.0, energy = 0.0, avisc_e = 0.0;
        double a11 = f->c11[i], a12 = f->c12[i], a13 = f->c13[i], a22 = f->c22[i], a23 = f->c23[i], a33 = f->c33[i];
        double gi[6] = {0, 0, 0, 0, 0, 0};
        if (avClean) gi[0] = f->dV11[i], gi[1] = f->dV12[i], gi[2] = f->dV13[i], gi[3] = f->dV22[i],
                     gi[4] = f->dV23[i], gi[5] = f->dV33[i];
        double eta_crit = std::cbrt(32.0 * M_PI / 3.0 / double(cnt + 1));
        for (unsigned pj = 0; pj < cnt; ++pj)
        {
            uint32_t j = nb[pj];
            double rx = xi - f->x[j], ry = yi - f->y[j], rz = zi - f->z[j];
            double vxj = f->vx[j], vyj = f->vy[j], vzj = f->vz[j];
            applyPBC(box, 2.0 * hi, rx, ry, rz);
            double r2   = rx * rx + ry * ry + rz * rz;
            double dist = std::sqrt(r2);
            double vxij = vxi - vxj, vyij = vyi - vyj, vzij = vzi - vzj;
            double hj = f->h[j], hjInv = 1.0 / hj;
            double v1 = dist * hiInv, v2 = dist * hjInv;
            double hjInv3 = hjInv * hjInv * hjInv;
            double Wi = hiInv3 * lookup(f->wh, v1);
            double Wj = hjInv3 * lookup(f->wh, v2);
            double t1i = -(a11 * rx + a12 * ry + a13 * rz) * Wi;
            double t2i = -(a12 * rx + a22 * ry + a23 * rz) * Wi;
            double t3i = -(a13 * rx + a23 * ry + a33 * rz) * Wi;
            double b11 = f->c11[j], b12 = f->c12[j], b13 = f->c13[j], b22 = f->c22[j], b23 = f->c23[j], b33 = f->c33[j];
            double t1j = -(b11 * rx + b12 * ry + b13 * rz) * Wj;
            double t2j = -(b12 * rx + b22 * ry + b23 * rz) * Wj;
            double t3j = -(b13 * rx + b23 * ry + b33 * rz) * Wj;
            double mj = f->m[j], cj = f->c[j], kxj = f->kx[j], xmassj = f->xm[j];
            double rhoj = kxj * mj / xmassj;
            double rv   = rx * vxij + ry * vyij + rz * vzij;
            if (avClean)
            {
                double R[3]  = {rx, ry, rz};
                double gj[6] = {f->dV11[j], f->dV12[j], f->dV13[j], f->dV22[j], f->dV23[j], f->dV33[j]};
                rv += avRvCorrection(R, (v2 < v1) ? v2 : v1, eta_crit, gi, gj);
            }
            double wij = rv / dist;
            // artificial_viscosity (sph/kernels.hpp:60-76), beta = 2
            double visc = 0.0;
            if (wij < 0.0)
            {
                double vij_signal = (alpha_i + f->alpha[j]) * 0.25 * (ci + cj) - 2.0 * wij;
                visc              = -vij_signal * wij;
            }
            double vijsignal = 0.5 * (ci + cj) - 2.0 * wij;
            maxv             = (vijsignal > maxv) ? vijsignal : maxv;
            double a_mom, b_mom;
            double Atwood = (std::abs(rhoi - rhoj)) / (rhoi + rhoj);
            if (Atwood < Atmin)
            {
                a_mom = xmassi * xmassi;
                b_mom = xmassj * xmassj;
            }
            else if (Atwood > Atmax)
            {
                a_mom = xmassi * xmassj;
                b_mom = a_mom;
            }
            else
            {
                double sigma_ij = ramp * (Atwood - Atmin);
                a_mom           = std::pow(xmassi, 2.0 - sigma_ij) * std::pow(xmassj, sigma_ij);
                b_mom           = std::pow(xmassj, 2.0 - sigma_ij) * std::pow(xmassi, sigma_ij);
            }
            double a_visc = mj / rhoi * visc, b_visc = mj / rhoj * visc;
            double avx = 0.5 * (a_visc * t1i + b_visc * t1j);
            double avy = 0.5 * (a_visc * t2i + b_visc * t2j);
            double avz = 0.5 * (a_visc * t3i + b_visc * t3j);
            avisc_e += avx * vxij + avy * vyij + avz * vzij;
            energy += mj * a_mom * (vxij * t1i + vyij * t2i + vzij * t3i);
            double mom_i = mj * prhoi * a_mom;
            double mom_j = mj * f->prho[j] * b_mom;
            mx += mom_i * t1i + mom_j * t1j + avx;
            my += mom_i * t2i + mom_j * t2j + avy;
            mz += mom_i * t3i + mom_j * t3j + avz;
        }
        avisc_e = (0.0 < avisc_e) ? avisc_e : 0.0;
        du[i]   = K * (prhoi * energy + 0.5 * avisc_e);
        gpx[i]  = -K * mx;
        gpy[i]  = -K * my;
        gpz[i]  = -K * mz;
        // tsKCourant (sph/kernels.hpp:10-17)
        double v  = maxv > 0.0 ? maxv : ci;
        double dt = hi > 0.0 ? double(Kcour * hi / v) : INFINITY;
        minDt     = std::min(minDt, dt);
    }
    return minDt;
}

// tabulateFunction of getSphKernelDerivative(sinc_n) (sph_kernel_tables.hpp:88-158,
// kernels.hpp:45-58): whd[k] = n sinc(pi v / 2)^(n-1) d/dv sinc(pi v / 2)
void orc_kernel_derivative_table(double sincIndex, double* whd)
{
    const int    numIntervals = kTableSize - 1;
    const double dx           = 2.0 / numIntervals;
    for (int k = 0; k < kTableSize; ++k)
    {
        double v = 0.0 + double(size_t(k)) * dx;
        double d = 0.0;
        if (v != 0.0)
        {
            const double Pv    = M_PI_2 * v;
            const double sincv = std::sin(Pv) / Pv;
            d                  = sincv * M_PI_2 * ((std::cos(Pv) / std::sin(Pv)) - 1.0 / Pv);
        }
        whd[k] = sincIndex * std::pow(wharmonic(v), sincIndex - 1) * d;
    }
}

} // extern "C"
