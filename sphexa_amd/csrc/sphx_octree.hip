// sphx_octree.hip - cornerstone leaf array, internal octree link, node geometry.
//
// Leaf construction.  The reference converges its leaf array by repeated
// rebalancing (split nodes with count > bucket, merge 8 sibling leaves whose
// parent count <= bucket; domain/include/cstone/tree/csarray.hpp:268-440).  The
// fixed point of that iteration is unique: a node is a leaf iff its count is
// <= bucket (or it sits at the maximum level 21) and its parent's count is
// > bucket.  Equivalently, with S = {nodes at levels 0..20 holding > bucket
// particles} ("split nodes"), the leaves are {0} U {first key of child c of s :
// s in S, c = 1..7} followed by 2^63, i.e. 7|S| + 1 leaves.
//
// On sorted keys every split node is identified by its first particle i: node
// (level l, prefix of key[i]) is in S iff i starts its level-l group
// (l > floor(commonPrefix(key[i-1], key[i]) / 3)) and key[i + bucket] shares the
// same level-l prefix (l <= floor(commonPrefix(key[i], key[i+bucket]) / 3)).  So a
// single pass over the keys counts the split nodes per particle, a scan places
// them, a second pass emits their 7 non-leading child keys, and one radix sort of
// 7|S| + 1 keys yields the leaves; binary searches give the counts.  No host
// round trip per refinement level, one 4-byte readback (|S|) in total.
#include <algorithm>

#include "sphx_internal.hpp"

namespace sphx
{

__device__ inline void splitLevels(const uint64_t* keys, uint32_t n, uint32_t bucket, uint32_t i, int& lo, int& hi)
{
    uint64_t k = keys[i];
    lo         = (i == 0) ? 0 : commonPrefix(keys[i - 1], k) / 3 + 1;
    if (uint64_t(i) + bucket < n)
    {
        int h = commonPrefix(k, keys[i + bucket]) / 3;
        hi    = h < int(kMaxLevel) - 1 ? h : int(kMaxLevel) - 1;
    }
    else { hi = -1; }
}

// numS[i]: the split nodes that start at particle i; *maxLevel: the deepest split level.
// Grid-stride over the particles with one atomic per block: a device-scope access per
// wave to the one address serializes (it made this pass 9x slower than its key reads).
__global__ __launch_bounds__(256) void csCountSplits(const uint64_t* __restrict__ keys, uint32_t n, uint32_t bucket,
                                                     uint32_t* __restrict__ numS, uint32_t* __restrict__ maxLevel)
{
    __shared__ int wtop[4];
    int            top = -1;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u)
    {
        int lo, hi;
        splitLevels(keys, n, bucket, i, lo, hi);
        numS[i] = hi >= lo ? uint32_t(hi - lo + 1) : 0u;
        if (hi >= lo && hi > top) top = hi;
    }
    top = waveMax(top);
    if ((threadIdx.x & 63u) == 0) wtop[threadIdx.x >> 6] = top;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        top = max(max(wtop[0], wtop[1]), max(wtop[2], wtop[3]));
        if (top >= 0) atomicMax(maxLevel, uint32_t(top));
    }
}

// sOfs: the exclusive scan of the split counts, n + 1 entries; a particle that starts no
// split node (sOfs[i + 1] == sOfs[i], nearly all of them) reads no key
__global__ __launch_bounds__(256) void csEmitChildren(const uint64_t* __restrict__ keys, uint32_t n, uint32_t bucket,
                                                      const uint32_t* __restrict__ sOfs,
                                                      uint64_t* __restrict__ childKeys)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || sOfs[i + 1] == sOfs[i]) return;
    int lo, hi;
    splitLevels(keys, n, bucket, i, lo, hi);
    uint64_t k    = keys[i];
    uint64_t base = uint64_t(sOfs[i]) * 7;
    for (int l = lo; l <= hi; ++l)
    {
        uint64_t nodeKey = k & ~(nodeRange(unsigned(l)) - 1);
        uint64_t r1      = nodeRange(unsigned(l) + 1);
        for (int c = 1; c < 8; ++c)
            childKeys[base + uint64_t(l - lo) * 7 + (c - 1)] = nodeKey + uint64_t(c) * r1;
    }
}

__global__ void csFinishLeaves(uint64_t* leaves, uint32_t numLeaves)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) leaves[numLeaves] = kRootRange;
}

__device__ inline uint32_t lowerBound(const uint64_t* a, uint32_t n, uint64_t v)
{
    uint32_t lo = 0, cnt = n;
    while (cnt > 0)
    {
        uint32_t step = cnt / 2, mid = lo + step;
        if (a[mid] < v)
        {
            lo = mid + 1;
            cnt -= step + 1;
        }
        else { cnt = step; }
    }
    return lo;
}

// layout[l] = first particle with key >= leaves[l]; counts = capped differences
__global__ __launch_bounds__(256) void csLayout(const uint64_t* __restrict__ keys, uint32_t n,
                                                const uint64_t* __restrict__ leaves, uint32_t numLeaves,
                                                uint32_t* __restrict__ layout)
{
    uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > numLeaves) return;
    layout[l] = (l == numLeaves) ? n : lowerBound(keys, n, leaves[l]);
}

__global__ __launch_bounds__(256) void csCounts(const uint32_t* __restrict__ layout, uint32_t numLeaves,
                                                uint32_t maxCount, uint32_t* __restrict__ counts)
{
    uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= numLeaves) return;
    uint32_t c = layout[l + 1] - layout[l];
    counts[l]  = c < maxCount ? c : maxCount;
}

// ---------------------------------------------------------------------------
// internal octree: the node layout of cstone::OctreeData (tree/octree.hpp:284-358):
// nodes sorted by placeholder-bit prefix (level-major, then SFC), internal nodes
// numbered in "unsorted" order by octIndex = (leafIdx + binaryKeyWeight) / 7
// (tree/octree.hpp:56-100); the arrays must match the reference element for
// element, so the unsorted index formula is the reference's format.
__device__ inline int digitWeight(int digit)
{
    int fourGeqMask = -int(digit >= 4);
    return ((7 - digit) & fourGeqMask) - (digit & ~fourGeqMask);
}

__global__ __launch_bounds__(256) void octUnsortedLayout(const uint64_t* __restrict__ leaves, int numInternal,
                                                         int numLeaves, uint64_t* __restrict__ prefixes,
                                                         uint32_t* __restrict__ nodeIdx)
{
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= numLeaves) return;
    uint64_t key   = leaves[tid];
    uint64_t next  = leaves[tid + 1];
    unsigned level = treeLevel(next - key);
    prefixes[tid + numInternal] = encodePlaceholderBit(key, 3 * int(level));
    nodeIdx[tid + numInternal]  = uint32_t(tid + numInternal);
    unsigned prefixLength = unsigned(commonPrefix(key, next));
    if (prefixLength % 3 == 0 && tid < numLeaves - 1)
    {
        int w = 0;
        for (unsigned l = 1; l <= prefixLength / 3 + 1; ++l)
            w += digitWeight(int(octalDigit(key, l)));
        int octIndex       = (tid + w) / 7;
        prefixes[octIndex] = encodePlaceholderBit(key, int(prefixLength));
        nodeIdx[octIndex]  = uint32_t(octIndex);
    }
}

__global__ __launch_bounds__(256) void octInvert(int32_t* __restrict__ internalToLeaf,
                                                 int32_t* __restrict__ leafToInternal, int numNodes, int numInternal)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numNodes) return;
    int v             = internalToLeaf[i];
    leafToInternal[v] = i;
    internalToLeaf[i] = v - numInternal;
}

__global__ void octLevelRange(const uint64_t* __restrict__ prefixes, int numNodes, int32_t* __restrict__ levelRange)
{
    int level = threadIdx.x;
    if (level <= int(kMaxLevel))
        levelRange[level] = int32_t(lowerBound(prefixes, uint32_t(numNodes), encodePlaceholderBit(0, 3 * level)));
    else if (level == int(kMaxLevel) + 1)
        levelRange[level] = numNodes;
}

__global__ __launch_bounds__(256) void octZero(int32_t* a, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ __launch_bounds__(256) void octLink(const uint64_t* __restrict__ prefixes, int numInternal,
                                               const int32_t* __restrict__ leafToInternal,
                                               const int32_t* __restrict__ levelRange,
                                               int32_t* __restrict__ childOffsets, int32_t* __restrict__ parents)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numInternal) return;
    int      idxA         = leafToInternal[i];
    uint64_t prefix       = prefixes[idxA];
    uint64_t nodeKey      = decodePlaceholderBit(prefix);
    unsigned prefixLength = decodePrefixLength(prefix);
    unsigned level        = prefixLength / 3;
    if (level >= kMaxLevel) return;
    uint64_t childPrefix = encodePlaceholderBit(nodeKey, int(prefixLength) + 3);
    int      s = levelRange[level + 1], e = levelRange[level + 2];
    int      childIdx = s + int(lowerBound(prefixes + s, uint32_t(e - s), childPrefix));
    if (childIdx != e && prefixes[childIdx] == childPrefix)
    {
        childOffsets[idxA]          = childIdx;
        parents[(childIdx - 1) / 8] = idxA;
    }
}

// node centres/sizes: hilbertIBox (sfc/hilbert.hpp:259-275) + centerAndSize
// (sfc/box.hpp:317-334) with separately rounded operations
__global__ __launch_bounds__(256) void octCenters(const uint64_t* __restrict__ prefixes, int numNodes, DBox box,
                                                  double* __restrict__ centers, double* __restrict__ sizes)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numNodes) return;
    uint64_t prefix   = prefixes[i];
    uint64_t startKey = decodePlaceholderBit(prefix);
    unsigned level    = decodePrefixLength(prefix) / 3;
    unsigned cube     = (1u << kMaxLevel) >> level;
    unsigned mask     = ~(cube - 1);
    unsigned ix, iy, iz;
    hilbertDecode(startKey, ix, iy, iz);
    int          imin[3] = {int(ix & mask), int(iy & mask), int(iz & mask)};
    const double uL      = 1.0 / double(1u << kMaxLevel);
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        int    imax        = imin[d] + int(cube);
        double halfUnit    = 0.5 * uL * box.len[d];
        centers[3 * i + d] = box.lo[d] + double(imax + imin[d]) * halfUnit;
        sizes[3 * i + d]   = double(imax - imin[d]) * halfUnit;
    }
}

} // namespace sphx

using namespace sphx;

extern "C"
{

int sphx_compute_octree(sphx_ctx* ctx, const uint64_t* keys, uint64_t n, uint32_t bucketSize, uint32_t maxCount,
                        int32_t* numLeaves, uint64_t** leaves, uint32_t** counts, uint32_t** layout)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!numLeaves || !leaves || !counts || !layout) return argError(ctx, "compute_octree: null output pointer");
    if (n > 0 && !keys) return argError(ctx, "compute_octree: null keys");
    if (bucketSize == 0) return argError(ctx, "compute_octree: bucketSize must be > 0");
    if (maxCount <= bucketSize) return argError(ctx, "compute_octree: maxCount must exceed bucketSize");
    if (n >= (uint64_t(1) << 31)) return argError(ctx, "compute_octree: n must be < 2^31");
    uint32_t nn = uint32_t(n);

    // sOfs[0..nn]: the scan of the split-node counts (sOfs[nn] = S); sOfs[nn + 1]: the
    // deepest split level, read back with S
    uint32_t  S = 0, maxSplit = 0;
    uint32_t* sOfs = scratchT<uint32_t>(ctx, S_CS_NUMS, uint64_t(nn) + 2);
    if (!sOfs) return SPHX_ERR_NOMEM;
    if (nn > 0)
    {
        SPHX_HIP(ctx, hipMemsetAsync(sOfs + nn + 1, 0, sizeof(uint32_t), ctx->stream));
        {
            KTimer kt_(ctx, "csCountSplits");
            const unsigned grid = std::min<unsigned>(gridFor(nn, 256), unsigned(ctx->numCU) * 8);
            csCountSplits<<<grid, 256, 0, ctx->stream>>>(keys, nn, bucketSize, sOfs, sOfs + nn + 1);
        }
        SPHX_LAUNCH_CHECK(ctx, "csCountSplits");
        SPHX_TRY(exclusiveScanU32(ctx, sOfs, sOfs, nn));
        uint32_t sm[2];
        SPHX_TRY(readBack(ctx, sOfs + nn, sm, sizeof(sm)));
        S = sm[0], maxSplit = sm[1];
    }
    uint64_t L = uint64_t(S) * 7 + 1;
    if (L >= (uint64_t(1) << 31)) return argError(ctx, "compute_octree: too many leaves");
    uint64_t* child  = scratchT<uint64_t>(ctx, S_CS_CHILD, L + 1);
    uint32_t* dummy  = scratchT<uint32_t>(ctx, S_CS_CHILDV, L + 1);
    uint32_t* cnts   = scratchT<uint32_t>(ctx, S_CS_COUNTS, L + 1);
    uint32_t* lay    = scratchT<uint32_t>(ctx, S_CS_LAYOUT, L + 1);
    if (!child || !dummy || !cnts || !lay) return SPHX_ERR_NOMEM;
    // child key 0 (the first leaf) sits at index 7S
    SPHX_HIP(ctx, hipMemsetAsync(child + uint64_t(S) * 7, 0, sizeof(uint64_t), ctx->stream));
    if (S > 0)
    {
        {
            KTimer kt_(ctx, "csEmitChildren");
            csEmitChildren<<<gridFor(nn, 256), 256, 0, ctx->stream>>>(keys, nn, bucketSize, sOfs, child);
        }
        SPHX_LAUNCH_CHECK(ctx, "csEmitChildren");
        // children of level-l nodes are multiples of the level-(l + 1) range 8^(20 - l): the
        // low 3 (20 - maxSplit) bits of every child key are zero.  The digits above all but
        // always vary (the root's children differ in bits 60..62, level maxSplit + 1
        // children in the lowest digit kept), so the sort runs them without the
        // constant-digit check and its host read (a pass over a constant digit changes
        // nothing)
        SPHX_TRY(radixSortPairs(ctx, child, nullptr, L, 64, 3 * (int(kMaxLevel) - 1 - int(maxSplit)), false));
    }
    csFinishLeaves<<<1, 64, 0, ctx->stream>>>(child, uint32_t(L));
    SPHX_LAUNCH_CHECK(ctx, "csFinishLeaves");
    {
        KTimer kt_(ctx, "csLayout");
        csLayout<<<gridFor(L + 1, 256), 256, 0, ctx->stream>>>(keys, nn, child, uint32_t(L), lay);
    }
    SPHX_LAUNCH_CHECK(ctx, "csLayout");
    csCounts<<<gridFor(L, 256), 256, 0, ctx->stream>>>(lay, uint32_t(L), maxCount, cnts);
    SPHX_LAUNCH_CHECK(ctx, "csCounts");
    ctx->csNumLeaves = int32_t(L);
    *numLeaves = int32_t(L);
    *leaves    = child;
    *counts    = cnts;
    *layout    = lay;
    return SPHX_OK;
}

int sphx_octree_sizes(int32_t numLeafNodes, int32_t* numInternalNodes, int32_t* numNodes, int32_t* numParents)
{
    if (numLeafNodes < 1) return SPHX_ERR_ARG;
    int32_t ni = (numLeafNodes - 1) / 7;
    int32_t nn = numLeafNodes + ni;
    if (numInternalNodes) *numInternalNodes = ni;
    if (numNodes) *numNodes = nn;
    if (numParents) *numParents = std::max(1, (nn - 1) / 8);
    return SPHX_OK;
}

int sphx_build_octree(sphx_ctx* ctx, const uint64_t* leaves, sphx_octree* t)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!leaves || !t) return argError(ctx, "build_octree: null pointer");
    int L = t->numLeafNodes;
    if (L < 1 || (L - 1) % 7 != 0) return argError(ctx, "build_octree: numLeafNodes must be 7k+1 (cornerstone)");
    if (!t->prefixes || !t->childOffsets || !t->parents || !t->levelRange || !t->internalToLeaf ||
        !t->leafToInternal)
        return argError(ctx, "build_octree: null tree array");
    int ni = (L - 1) / 7, nn = L + ni;
    t->numInternalNodes = ni;
    t->numNodes         = nn;

    {
        KTimer kt_(ctx, "octUnsortedLayout");
        octUnsortedLayout<<<gridFor(L, 256), 256, 0, ctx->stream>>>(leaves, ni, L, t->prefixes,
                                                                reinterpret_cast<uint32_t*>(t->internalToLeaf));
    }
    SPHX_LAUNCH_CHECK(ctx, "octUnsortedLayout");
    SPHX_TRY(radixSortPairs(ctx, t->prefixes, reinterpret_cast<uint32_t*>(t->internalToLeaf), uint64_t(nn), 64));
    octInvert<<<gridFor(nn, 256), 256, 0, ctx->stream>>>(t->internalToLeaf, t->leafToInternal, nn, ni);
    SPHX_LAUNCH_CHECK(ctx, "octInvert");
    octLevelRange<<<1, 64, 0, ctx->stream>>>(t->prefixes, nn, t->levelRange);
    SPHX_LAUNCH_CHECK(ctx, "octLevelRange");
    octZero<<<gridFor(nn + 1, 256), 256, 0, ctx->stream>>>(t->childOffsets, nn + 1);
    SPHX_LAUNCH_CHECK(ctx, "octZero");
    if (ni > 0)
    {
        {
            KTimer kt_(ctx, "octLink");
            octLink<<<gridFor(ni, 256), 256, 0, ctx->stream>>>(t->prefixes, ni, t->leafToInternal, t->levelRange,
                                                           t->childOffsets, t->parents);
        }
        SPHX_LAUNCH_CHECK(ctx, "octLink");
    }
    else
    {
        // a single-leaf tree still has one parent slot (OctreeData::resize); the reference leaves it as is
        octZero<<<1, 64, 0, ctx->stream>>>(t->parents, 1);
        SPHX_LAUNCH_CHECK(ctx, "octZero");
    }
    return SPHX_OK;
}

int sphx_node_centers(sphx_ctx* ctx, const uint64_t* prefixes, int32_t numNodes, const sphx_box* box,
                      double* centers, double* sizes)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (numNodes <= 0) return SPHX_OK;
    if (!prefixes || !centers || !sizes) return argError(ctx, "node_centers: null pointer");
    if (!validBox(box)) return argError(ctx, "node_centers: invalid box");
    {
        KTimer kt_(ctx, "octCenters");
        octCenters<<<gridFor(numNodes, 256), 256, 0, ctx->stream>>>(prefixes, numNodes, makeDBox(box), centers, sizes);
    }
    SPHX_LAUNCH_CHECK(ctx, "octCenters");
    return SPHX_OK;
}

} // extern "C"
