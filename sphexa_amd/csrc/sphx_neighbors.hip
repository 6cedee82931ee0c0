// sphx_neighbors.hip - wave64 target-group neighbor search.
//
// One wave owns one target group (<= 64 SFC-consecutive particles, one per lane).
//   nsGroupRanges : breadth-first walk of the octree with the group's search box
//                   (union of the targets' 2h spheres, minimum image in periodic
//                   dimensions); overlapping leaves are collected in LDS, sorted
//                   (leaf order = SFC order = particle order) and merged into
//                   contiguous particle ranges.
//   nsCandidates  : the ranges are streamed in chunks of 64 particles; each chunk is
//                   filtered against the regions of the wave's four 16-target quarters
//                   (points within the largest 2h of the targets' bounding box) and
//                   queued per quarter in LDS; a quarter's queue is tested
//                   64 candidates at a time (lane = candidate, one target per step)
//                   with the reference predicate d^2 < (2h)^2 evaluated in FP64 with
//                   separately rounded operations (findneighbors.hpp:33-60,115-140),
//                   so the neighbor sets are bit-identical to the CPU path; hits are
//                   appended in ascending j, which is also the CPU's depth-first
//                   order, so truncation at ngmax keeps the same entries.
// The smoothing-length iteration of findNeighborsSph (sph/find_neighbors.hpp:10-41)
// re-runs only the groups that still have particles outside [ng0/4, ngmax + 1].
#include <algorithm>
#include <climits>
#include <cstddef>
#include <type_traits>

#include "sphx_internal.hpp"

namespace sphx
{

// traversal capacities: every group is first walked with the small ones (LDS for more
// waves per CU); the groups that exceed them are walked again with the large ones
constexpr int NS_FCAP_SMALL = 512;  // BFS frontier capacity (per buffer)
constexpr int NS_LCAP_SMALL = 512;  // candidate leaves per group
constexpr int NS_FCAP       = 2048;
constexpr int NS_LCAP       = 1024; // also the stride of the per-group range lists

struct NsFlags
{
    uint32_t           redoCount;  // groups appended to the redo list
    uint32_t           overflow;   // traversal capacity exceeded
    uint32_t           badGroup;   // a group with more than 64 particles
    uint32_t           retryCount; // groups that exceeded the small traversal capacities
    uint32_t           hfacMiss;   // a count to update h from lies beyond the factor table
    unsigned long long tests;      // exact (target, candidate) distance tests of this pass
};

__device__ inline bool nodeOverlapsBox(const double* __restrict__ centers, const double* __restrict__ sizes, int node,
                                       const double gc[3], const double gh[3], const DBox& box)
{
    bool ov = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        double dx = centers[3 * node + d] - gc[d];
        if (box.pbc[d]) dx -= box.len[d] * rint(dx * box.ilen[d]);
        ov = ov && (fabs(dx) <= sizes[3 * node + d] + gh[d]);
    }
    return ov;
}

// The ranges of one group g (all 64 lanes of the wave call this).  A group whose walk
// exceeds the capacities gets no ranges and is appended to retryList, or, without a
// retry list, raises the overflow flag.
template<class Th, int FCAP, int LCAP>
__device__ void groupRangesOf(uint32_t g, GroupsDev G, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, const Th* __restrict__ h,
                              const int32_t* __restrict__ childOffsets, const int32_t* __restrict__ internalToLeaf,
                              const uint32_t* __restrict__ layout, const double* __restrict__ centers,
                              const double* __restrict__ sizes, const DBox& box, uint2* __restrict__ ranges,
                              uint32_t* __restrict__ rangeCount, uint32_t* __restrict__ retryList,
                              NsFlags* __restrict__ flags)
{
    static_assert((LCAP & (LCAP - 1)) == 0 && LCAP <= NS_LCAP, "leaf capacity: a power of two within the range stride");
    __shared__ int32_t front[2][FCAP];
    __shared__ int32_t leafList[LCAP];

    const unsigned lane = threadIdx.x;
    uint32_t       s, e;
    groupRange(G, g, s, e);
    if (e - s > 64)
    {
        if (lane == 0) atomicOr(&flags->badGroup, 1u);
        e = s + 64;
    }
    uint32_t i   = s + lane;
    bool     act = i < e;

    double lo[3], hi[3];
    {
        double px[3] = {0, 0, 0};
        double r     = 0;
        if (act)
        {
            px[0] = x[i], px[1] = y[i], px[2] = z[i];
            r     = 2.0 * double(h[i]);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            lo[d] = waveMin(act ? px[d] - r : INFINITY);
            hi[d] = waveMax(act ? px[d] + r : -INFINITY);
        }
    }
    if (!(lo[0] <= hi[0]))
    {
        if (lane == 0) rangeCount[g] = 0;
        return;
    }
    double gc[3], gh[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        gc[d] = 0.5 * (lo[d] + hi[d]);
        gh[d] = 0.5 * (hi[d] - lo[d]);
        gh[d] += 1e-12 * (fabs(gc[d]) + gh[d]); // never under-cover the union of spheres
    }

    // ---- breadth-first traversal
    int  lcount   = 0;
    bool overflow = false;
    int  fcount   = 0;
    int  cur      = 0;
    if (childOffsets[0] == 0)
    {
        if (lane == 0) leafList[0] = internalToLeaf[0];
        lcount = 1;
    }
    else
    {
        if (lane < 8) front[0][lane] = childOffsets[0] + int(lane);
        fcount = 8;
    }
    __syncthreads();
    const uint64_t lt = lanemaskLt();
    while (fcount > 0 && !overflow)
    {
        int ncount = 0;
        int nxt    = cur ^ 1;
        for (int b = 0; b < fcount; b += 64)
        {
            int idx  = b + int(lane);
            int node = idx < fcount ? front[cur][idx] : -1;
            // the node's child offset and leaf index are loaded with its box (one memory
            // round trip per level), whether or not the box overlaps
            int  childOf = 0, leafOf = 0;
            if (node >= 0) childOf = childOffsets[node], leafOf = internalToLeaf[node];
            bool ov    = node >= 0 && nodeOverlapsBox(centers, sizes, node, gc, gh, box);
            int  child = ov ? childOf : 0;
            bool leaf  = ov && child == 0;
            bool inner = ov && child != 0;

            uint64_t mL   = ballot(leaf);
            int      posL = lcount + int(popc64(mL & lt));
            if (leaf && posL < LCAP) leafList[posL] = leafOf;
            lcount += int(popc64(mL));

            uint64_t mI   = ballot(inner);
            int      posI = ncount + 8 * int(popc64(mI & lt));
            if (inner && posI + 8 <= FCAP)
            {
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    front[nxt][posI + c] = child + c;
            }
            ncount += 8 * int(popc64(mI));
        }
        __syncthreads();
        if (lcount > LCAP || ncount > FCAP) overflow = true;
        cur    = nxt;
        fcount = ncount;
    }
    if (overflow)
    {
        if (lane == 0)
        {
            rangeCount[g] = 0;
            if (retryList) retryList[atomicAdd(&flags->retryCount, 1u)] = g;
            else atomicOr(&flags->overflow, 1u);
        }
        return;
    }

    // ---- bitonic sort of the candidate leaves (ascending)
    int P = 1;
    while (P < lcount)
        P <<= 1;
    for (int k = lcount + int(lane); k < P; k += 64)
        leafList[k] = INT_MAX;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1)
    {
        for (int stride = size >> 1; stride > 0; stride >>= 1)
        {
            for (int t = int(lane); t < P / 2; t += 64)
            {
                int  a   = 2 * t - (t & (stride - 1));
                int  b2  = a + stride;
                bool up  = ((a & size) == 0);
                int  va  = leafList[a], vb = leafList[b2];
                bool swp = up ? (va > vb) : (va < vb);
                if (swp)
                {
                    leafList[a]  = vb;
                    leafList[b2] = va;
                }
            }
            __syncthreads();
        }
    }

    // ---- merge consecutive leaves into particle ranges
    int rcount = 0;
    for (int b = 0; b < lcount; b += 64)
    {
        int      k       = b + int(lane);
        bool     inRange = k < lcount;
        bool     isStart = inRange && (k == 0 || leafList[k] != leafList[k - 1] + 1);
        uint64_t mS      = ballot(isStart);
        int      r       = rcount + int(popc64(mS & lt));
        if (isStart)
        {
            // find the end of this run
            int kk = k;
            while (kk + 1 < lcount && leafList[kk + 1] == leafList[kk] + 1)
                ++kk;
            uint32_t ps = layout[leafList[k]];
            uint32_t pe = layout[leafList[kk] + 1];
            ranges[size_t(g) * NS_LCAP + r] = make_uint2(ps, pe);
        }
        rcount += int(popc64(mS));
    }
    if (lane == 0) rangeCount[g] = uint32_t(rcount);
}

// all groups (or those of groupList) with the small capacities
template<class Th>
__global__ __launch_bounds__(64) void nsGroupRanges(GroupsDev G, const uint32_t* __restrict__ groupList,
                                                    const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const Th* __restrict__ h,
                                                    const int32_t* __restrict__ childOffsets,
                                                    const int32_t* __restrict__ internalToLeaf,
                                                    const uint32_t* __restrict__ layout,
                                                    const double* __restrict__ centers,
                                                    const double* __restrict__ sizes, DBox box,
                                                    uint2* __restrict__ ranges, uint32_t* __restrict__ rangeCount,
                                                    uint32_t* __restrict__ retryList, NsFlags* __restrict__ flags)
{
    const uint32_t blk = xcdContiguous(blockIdx.x, gridDim.x);
    const uint32_t g   = groupList ? groupList[blk] : blk;
    groupRangesOf<Th, NS_FCAP_SMALL, NS_LCAP_SMALL>(g, G, x, y, z, h, childOffsets, internalToLeaf, layout, centers,
                                                    sizes, box, ranges, rangeCount, retryList, flags);
}

// the groups of the retry list with the large capacities (grid-stride over the list,
// whose length the first walk left in flags->retryCount)
template<class Th>
__global__ __launch_bounds__(64) void nsGroupRangesRetry(GroupsDev G, const double* __restrict__ x,
                                                         const double* __restrict__ y, const double* __restrict__ z,
                                                         const Th* __restrict__ h,
                                                         const int32_t* __restrict__ childOffsets,
                                                         const int32_t* __restrict__ internalToLeaf,
                                                         const uint32_t* __restrict__ layout,
                                                         const double* __restrict__ centers,
                                                         const double* __restrict__ sizes, DBox box,
                                                         uint2* __restrict__ ranges, uint32_t* __restrict__ rangeCount,
                                                         const uint32_t* __restrict__ retryList,
                                                         NsFlags* __restrict__ flags)
{
    const uint32_t n = flags->retryCount;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x)
    {
        groupRangesOf<Th, NS_FCAP, NS_LCAP>(retryList[k], G, x, y, z, h, childOffsets, internalToLeaf, layout,
                                             centers, sizes, box, ranges, rangeCount, nullptr, flags);
        __syncthreads(); // the next group reuses the LDS lists
    }
}

// Candidate test.  Sph = true: findNeighborsSph semantics (nc counts self, h updates,
// redo flags) and the SPH list format below; Sph = false: findNeighbors semantics
// (count without self, u32 lists in the reference CPU layout).
//
// SPH list format, per sub-group of 16 consecutive targets (quarter q of the wave,
// sg = 4 g + q): the sorted union of the stored neighbors of its 16 targets
// (unions[sg][0..usize)) and per target a bitmask over that union
// (masks[sg][w][t]).
//
// Thread mapping (transposed test).  The group's candidate ranges are streamed in
// chunks of 64 particles, lane = candidate; every chunk is filtered against the
// regions of the four quarters (16 targets each) and the survivors are appended, in
// ascending j, to per-quarter LDS queues.  A queue holding 64 candidates
// (or the rest, at the end) is tested as one batch: lane = candidate, and the wave
// walks the quarter's 16 targets, one target per step, so a step costs a handful of
// FP64 instructions for 64 tests and its ballot is the target's hit mask over the
// batch.  The hit masks give the batch's union entries (candidates stored by some
// target, in lane order = ascending j); a lane permutation compacts the entries and
// one ballot per target turns the compacted bits into that target's mask bits, which
// the target's lane appends to its mask word stream.
constexpr uint32_t NS_QCAP = 128; // per-quarter queue capacity (ring buffer)

// v_writelane_b32: `old` with lane `lane` replaced by the wave-uniform value v
__device__ int writeLane(int v, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

template<class Th, bool Sph>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void nsCandidates(GroupsDev G, const uint32_t* __restrict__ groupList,
                                                   const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, Th* __restrict__ h,
                                                   uint8_t* __restrict__ iter, const uint2* __restrict__ ranges,
                                                   const uint32_t* __restrict__ rangeCount, DBox box, uint32_t ngmax,
                                                   uint32_t ng0, HFactors hfac, uint32_t* __restrict__ nbrs, SphLists out,
                                                   uint32_t* __restrict__ nc, uint32_t* __restrict__ redoList,
                                                   NsFlags* __restrict__ flags,
                                                   unsigned long long* __restrict__ groupTests)
{
#pragma clang fp contract(off)
    __shared__ uint32_t qj[4][NS_QCAP]; // queued candidate indices per quarter
    __shared__ double   tgt[64][4];     // target of lane l: x, y, z, (2h)^2
    __shared__ uint32_t tid[64];        // its particle index
    __shared__ double   qbox[4][7];     // quarter: box centre xyz, half-widths xyz, (radius)^2

    const uint32_t blk  = xcdContiguous(blockIdx.x, gridDim.x);
    const uint32_t g    = groupList ? groupList[blk] : blk;
    const unsigned lane = threadIdx.x;
    uint32_t       s, e;
    groupRange(G, g, s, e);
    if (e - s > 64) e = s + 64;
    const uint32_t i   = s + lane;
    const bool     act = i < e;

    double xi = 0, yi = 0, zi = 0, hi = 0;
    if (act) xi = x[i], yi = y[i], zi = z[i], hi = double(h[i]);
    const double radiusSq = 4.0 * hi * hi;
    bool         inside   = true;
    {
        double p[3] = {xi, yi, zi};
#pragma unroll
        for (int d = 0; d < 3; ++d)
            inside = inside && (p[d] - 2.0 * hi >= box.lo[d]) && (p[d] + 2.0 * hi <= box.hi[d]);
    }
    const bool   usePbc = box.anyPbc && !inside;
    const double px     = box.pbc[0] ? box.len[0] : 0.0;
    const double py     = box.pbc[1] ? box.len[1] : 0.0;
    const double pz     = box.pbc[2] ? box.len[2] : 0.0;

    // quarter regions: the bounding box of the 16 targets' positions (centre, half-widths
    // widened by a relative 1e-12) and the largest search radius 2h of the quarter.  A
    // neighbor of any target of the quarter lies within that radius of the box (its image
    // nearest the box centre, when the quarter uses periodic images), so the filter keeps
    // candidates whose distance to the box is at most the radius, with a relative slack of
    // 1e-9 on the squared radius; a quarter without targets keeps nothing.
    const unsigned q = lane >> 4, t = lane & 15u;
    const uint64_t actMask = ballot(act);
    const uint64_t pbcMask = ballot(usePbc);
    {
        double pp[3] = {xi, yi, zi};
        double lo[3], hh[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            double a = act ? pp[d] : INFINITY, b = act ? pp[d] : -INFINITY;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1)
            {
                a = fmin(a, __shfl_xor(a, o, 64));
                b = fmax(b, __shfl_xor(b, o, 64));
            }
            lo[d] = a, hh[d] = b;
        }
        double rmax = act ? 2.0 * hi : -1.0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1)
            rmax = fmax(rmax, __shfl_xor(rmax, o, 64));
        if (t == 0)
        {
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                double c = 0.5 * (lo[d] + hh[d]), w = 0.5 * (hh[d] - lo[d]);
                qbox[q][d]     = c;
                qbox[q][3 + d] = w + 1e-12 * (fabs(c) + w + box.len[d]);
            }
            qbox[q][6] = rmax < 0.0 ? -1.0 : rmax * rmax * (1.0 + 1e-9);
        }
        tgt[lane][0] = xi, tgt[lane][1] = yi, tgt[lane][2] = zi, tgt[lane][3] = radiusSq;
        tid[lane]    = act ? i : 0xFFFFFFFFu;
    }
    __syncthreads();

    const uint64_t lt = lanemaskLt();
    // per-target state, in the target's lane
    uint32_t count = 0; // hits so far (stored or not)
    uint32_t word  = 0; // SPH: mask bits of union entries 32 * (usz / 32) ...
    uint32_t extra = 0; // hits of the current batch beyond the per-target cap (SPH)
    uint32_t delta = 0; // hits of the current batch (non-SPH)
    // per-quarter state (wave-uniform; indexed by compile-time quarter numbers only)
    uint32_t qhead[4] = {0, 0, 0, 0}, qlen[4] = {0, 0, 0, 0}, usz[4] = {0, 0, 0, 0};
    unsigned long long tests = 0;

    // test the first nb queued candidates of quarter Q: lane = candidate
    auto batch = [&](auto Qc, uint32_t nb) {
        constexpr uint32_t Q  = decltype(Qc)::value;
        constexpr uint32_t qb = 16u * Q;
        const bool         valid = lane < nb;
        uint32_t           jc    = 0;
        double             xc = 0, yc = 0, zc = 0;
        if (valid)
        {
            jc = qj[Q][(qhead[Q] + lane) & (NS_QCAP - 1)];
            xc = x[jc], yc = y[jc], zc = z[jc];
        }
        tests += (unsigned long long)nb * popc64((actMask >> qb) & 0xFFFFu);
        uint32_t hb = 0; // bit tt: target tt of the quarter hits this candidate
        // straight-line steps; an inactive target has (2h)^2 = 0 and never hits
        if (((pbcMask >> qb) & 0xFFFFu) == 0)
        {
#pragma unroll
            for (uint32_t tt = 0; tt < 16; ++tt)
            {
                const double dx = xc - tgt[qb + tt][0];
                const double dy = yc - tgt[qb + tt][1];
                const double dz = zc - tgt[qb + tt][2];
                const double d2 = dx * dx + dy * dy + dz * dz;
                hb |= (valid & (d2 < tgt[qb + tt][3])) ? (1u << tt) : 0u;
            }
        }
        else
        {
            // quarters with periodic targets (near the box faces): the targets that use
            // periodic images fold the offsets (volatile reads keep this rare path's loads
            // from being merged with the common path)
            const volatile double* tv = &tgt[qb][0];
#pragma unroll 4
            for (uint32_t tt = 0; tt < 16; ++tt)
            {
                double dx = xc - tv[4 * tt + 0];
                double dy = yc - tv[4 * tt + 1];
                double dz = zc - tv[4 * tt + 2];
                if ((pbcMask >> (qb + tt)) & 1u)
                {
                    dx -= px * rint(dx * box.ilen[0]);
                    dy -= py * rint(dy * box.ilen[1]);
                    dz -= pz * rint(dz * box.ilen[2]);
                }
                const double d2 = dx * dx + dy * dy + dz * dz;
                hb |= (valid & (d2 < tv[4 * tt + 3])) ? (1u << tt) : 0u;
            }
        }
        // a target is not its own neighbor
        {
            const uint32_t self = jc - (s + qb);
            if (valid & (self < 16u)) hb &= ~(1u << self);
        }

        const bool mine = q == Q;
        if constexpr (Sph)
        {
            // union entries of the batch (candidates stored by some target) compacted into
            // lanes 0..nU-1 by a lane permutation; target lane qb + tt receives its mask
            // bits over them in (cmHi:cmLo)
            uint32_t nU = 0, jp = 0, cmLo = 0, cmHi = 0;
            for (int pass = 0; pass < 2; ++pass)
            {
                const uint64_t U = ballot(hb != 0);
                nU               = popc64(U);
                cmLo = cmHi = 0;
                if (nU)
                {
                    const uint32_t dst = hb ? rankIn(U) : nU + rankIn(~U);
                    const uint32_t hp  = uint32_t(__builtin_amdgcn_ds_permute(int(4u * dst), int(hb)));
                    jp                 = uint32_t(__builtin_amdgcn_ds_permute(int(4u * dst), int(jc)));
#pragma unroll
                    for (uint32_t tt = 0; tt < 16; ++tt)
                    {
                        const uint64_t cm = ballot((hp >> tt) & 1u);
                        cmLo = uint32_t(writeLane(int(uint32_t(cm)), int(qb + tt), int(cmLo)));
                        cmHi = uint32_t(writeLane(int(uint32_t(cm >> 32)), int(qb + tt), int(cmHi)));
                    }
                }
                const uint32_t gain = uint32_t(__builtin_popcount(cmLo)) + uint32_t(__builtin_popcount(cmHi));
                if (pass == 1 || ballot(mine && count + gain > ngmax) == 0) break;
                // a target reaches its cap in this batch: keep only its first hits, count
                // the others, and compact again
#pragma unroll 1
                for (uint32_t tt = 0; tt < 16; ++tt)
                {
                    const bool     hit  = (hb >> tt) & 1u;
                    const uint64_t mh   = ballot(hit);
                    const uint32_t pc   = popc64(mh);
                    const uint32_t ct   = uint32_t(__builtin_amdgcn_readlane(int(count), int(qb + tt)));
                    const uint32_t room = ct < ngmax ? ngmax - ct : 0u;
                    if (pc > room)
                    {
                        if (hit && rankIn(mh) >= room) hb &= ~(1u << tt);
                        extra = uint32_t(writeLane(int(pc - room), int(qb + tt), int(extra)));
                    }
                }
            }
            if (lane < nU) out.unions[size_t(4u * g + Q) * out.capU + usz[Q] + lane] = jp;
            if (mine)
            {
                count += uint32_t(__builtin_popcount(cmLo)) + uint32_t(__builtin_popcount(cmHi)) + extra;
                extra = 0;
                if (nU)
                {
                    // append nU mask bits at bit offset usz of the target's word stream
                    const uint32_t u0 = usz[Q], o = u0 & 31u, end = o + nU;
                    const uint64_t cm64 = (uint64_t(cmHi) << 32) | cmLo;
                    const uint64_t lo64 = uint64_t(word) | (cm64 << o);
                    const uint32_t w2   = o ? uint32_t(cm64 >> (64u - o)) : 0u;
                    uint32_t* wp = out.masks + (size_t(4u * g + Q) * out.words + (u0 >> 5)) * 16 + t;
                    if (end >= 32) wp[0] = uint32_t(lo64);
                    if (end >= 64) wp[16] = uint32_t(lo64 >> 32);
                    word = end >= 64 ? w2 : (end >= 32 ? uint32_t(lo64 >> 32) : uint32_t(lo64));
                }
            }
            usz[Q] += nU;
        }
        else
        {
            // reference CPU layout: the k-th hit of target i goes to nbrs[(i - first) * ngmax + k]
#pragma unroll 1
            for (uint32_t tt = 0; tt < 16; ++tt)
            {
                const bool     hit = (hb >> tt) & 1u;
                const uint64_t mh  = ballot(hit);
                if (mh == 0) continue;
                const uint32_t ct  = uint32_t(__builtin_amdgcn_readlane(int(count), int(qb + tt)));
                const uint32_t pos = ct + rankIn(mh);
                if (hit && nbrs && pos < ngmax) nbrs[size_t(s + qb + tt - G.first) * ngmax + pos] = jc;
                delta = uint32_t(writeLane(int(popc64(mh)), int(qb + tt), int(delta)));
            }
            if (mine) count += delta;
            delta = 0;
        }
        qhead[Q] += nb;
        qlen[Q] -= nb;
    };

    // Stream the candidates in chunks of 64 (lane = candidate): filter each chunk against
    // the 4 quarter boxes and queue the survivors; then every quarter whose queue holds
    // 64 candidates tests one batch (a chunk adds at most 64, so queues stay below 128).
    // After the last chunk the remaining partial queues are tested.
    const uint32_t R   = rangeCount[g];
    uint32_t       r   = 0;
    uint2          rg  = R ? ranges[size_t(g) * NS_LCAP] : make_uint2(0u, 0u);
    uint32_t       cur = rg.x;
    for (;;)
    {
        const bool done = r >= R;
        if (!done)
        {
            const uint32_t m  = min(64u, rg.y - cur);
            const bool     in = lane < m;
            const uint32_t j  = cur + lane;
            double         ax = 0, ay = 0, az = 0;
            if (in) ax = x[j], ay = y[j], az = z[j];
#pragma unroll
            for (uint32_t qq = 0; qq < 4; ++qq)
            {
                double dx = ax - qbox[qq][0], dy = ay - qbox[qq][1], dz = az - qbox[qq][2];
                if ((pbcMask >> (16u * qq)) & 0xFFFFu)
                {
                    dx -= px * rint(dx * box.ilen[0]);
                    dy -= py * rint(dy * box.ilen[1]);
                    dz -= pz * rint(dz * box.ilen[2]);
                }
                const double ex   = fmax(fabs(dx) - qbox[qq][3], 0.0);
                const double ey   = fmax(fabs(dy) - qbox[qq][4], 0.0);
                const double ez   = fmax(fabs(dz) - qbox[qq][5], 0.0);
                const bool   pass = in & (ex * ex + ey * ey + ez * ez <= qbox[qq][6]);
                const uint64_t mk = ballot(pass);
                if (pass) qj[qq][(qhead[qq] + qlen[qq] + rankIn(mk)) & (NS_QCAP - 1)] = j;
                qlen[qq] += popc64(mk);
            }
            cur += m;
            if (cur >= rg.y && ++r < R)
            {
                rg  = ranges[size_t(g) * NS_LCAP + r];
                cur = rg.x;
            }
            __syncthreads();
        }
        if (qlen[0] >= 64 || (done && qlen[0])) batch(std::integral_constant<uint32_t, 0>{}, min(qlen[0], 64u));
        if (qlen[1] >= 64 || (done && qlen[1])) batch(std::integral_constant<uint32_t, 1>{}, min(qlen[1], 64u));
        if (qlen[2] >= 64 || (done && qlen[2])) batch(std::integral_constant<uint32_t, 2>{}, min(qlen[2], 64u));
        if (qlen[3] >= 64 || (done && qlen[3])) batch(std::integral_constant<uint32_t, 3>{}, min(qlen[3], 64u));
        if (done) break;
    }
    // exact pair tests of this pass
    // (stored per launched group and summed by nsSumTests: one atomic per group on a
    // single counter would serialize)
    if (lane == 0) groupTests[blk] = tests;

    if constexpr (Sph)
    {
        const uint32_t sg    = 4u * g + q;
        const uint32_t myUsz = q == 0 ? usz[0] : q == 1 ? usz[1] : q == 2 ? usz[2] : usz[3];
        if (t == 0) out.usize[sg] = myUsz;
        if (myUsz & 31u) out.masks[(size_t(sg) * out.words + myUsz / 32u) * 16 + t] = word;
        bool redo = false;
        if (act)
        {
            uint32_t ncSph = count + 1;
            uint32_t ngmin = ng0 / 4;
            bool     outR  = (ngmin > ncSph) || (ncSph - 1 > ngmax);
            nc[i] = ncSph; // dataset counts are indexed by particle (sph/find_neighbors.hpp:51)
            if (outR && iter[i] < 10)
            {
                // updateH (sph/kernels.hpp:27-33) in the precision of h, factor from the table
                Th hn = Th(hi);
                if (updateHTab(hfac, ncSph, hn))
                {
                    h[i] = hn;
                    iter[i] += 1;
                    redo = true;
                }
                else
                    flags->hfacMiss = 1;
            }
        }
        uint64_t anyRedo = ballot(redo);
        if (anyRedo && lane == 0)
        {
            uint32_t pos  = atomicAdd(&flags->redoCount, 1u);
            redoList[pos] = g;
        }
    }
    else
    {
        if (act) nc[i - G.first] = count;
    }
}

__global__ __launch_bounds__(256) void nsInitState(uint8_t* iter, uint32_t first, uint32_t last)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) return;
    iter[i] = 0;
}

__global__ void nsResetFlags(NsFlags* f, int clearErrors)
{
    if (threadIdx.x == 0)
    {
        f->redoCount  = 0;
        f->retryCount = 0;
        f->tests      = 0;
        if (clearErrors)
        {
            f->overflow = 0;
            f->badGroup = 0;
            f->hfacMiss = 0;
        }
    }
}

// flags->tests += sum of the per-group test counts (one atomic per block)
__global__ __launch_bounds__(256) void nsSumTests(const unsigned long long* __restrict__ groupTests, uint32_t n,
                                                  NsFlags* __restrict__ f)
{
    __shared__ unsigned long long part[4];
    unsigned long long            v = 0;
    for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < n; k += gridDim.x * 256u)
        v += groupTests[k];
    v = waveSum(v);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long tot = part[0] + part[1] + part[2] + part[3];
        if (tot) atomicAdd(&f->tests, tot);
    }
}

// export the SPH lists into the reference CPU layout [i - first][k]
__global__ __launch_bounds__(256) void nsExport(GroupsDev G, SphLists L, uint32_t ngmax,
                                                const uint32_t* __restrict__ nc, uint32_t* __restrict__ out)
{
    uint32_t g    = blockIdx.x * 4 + (threadIdx.x >> 6);
    uint32_t lane = threadIdx.x & 63u;
    if (g >= G.numGroups) return;
    uint32_t s, e;
    groupRange(G, g, s, e);
    uint32_t i = s + lane;
    if (i >= e || lane >= 64) return;
    uint32_t        sg   = 4u * g + (lane >> 4);
    const uint32_t* M    = maskRow(L, sg, lane & 15u);
    const uint32_t* uni  = L.unions + size_t(sg) * L.capU;
    const uint32_t  nw   = (L.usize[sg] + 31u) / 32u;
    uint32_t        k    = 0;
    for (uint32_t w = 0; w < nw; ++w)
    {
        uint32_t m = M[size_t(w) * 16];
        while (m)
        {
            uint32_t b = uint32_t(__builtin_ctz(m));
            m &= m - 1u;
            out[size_t(i - G.first) * ngmax + k++] = uni[32u * w + b];
        }
    }
}


// ---------------------------------------------------------------------------
// target groups: computeGroupSplits (domain/include/cstone/traversal/groups_gpu.cu:108-137,
// groups_gpu.cuh:37-77,148-216).  Fixed 64-particle groups are split after particle p
// when |X(p+1) - X(p)|^2 > min(distCrit^2, (2 h_p / minExtent)^2) in box-length units,
// distCrit = tolFactor * cbrt(volume of the smallest leaf holding a particle of the group).
//
// One wave per fixed group.  The leaves holding the group's particles p0..p1 are the
// non-empty leaves from leaf(p0) (the last leaf with layout <= p0; layout[0] = 0) through
// leaf(p1): the wave finds leaf(p0) with a 64-way search (four dependent probes for 2^24
// leaves) and then walks the leaves after it 64 at a time until one starts beyond p1,
// which for a group of 64 particles is nearly always one step.
template<class Th>
__global__ __launch_bounds__(256) void grpSplitMasks(uint32_t first, uint32_t last, uint32_t numFixed,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, const Th* __restrict__ h,
                                                     const uint64_t* __restrict__ leaves, int numLeaves,
                                                     const uint32_t* __restrict__ layout, DBox box, float tolFactor,
                                                     uint64_t* __restrict__ masks, uint32_t* __restrict__ numSegments)
{
#pragma clang fp contract(off)
    const uint32_t g    = blockIdx.x * 4u + (threadIdx.x >> 6);
    const unsigned lane = threadIdx.x & 63u;
    if (g >= numFixed) return; // whole waves
    const uint32_t p0 = first + 64u * g;
    const uint32_t p1 = min(p0 + 63u, last - 1);
    const uint32_t bi = min(p0 + lane, p1);
    // the particle's fields load while the leaf search below runs (they do not depend on it)
    const double xb = x[bi], yb = y[bi], zb = z[bi];
    const double hb = h ? double(h[bi]) : 0.0;

    int lo = 0, hi = numLeaves; // layout[lo] <= p0; leaf(p0) in [lo, hi)
    while (hi - lo > 1)
    {
        const int  stride = (hi - lo + 63) / 64;
        const int  m      = lo + int(lane) * stride;
        const bool le     = m < hi && layout[m] <= p0;
        const int  k      = 63 - int(clz64(ballot(le))); // lane 0 always holds
        hi                = min(hi, lo + (k + 1) * stride);
        lo                = lo + k * stride;
    }
    unsigned level = 0;
    for (int c = lo;; c += 64)
    {
        const int  m  = c + int(lane);
        const bool in = m < numLeaves && layout[m] <= p1; // m <= leaf(p1)
        if (in && layout[m + 1] > layout[m]) level = max(level, treeLevel(leaves[m + 1] - leaves[m]));
        if (ballot(in) != ~uint64_t(0)) break;
    }
    // unit-box leaf volume 2^(-3 level); its cube root 2^(-level) is exact
    int      maxLevel = int(waveMax(int(level)));
    double   distCrit = ldexp(1.0, -maxLevel) * double(tolFactor);
    double   minExt   = fmin(fmin(box.len[0], box.len[1]), box.len[2]);
    double   px = xb * box.ilen[0], py = yb * box.ilen[1], pz = zb * box.ilen[2];
    double   pw = h ? 2.0 * hb / minExt : 1.0;
    double   qx = __shfl_down(px, 1, 64), qy = __shfl_down(py, 1, 64), qz = __shfl_down(pz, 1, 64);
    if (lane == 63) qx = px, qy = py, qz = pz;
    double   dx = qx - px, dy = qy - py, dz = qz - pz;
    double   d2 = dx * dx + dy * dy + dz * dz;
    double   crit2 = distCrit * distCrit;
    bool     split = d2 > fmin(crit2, pw * pw);
    uint64_t m     = ballot(split);
    if (lane == 0)
    {
        masks[g]       = m;
        numSegments[g] = 1u + popc64(m);
    }
}

// group boundaries from the split masks (makeSplits, groups_gpu.cuh:91-114)
__global__ __launch_bounds__(256) void grpBoundaries(uint32_t first, uint32_t last, uint32_t numFixed,
                                                     const uint64_t* __restrict__ masks,
                                                     const uint32_t* __restrict__ segOfs,
                                                     uint32_t* __restrict__ bounds)
{
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= numFixed) return;
    uint32_t base = first + 64u * g;
    uint32_t o    = segOfs[g];
    uint64_t m    = masks[g];
    bounds[o++]   = base;
    while (m)
    {
        uint32_t b  = uint32_t(__ffsll((unsigned long long)m) - 1);
        bounds[o++] = base + b + 1;
        m &= m - 1;
    }
    if (g + 1 == numFixed) bounds[o] = last;
}

// ---------------------------------------------------------------------------
// host driver shared by both entry points
template<class Th, bool Sph>
static int runSearch(sphx_ctx* ctx, const GroupsDev& G, const double* x, const double* y, const double* z, Th* h,
                     const sphx_box* box, const sphx_octree* t, uint32_t ngmax, uint32_t ng0, uint32_t* nbrs,
                     SphLists lists, uint32_t* nc)
{
    if (G.numGroups == 0) return SPHX_OK;
    uint32_t  numG   = G.numGroups;
    uint2*    ranges = scratchT<uint2>(ctx, S_NS_RANGES, size_t(numG) * NS_LCAP);
    uint32_t* rcnt   = scratchT<uint32_t>(ctx, S_NS_RCOUNT, numG);
    uint8_t*  iter   = scratchT<uint8_t>(ctx, S_NS_STATE, size_t(G.last));
    uint32_t* redo   = scratchT<uint32_t>(ctx, S_NS_REDO, 2 * size_t(numG));
    uint32_t* retry  = scratchT<uint32_t>(ctx, S_NS_RETRY, numG);
    auto*     gtests = scratchT<unsigned long long>(ctx, S_NS_TESTS, numG);
    NsFlags*  flags  = scratchT<NsFlags>(ctx, S_SMALL, 1);
    if (!ranges || !rcnt || !iter || !redo || !retry || !gtests || !flags) return SPHX_ERR_NOMEM;
    DBox     db = makeDBox(box);
    HFactors hfac;
    if (Sph) SPHX_TRY(hFactors(ctx, ng0, hfac));

    if (G.last > G.first)
    {
        nsInitState<<<gridFor(G.last - G.first, 256), 256, 0, ctx->stream>>>(iter, G.first, G.last);
        SPHX_LAUNCH_CHECK(ctx, "nsInitState");
    }
    nsResetFlags<<<1, 64, 0, ctx->stream>>>(flags, 1);
    SPHX_LAUNCH_CHECK(ctx, "nsResetFlags");
    ctx->nsTests = 0;

    const uint32_t* list      = nullptr; // nullptr: all groups
    uint32_t        numLaunch = numG;
    int             pingpong  = 0;
    for (int pass = 0; pass < 12 && numLaunch > 0; ++pass)
    {
        uint32_t* outList = redo + size_t(pingpong) * numG;
        {
            KTimer kt_(ctx, "nsGroupRanges");
            nsGroupRanges<Th><<<numLaunch, 64, 0, ctx->stream>>>(G, list, x, y, z, h, t->childOffsets,
                                                                t->internalToLeaf, t->layout, t->centers, t->sizes,
                                                                db, ranges, rcnt, retry, flags);
        }
        SPHX_LAUNCH_CHECK(ctx, "nsGroupRanges");
        {
            KTimer   kt_(ctx, "nsGroupRangesRetry");
            uint32_t grid = std::min<uint32_t>(numLaunch, uint32_t(ctx->numCU) * 4u);
            nsGroupRangesRetry<Th><<<grid, 64, 0, ctx->stream>>>(G, x, y, z, h, t->childOffsets, t->internalToLeaf,
                                                                 t->layout, t->centers, t->sizes, db, ranges, rcnt,
                                                                 retry, flags);
        }
        SPHX_LAUNCH_CHECK(ctx, "nsGroupRangesRetry");
        {
            KTimer kt_(ctx, "nsCandidates", true);
            nsCandidates<Th, Sph><<<numLaunch, 64, 0, ctx->stream>>>(G, list, x, y, z, h, iter, ranges, rcnt, db,
                                                                    ngmax, ng0, hfac, nbrs, lists, nc, outList,
                                                                    flags, gtests);
        }
        SPHX_LAUNCH_CHECK(ctx, "nsCandidates");
        nsSumTests<<<std::min<uint32_t>(256u, (numLaunch + 1023u) / 1024u), 256, 0, ctx->stream>>>(gtests, numLaunch,
                                                                                                flags);
        SPHX_LAUNCH_CHECK(ctx, "nsSumTests");
        NsFlags hf;
        SPHX_TRY(readBack(ctx, flags, &hf, sizeof(NsFlags)));
        ctx->nsTests += double(hf.tests);
        if (hf.overflow) return fail(ctx, SPHX_ERR_STACK, "neighbor search: traversal capacity exceeded");
        if (hf.badGroup) return argError(ctx, "neighbor search: a target group holds more than 64 particles");
        if (hf.hfacMiss)
            return fail(ctx, SPHX_ERR_STACK,
                        "neighbor search: a neighbor count to update h from exceeds the factor table (" +
                            std::to_string(kHFactorCap) + ")");
        if (!Sph || hf.redoCount == 0) break;
        numLaunch = hf.redoCount;
        list      = outList;
        pingpong ^= 1;
        nsResetFlags<<<1, 64, 0, ctx->stream>>>(flags, 0);
        SPHX_LAUNCH_CHECK(ctx, "nsResetFlags");
    }
    return SPHX_OK;
}

int groupsFromArgs(sphx_ctx* ctx, const sphx_groups* groups, uint32_t first, uint32_t last, GroupsDev& G)
{
    if (groups && groups->groupStart)
    {
        if (!groups->groupEnd) return argError(ctx, "groups: groupEnd is null");
        G.first     = groups->firstBody;
        G.last      = groups->lastBody;
        G.numGroups = groups->numGroups;
        G.start     = groups->groupStart;
        G.end       = groups->groupEnd;
    }
    else
    {
        G.first     = groups ? groups->firstBody : first;
        G.last      = groups ? groups->lastBody : last;
        G.numGroups = (G.last - G.first + 63) / 64;
        G.start     = nullptr;
        G.end       = nullptr;
    }
    if (G.last < G.first) return argError(ctx, "groups: lastBody < firstBody");
    return SPHX_OK;
}

static bool validTree(const sphx_octree* t)
{
    return t && t->numNodes > 0 && t->childOffsets && t->internalToLeaf && t->layout && t->centers && t->sizes;
}

} // namespace sphx

using namespace sphx;

extern "C"
{

int sphx_compute_groups(sphx_ctx* ctx, uint32_t first, uint32_t last, const double* x, const double* y,
                        const double* z, const void* h, int32_t precision, const sphx_octree* tree,
                        const sphx_box* box, float tolFactor, sphx_groups* groups)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!groups) return argError(ctx, "compute_groups: null output");
    if (last < first) return argError(ctx, "compute_groups: last < first");
    if (!validBox(box)) return argError(ctx, "compute_groups: invalid box");
    if (!tree || tree->numLeafNodes < 1 || !tree->leaves || !tree->layout)
        return argError(ctx, "compute_groups: tree needs numLeafNodes, leaves and layout");
    if (precision != SPHX_FP64 && precision != SPHX_MIXED) return argError(ctx, "compute_groups: unknown precision");
    uint32_t numFixed = (last - first + 63) / 64;
    *groups           = sphx_groups{first, last, 0, nullptr, nullptr};
    if (numFixed == 0) return SPHX_OK;
    if (!x || !y || !z) return argError(ctx, "compute_groups: null coordinates");
    uint64_t* masks = scratchT<uint64_t>(ctx, S_GROUPS_TMP, numFixed);
    uint32_t* segs  = scratchT<uint32_t>(ctx, S_NS_RCOUNT, size_t(numFixed) + 1);
    if (!masks || !segs) return SPHX_ERR_NOMEM;
    DBox db = makeDBox(box);
    {
        KTimer   kt_(ctx, "grpSplitMasks");
        uint32_t grid = (numFixed + 3u) / 4u;
        if (precision == SPHX_FP64)
            grpSplitMasks<double><<<grid, 256, 0, ctx->stream>>>(first, last, numFixed, x, y, z,
                                                                 static_cast<const double*>(h), tree->leaves,
                                                                 tree->numLeafNodes, tree->layout, db, tolFactor,
                                                                 masks, segs);
        else
            grpSplitMasks<float><<<grid, 256, 0, ctx->stream>>>(first, last, numFixed, x, y, z,
                                                                static_cast<const float*>(h), tree->leaves,
                                                                tree->numLeafNodes, tree->layout, db, tolFactor,
                                                                masks, segs);
    }
    SPHX_LAUNCH_CHECK(ctx, "grpSplitMasks");
    SPHX_TRY(exclusiveScanU32(ctx, segs, segs, numFixed));
    uint32_t numGroups = 0;
    SPHX_TRY(readBack(ctx, segs + numFixed, &numGroups, sizeof(uint32_t)));
    uint32_t* bounds = scratchT<uint32_t>(ctx, S_GROUPS, size_t(numGroups) + 1);
    if (!bounds) return SPHX_ERR_NOMEM;
    grpBoundaries<<<gridFor(numFixed, 256), 256, 0, ctx->stream>>>(first, last, numFixed, masks, segs, bounds);
    SPHX_LAUNCH_CHECK(ctx, "grpBoundaries");
    groups->numGroups  = numGroups;
    groups->groupStart = bounds;
    groups->groupEnd   = bounds + 1;
    return SPHX_OK;
}

int sphx_find_neighbors(sphx_ctx* ctx, const double* x, const double* y, const double* z, const double* h,
                        uint32_t first, uint32_t last, const sphx_box* box, const sphx_octree* tree, uint32_t ngmax,
                        uint32_t* nbrs, uint32_t* nc)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (last < first) return argError(ctx, "find_neighbors: last < first");
    if (last == first) return SPHX_OK;
    if (!x || !y || !z || !h || !nc) return argError(ctx, "find_neighbors: null pointer");
    if (!validBox(box)) return argError(ctx, "find_neighbors: invalid box");
    if (!validTree(tree)) return argError(ctx, "find_neighbors: incomplete tree view");
    GroupsDev G;
    SPHX_TRY(groupsFromArgs(ctx, nullptr, first, last, G));
    return runSearch<double, false>(ctx, G, x, y, z, const_cast<double*>(h), box, tree, ngmax, 0, nbrs, SphLists{},
                                    nc);
}

int sphx_find_neighbors_sph(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box,
                            const sphx_octree* tree)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!d || !d->x || !d->y || !d->z || !d->h || !d->nc) return argError(ctx, "find_neighbors_sph: null field");
    if (!validBox(box)) return argError(ctx, "find_neighbors_sph: invalid box");
    if (!validTree(tree)) return argError(ctx, "find_neighbors_sph: incomplete tree view");
    if (d->ng0 > d->ngmax) return fail(ctx, SPHX_ERR_NGMAX, "ng0 should be smaller than ngmax");
    if (d->ngmax == 0) return argError(ctx, "find_neighbors_sph: ngmax must be > 0");
    GroupsDev G;
    SPHX_TRY(groupsFromArgs(ctx, groups, 0, uint32_t(d->n), G));
    if (G.last > d->n) return argError(ctx, "find_neighbors_sph: group range exceeds n");
    if (d->ngmax > 4096) return argError(ctx, "find_neighbors_sph: ngmax must be <= 4096");
    ctx->nsValid = false;
    size_t   numSub = size_t(G.numGroups) * 4;
    SphLists L;
    L.capU   = 32u * ((16u * d->ngmax + 31u) / 32u);
    L.words  = L.capU / 32u;
    L.masks  = scratchT<uint32_t>(ctx, S_NS_LISTS, numSub * L.words * 16);
    L.unions = scratchT<uint32_t>(ctx, S_NS_UNION, numSub * L.capU);
    L.usize  = scratchT<uint32_t>(ctx, S_NS_USIZE, numSub);
    if (!L.masks || !L.unions || !L.usize) return SPHX_ERR_NOMEM;
    int st;
    if (d->precision == SPHX_FP64)
        st = runSearch<double, true>(ctx, G, d->x, d->y, d->z, static_cast<double*>(d->h), box, tree, d->ngmax,
                                     d->ng0, nullptr, L, d->nc);
    else if (d->precision == SPHX_MIXED)
        st = runSearch<float, true>(ctx, G, d->x, d->y, d->z, static_cast<float*>(d->h), box, tree, d->ngmax,
                                    d->ng0, nullptr, L, d->nc);
    else
        return argError(ctx, "find_neighbors_sph: unknown precision");
    if (st != SPHX_OK) return st;
    ctx->nsLists  = L;
    ctx->nsValid  = true;
    ctx->nsGroups = G;
    ctx->nsNgmax  = d->ngmax;
    ctx->nsNc     = d->nc;
    return SPHX_OK;
}

int sphx_export_neighbors(sphx_ctx* ctx, uint32_t* nbrs)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!ctx->nsValid) return fail(ctx, SPHX_ERR_STATE, "export_neighbors: no neighbor lists in context");
    if (!nbrs) return argError(ctx, "export_neighbors: null pointer");
    const GroupsDev& G = ctx->nsGroups;
    if (G.numGroups == 0) return SPHX_OK;
    nsExport<<<gridFor(G.numGroups, 4), 256, 0, ctx->stream>>>(G, ctx->nsLists, ctx->nsNgmax, ctx->nsNc, nbrs);
    SPHX_LAUNCH_CHECK(ctx, "nsExport");
    return SPHX_OK;
}

} // extern "C"
