// sphx_internal.hpp - shared context, scratch arena and device helpers of libsphx.
//
// Target: gfx950 (CDNA4), wave64.  All kernels are launched on ctx->stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/sphx.h"

namespace sphx
{

// ---------------------------------------------------------------------------
// scratch arena: one grow-only device buffer per slot
enum Slot : int
{
    S_SORT_KEYS = 0, // radix sort ping-pong keys
    S_SORT_VALS,     // radix sort ping-pong values
    S_SORT_HIST,     // per-tile digit histograms (+ the digit totals)
    S_SCAN0,         // device-wide scan block sums, level 0
    S_SCAN1,         // level 1
    S_SCAN2,         // level 2
    S_CS_NUMS,       // cornerstone: split-node count per particle (+ scan)
    S_CS_CHILD,      // cornerstone: emitted child keys -> sorted leaves
    S_CS_CHILDV,     // cornerstone: dummy sort values
    S_CS_COUNTS,     // cornerstone: leaf counts
    S_CS_LAYOUT,     // cornerstone: leaf layout
    S_NS_RANGES,     // neighbor search: candidate particle ranges per group
    S_NS_RCOUNT,     // neighbor search: range count per group
    S_NS_LISTS,      // neighbor bitmasks [sub-group][word][16]
    S_NS_UNION,      // neighbor unions per sub-group
    S_NS_USIZE,      // union sizes
    S_NS_STATE,      // per-particle h-iteration counters / active flags
    S_NS_REDO,       // groups that need another h iteration
    S_NS_RETRY,      // groups whose tree walk exceeded the small capacities
    S_NS_TESTS,      // candidate tests per launched group
    S_GROUPS,        // computed group boundaries
    S_GROUPS_TMP,    // group split scratch
    S_SMALL,         // small device scalars (flags, counters, reductions)
    S_TABLE64,       // sinc table, FP64
    S_TABLE32,       // sinc table, FP32
    S_TABLE_D,       // kernel-derivative table (VE gradh loop), FP64
    S_HY_REC,        // packed neighbor-side records of the SPH loops
    S_HFAC,          // smoothing-length update factors per neighbor count (FP64, then FP32)
    S_SORT_VAR,      // radix sort: per-tile OR and AND of the keys (first pass)
    S_MINDT,         // SPH loops: per-wave time-step minima before their reduction
    S_HY_REC_IAD,    // fused std-hydro step: IAD records written by the density loop
    S_HY_REC_ME,     // fused std-hydro step: momentum-energy records written by the IAD loop
    S_SORT_KEYS2,    // three-pass prefix sort: second work keys
    S_SORT_VALS2,    // three-pass prefix sort: second work values
    S_NUM_SLOTS
};

struct DevBuf
{
    void*  p     = nullptr;
    size_t bytes = 0;
};

// target groups on the device: explicit [start, end) arrays or fixed 64-particle groups
struct GroupsDev
{
    uint32_t        first = 0, last = 0, numGroups = 0;
    const uint32_t* start = nullptr;
    const uint32_t* end   = nullptr;
};

// Blocks are dealt round-robin to the 8 XCDs of the device (block b runs on XCD
// b mod 8).  This bijection of [0, numBlocks) gives each XCD one contiguous range of
// work items, so that spatially adjacent items share that XCD's L2.
__device__ inline uint32_t xcdContiguous(uint32_t b, uint32_t numBlocks)
{
    uint32_t perXcd = (numBlocks + 7u) / 8u;
    uint32_t xcd = b & 7u, idx = b >> 3;
    uint32_t full = numBlocks - 8u * (perXcd - 1u); // XCDs that get perXcd blocks
    return xcd < full ? xcd * perXcd + idx : full * perXcd + (xcd - full) * (perXcd - 1u) + idx;
}

__host__ __device__ inline void groupRange(const GroupsDev& G, uint32_t g, uint32_t& s, uint32_t& e)
{
    if (G.start)
    {
        s = G.start[g];
        e = G.end[g];
    }
    else
    {
        s = G.first + 64u * g;
        e = (s + 64u < G.last) ? s + 64u : G.last;
    }
}

// Smoothing-length update factors of updateH (sph/kernels.hpp:27-33):
// h' = (h * 0.5) * f(nc), f(nc) = pow(1 + 1023 ng0 / nc, 0.1).  The factor depends only on
// the integers ng0 and nc, so it is tabulated on the host with the host libm (the pow the
// reference CPU path calls) for nc in [0, cap): f64 in double (updateH<double>), f32 in
// float arithmetic with powf (updateH<float>, the reference app's FP32 h).  A count at or
// beyond cap raises an error instead of falling back to the device pow.
constexpr uint32_t kHFactorCap = 1u << 17;
struct HFactors
{
    const double* f64 = nullptr;
    const float*  f32 = nullptr;
    uint32_t      cap = 0;
};

template<class T>
__device__ inline const T* hFactorsOf(const HFactors& F)
{
    if constexpr (sizeof(T) == 8) return F.f64;
    else return F.f32;
}

// updateH with the tabulated factor; false (h unchanged) when nc >= F.cap
template<class T>
__device__ inline bool updateHTab(const HFactors& F, uint32_t nc, T& h)
{
    if (nc >= F.cap) return false;
    h = (h * T(0.5)) * hFactorsOf<T>(F)[nc];
    return true;
}

// neighbor-list layout: GPU [group][k][lane] or the reference CPU layout [i - first][k]
struct ListLayout
{
    uint32_t strideK; // distance between consecutive entries of one particle
    uint32_t cpu;     // 1: [i - first][k]
};

__host__ __device__ inline size_t listBase(const ListLayout& lay, const GroupsDev& G, uint32_t g, uint32_t s,
                                           uint32_t lane, uint32_t ngmax)
{
    return lay.cpu ? size_t(s + lane - G.first) * ngmax : size_t(g) * ngmax * 64 + lane;
}

// SPH neighbor lists: per 16-target sub-group sg (targets 16 q .. 16 q + 15 of group
// g, sg = 4 g + q), the ascending union of the targets' stored neighbors, and per
// target a bitmask over that union: bit b of masks[sg][w][t] set = union entry
// 32 w + b is a stored neighbor of target t.  Set bits ascend with j, so a target's
// neighbors read in bit order come in ascending j, the reference's list order.
struct SphLists
{
    uint32_t* masks  = nullptr; // [sg][words][16]
    uint32_t* unions = nullptr; // [sg][capU]
    uint32_t* usize  = nullptr; // [sg]
    uint32_t  capU   = 0;       // 16 * ngmax, a multiple of 32
    uint32_t  words  = 0;       // capU / 32
};

__host__ __device__ inline const uint32_t* maskRow(const SphLists& L, uint32_t sg, uint32_t t)
{
    return L.masks + size_t(sg) * L.words * 16 + t;
}

// entries of a particle's list that the loops read: min(nc - 1, ngmax), 0 when nc == 0
__host__ __device__ inline uint32_t listCount(uint32_t ncSelf, uint32_t ngmax)
{
    return ncSelf ? (ncSelf - 1u < ngmax ? ncSelf - 1u : ngmax) : 0u;
}

} // namespace sphx

struct sphx_ctx
{
    int             device = 0;
    hipStream_t     stream = nullptr;
    int             numCU  = 256;
    std::string     err;
    sphx::DevBuf    buf[sphx::S_NUM_SLOTS];
    // neighbor-list bookkeeping: the SPH loops require lists of the same particle set
    sphx::GroupsDev nsGroups;
    sphx::SphLists  nsLists;
    uint32_t        nsNgmax = 0;
    const uint32_t* nsNc    = nullptr;
    bool            nsValid = false;
    int             fp64Exact = 0; // 1: FP64 loops without FMA contraction (bit-exact vs the CPU path)
    int             loopShape[2] = {0, 0}; // per precision (FP64, MIXED): 0 LDS-staged, 1 sequential
    int             sortPrefixSkip = 0;    // key sorts left to run whole after a prefix sort fell back
    bool            sortPrefix24   = false; // the last prefix sort found every 24-bit prefix run short
    double          nsTests      = 0;      // candidate tests of the last neighbor search
    // per-kernel HIP-event timing (SPHX_OPT_TIMING)
    struct TimedLaunch
    {
        const char* name;
        hipEvent_t  a, b;
    };
    int                      timing = 0;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t>  eventPool;
    size_t                   eventsUsed = 0;
    // cornerstone results (context-owned)
    int32_t csNumLeaves = 0;
    // computed groups
    uint32_t grpNumGroups = 0;
    // smoothing-length factor tables in S_HFAC, built for this ng0 (-1: none yet)
    long long hfacNg0 = -1;
    // this device's address of the allocation-free sort's work array (nullptr: unavailable)
    void* noallocScratch = nullptr;
    // pinned host staging for small device-to-host reads (kHostStageBytes; nullptr: pageable
    // reads) and the event that marks a started read
    void*      hostStage = nullptr;
    hipEvent_t readEvent = nullptr;
    // deferred Courant time step (SPHX_OPT_DEFER_MIN_DT): the momentum-energy entries copy
    // the reduced minimum into minDtHost (pinned) and record minDtEvent; sphx_min_dt_courant
    // waits for it.  minDtLast: the last value read.
    int        deferMinDt   = 0;
    double*    minDtHost    = nullptr;
    hipEvent_t minDtEvent   = nullptr;
    bool       minDtPending = false;
    double     minDtLast    = INFINITY;
};

namespace sphx
{

int fail(sphx_ctx* ctx, int code, const std::string& msg);

// Small device-to-host reads through the context's pinned staging buffer.  readBack copies
// bytes from src and waits for the stream; readBackStart enqueues the copy and marks it
// with an event, and readBackFinish waits for that copy alone (work enqueued after the
// start keeps running) and delivers the bytes.  One started read at a time.
constexpr size_t kHostStageBytes = 4096;
int              readBack(sphx_ctx* ctx, const void* src, void* dst, size_t bytes);
int              readBackStart(sphx_ctx* ctx, const void* src, size_t bytes);
int              readBackFinish(sphx_ctx* ctx, void* dst, size_t bytes);
int failHip(sphx_ctx* ctx, hipError_t e, const char* what);
int argError(sphx_ctx* ctx, const char* what);

// the smoothing-length factor tables for ng0 (built on the host on first use per ng0)
int hFactors(sphx_ctx* ctx, uint32_t ng0, HFactors& out);

// grow-only scratch; returns nullptr (and sets the error) on allocation failure
void* scratch(sphx_ctx* ctx, Slot s, size_t bytes);

template<class T>
T* scratchT(sphx_ctx* ctx, Slot s, size_t count)
{
    return static_cast<T*>(scratch(ctx, s, count * sizeof(T) + 256));
}

// launch check after a kernel launch
#define SPHX_LAUNCH_CHECK(ctx, name)                                                                                   \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = hipGetLastError();                                                                             \
        if (e_ != hipSuccess) return ::sphx::failHip(ctx, e_, name);                                                   \
    } while (0)

#define SPHX_HIP(ctx, call)                                                                                            \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return ::sphx::failHip(ctx, e_, #call);                                                  \
    } while (0)

#define SPHX_TRY(expr)                                                                                                 \
    do                                                                                                                 \
    {                                                                                                                  \
        int s_ = (expr);                                                                                               \
        if (s_ != SPHX_OK) return s_;                                                                                  \
    } while (0)

inline unsigned gridFor(uint64_t n, unsigned block) { return unsigned((n + block - 1) / block); }

// records HIP events around one kernel launch on ctx->stream when timing is enabled
hipEvent_t timingEvent(sphx_ctx* ctx);
struct KTimer
{
    sphx_ctx*   ctx;
    const char* name;
    hipEvent_t  a = nullptr;
    // recorded when the context times every hot kernel (timing 1), or only the heavy ones
    // (timing 2: the neighbor candidates and the SPH particle loops) and this is one of them
    KTimer(sphx_ctx* c, const char* n, bool particleLoop = false)
        : ctx(c)
        , name(n)
    {
        if (ctx && (ctx->timing == 1 || (ctx->timing == 2 && particleLoop)) && (a = timingEvent(ctx)))
            (void)hipEventRecord(a, ctx->stream);
    }
    void stop()
    {
        if (!a) return;
        hipEvent_t b = timingEvent(ctx);
        if (b)
        {
            (void)hipEventRecord(b, ctx->stream);
            ctx->timed.push_back({name, a, b});
        }
        a = nullptr;
    }
    ~KTimer() { stop(); }
};

// ---------------------------------------------------------------------------
// device-wide primitives (sphx_primitives.hip)

// exclusive scan of n uint32 values into out[0..n] (out[n] = total); in may equal out
int exclusiveScanU32(sphx_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n);

// stable LSD radix sort of 64-bit keys (+ optional 32-bit values) in place;
// only the bits in [0, keyBits) are examined
// lowBits: bits below it are known to be equal across all keys (passes start at the digit
// holding bit lowBits)
// checkDigits = false: no constant-digit check (no host read; every pass from lowBits runs)
int radixSortPairs(sphx_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n, int keyBits, int lowBits = 0,
                   bool checkDigits = true);

// ---------------------------------------------------------------------------
// key bit algebra (64-bit Hilbert keys, 21 levels, 63 key bits)
constexpr unsigned kMaxLevel  = 21;
constexpr uint64_t kRootRange = uint64_t(1) << 63;
constexpr uint64_t kRemoveKey = uint64_t(1) << 63;

__host__ __device__ inline unsigned clz64(uint64_t k)
{
#ifdef __HIP_DEVICE_COMPILE__
    return k == 0 ? 64u : unsigned(__clzll(k));
#else
    return k == 0 ? 64u : unsigned(__builtin_clzll(k));
#endif
}
__host__ __device__ inline uint64_t nodeRange(unsigned level) { return uint64_t(1) << (3u * (kMaxLevel - level)); }
__host__ __device__ inline unsigned treeLevel(uint64_t range) { return (clz64(range - 1) - 1) / 3; }
__host__ __device__ inline int commonPrefix(uint64_t a, uint64_t b) { return int(clz64(a ^ b)) - 1; }
__host__ __device__ inline uint64_t encodePlaceholderBit(uint64_t code, int prefixLength)
{
    return (uint64_t(1) << prefixLength) | (code >> (3 * int(kMaxLevel) - prefixLength));
}
__host__ __device__ inline unsigned decodePrefixLength(uint64_t code) { return 63u - clz64(code); }
__host__ __device__ inline uint64_t decodePlaceholderBit(uint64_t code)
{
    unsigned len = decodePrefixLength(code);
    return (code ^ (uint64_t(1) << len)) << (3 * kMaxLevel - len);
}
__host__ __device__ inline unsigned octalDigit(uint64_t code, unsigned position)
{
    return unsigned(code >> (3u * (kMaxLevel - position))) & 7u;
}

// Hilbert encode: octant order 0,1,3,2,7,6,4,5 packed as 3-bit fields
__host__ __device__ inline uint64_t hilbertKey(unsigned px, unsigned py, unsigned pz)
{
    constexpr uint32_t m2h = (0u << 0) | (1u << 3) | (3u << 6) | (2u << 9) | (7u << 12) | (6u << 15) | (4u << 18) |
                             (5u << 21);
    uint64_t key = 0;
    for (int level = int(kMaxLevel) - 1; level >= 0; --level)
    {
        unsigned xi = (px >> level) & 1u;
        unsigned yi = (py >> level) & 1u;
        unsigned zi = (pz >> level) & 1u;
        unsigned octant = (xi << 2) | (yi << 1) | zi;
        key = (key << 3) + ((m2h >> (3u * octant)) & 7u);

        px ^= 0u - (xi & ((!yi) | zi));
        py ^= 0u - ((xi & (yi | zi)) | (yi & (!zi)));
        pz ^= 0u - ((xi & (!yi) & (!zi)) | (yi & (!zi)));

        unsigned tx = px, ty = py, tz = pz;
        if (zi) { px = ty, py = tz, pz = tx; }       // cyclic rotation
        else if (!yi) { px = tz, pz = tx; }          // swap x and z
    }
    return key;
}

// hilbertKey as a state machine.  Encoding a level reflects and permutes the remaining
// coordinate bits, so before each level the coordinates are seen through a state =
// (which axis feeds each of the three slots, which slots are reflected): 6 x 8 = 48
// states, numbered 8 * permutation + reflections; the initial state 0 is the identity.
// hilbertStep encodes one level from the state and the level's raw bits
// (x << 2 | y << 1 | z): it returns the output digit and advances the state, the same
// updates hilbertKey applies to px, py, pz.
__host__ __device__ inline uint32_t hilbertStep(uint32_t& state, uint32_t raw)
{
    constexpr uint32_t m2h = (0u << 0) | (1u << 3) | (3u << 6) | (2u << 9) | (7u << 12) | (6u << 15) | (4u << 18) |
                             (5u << 21);
    // slot axes of the 6 permutations, 2 bits per slot (slot 0 lowest), 6 bits per permutation
    constexpr uint64_t perms = uint64_t(0u | 1u << 2 | 2u << 4) | uint64_t(0u | 2u << 2 | 1u << 4) << 6 |
                               uint64_t(1u | 0u << 2 | 2u << 4) << 12 | uint64_t(1u | 2u << 2 | 0u << 4) << 18 |
                               uint64_t(2u | 0u << 2 | 1u << 4) << 24 | uint64_t(2u | 1u << 2 | 0u << 4) << 30;
    const uint32_t pa = uint32_t(perms >> (6u * (state >> 3))) & 63u;
    uint32_t       a0 = pa & 3u, a1 = (pa >> 2) & 3u, a2 = (pa >> 4) & 3u;
    uint32_t       f0 = state & 1u, f1 = (state >> 1) & 1u, f2 = (state >> 2) & 1u;
    const uint32_t xi = ((raw >> (2u - a0)) & 1u) ^ f0;
    const uint32_t yi = ((raw >> (2u - a1)) & 1u) ^ f1;
    const uint32_t zi = ((raw >> (2u - a2)) & 1u) ^ f2;
    const uint32_t digit = (m2h >> (3u * ((xi << 2) | (yi << 1) | zi))) & 7u;
    f0 ^= xi & ((!yi) | zi);
    f1 ^= (xi & (yi | zi)) | (yi & (!zi));
    f2 ^= (xi & (!yi) & (!zi)) | (yi & (!zi));
    if (zi)
    {
        const uint32_t ta = a0, tf = f0;
        a0 = a1, f0 = f1;
        a1 = a2, f1 = f2;
        a2 = ta, f2 = tf;
    }
    else if (!yi)
    {
        const uint32_t ta = a0, tf = f0;
        a0 = a2, f0 = f2;
        a2 = ta, f2 = tf;
    }
    const uint32_t code = a0 | a1 << 2 | a2 << 4;
    uint32_t       p    = 0;
    for (uint32_t k = 1; k < 6; ++k)
        if ((uint32_t(perms >> (6u * k)) & 63u) == code) p = k;
    state = 8u * p + (f0 | f1 << 1 | f2 << 2);
    return digit;
}

__host__ __device__ inline void hilbertDecode(uint64_t key, unsigned& ox, unsigned& oy, unsigned& oz)
{
    unsigned px = 0, py = 0, pz = 0;
    for (unsigned level = 0; level < kMaxLevel; ++level)
    {
        unsigned octant = unsigned(key >> (3 * level)) & 7u;
        unsigned xi = octant >> 2u, yi = (octant >> 1u) & 1u, zi = octant & 1u;
        unsigned tx = px, ty = py, tz = pz;
        if (yi ^ zi) { px = tz, pz = ty, py = tx; }
        else if ((!xi & !yi & !zi) || (xi & yi & zi)) { px = tz, pz = tx; }
        unsigned mask = (1u << level) - 1;
        px ^= mask & (0u - (xi & (yi | zi)));
        py ^= mask & (0u - ((xi & ((!yi) | (!zi))) | ((!xi) & yi & zi)));
        pz ^= mask & (0u - ((xi & (!yi) & (!zi)) | (yi & zi)));
        px |= (xi << level);
        py |= ((xi ^ yi) << level);
        pz |= ((yi ^ zi) << level);
    }
    ox = px, oy = py, oz = pz;
}

// ---------------------------------------------------------------------------
// box with the derived quantities of cstone::Box
struct DBox
{
    double lo[3], hi[3], len[3], ilen[3];
    int    pbc[3]; // 1 if periodic
    int    anyPbc;
};

inline DBox makeDBox(const sphx_box* b)
{
    DBox r;
    r.lo[0] = b->xmin, r.hi[0] = b->xmax;
    r.lo[1] = b->ymin, r.hi[1] = b->ymax;
    r.lo[2] = b->zmin, r.hi[2] = b->zmax;
    int bc[3] = {b->bx, b->by, b->bz};
    r.anyPbc = 0;
    for (int d = 0; d < 3; ++d)
    {
        r.len[d]  = r.hi[d] - r.lo[d];
        r.ilen[d] = 1.0 / (r.hi[d] - r.lo[d]);
        r.pbc[d]  = (bc[d] == SPHX_BC_PERIODIC);
        r.anyPbc |= r.pbc[d];
    }
    return r;
}

bool validBox(const sphx_box* b);

// ---------------------------------------------------------------------------
// wave64 helpers
__device__ inline unsigned laneId() { return __lane_id(); }
__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ inline unsigned popc64(uint64_t m) { return unsigned(__popcll(m)); }
__device__ inline uint64_t lanemaskLt() { return (uint64_t(1) << laneId()) - 1; }
// number of set bits of m below this lane (v_mbcnt)
__device__ inline uint32_t rankIn(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

template<class T>
__device__ inline T waveMin(T v)
{
    for (int o = 32; o > 0; o >>= 1)
    {
        T w = __shfl_xor(v, o, 64);
        v   = w < v ? w : v;
    }
    return v;
}
template<class T>
__device__ inline T waveMax(T v)
{
    for (int o = 32; o > 0; o >>= 1)
    {
        T w = __shfl_xor(v, o, 64);
        v   = w > v ? w : v;
    }
    return v;
}
template<class T>
__device__ inline T waveSum(T v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

} // namespace sphx
