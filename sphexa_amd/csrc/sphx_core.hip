// sphx_core.hip - context, scratch arena, device-wide scan, LSD radix sort,
// gather and Hilbert keys.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "sphx_internal.hpp"

namespace sphx
{

int fail(sphx_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int failHip(sphx_ctx* ctx, hipError_t e, const char* what)
{
    return fail(ctx, SPHX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int argError(sphx_ctx* ctx, const char* what) { return fail(ctx, SPHX_ERR_ARG, what); }

int readBackStart(sphx_ctx* ctx, const void* src, size_t bytes)
{
    if (!ctx->hostStage || !ctx->readEvent || bytes > kHostStageBytes)
        return fail(ctx, SPHX_ERR_STATE, "readBackStart: no pinned staging for this read");
    SPHX_HIP(ctx, hipMemcpyAsync(ctx->hostStage, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SPHX_HIP(ctx, hipEventRecord(ctx->readEvent, ctx->stream));
    return SPHX_OK;
}

int readBackFinish(sphx_ctx* ctx, void* dst, size_t bytes)
{
    SPHX_HIP(ctx, hipEventSynchronize(ctx->readEvent));
    std::memcpy(dst, ctx->hostStage, bytes);
    return SPHX_OK;
}

int readBack(sphx_ctx* ctx, const void* src, void* dst, size_t bytes)
{
    if (ctx->hostStage && bytes <= kHostStageBytes)
    {
        SPHX_HIP(ctx, hipMemcpyAsync(ctx->hostStage, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        SPHX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        std::memcpy(dst, ctx->hostStage, bytes);
        return SPHX_OK;
    }
    SPHX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SPHX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SPHX_OK;
}

void* scratch(sphx_ctx* ctx, Slot s, size_t bytes)
{
    DevBuf& b = ctx->buf[s];
    if (b.bytes >= bytes && b.p) return b.p;
    if (b.p)
    {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(b.p);
        b.p     = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes + bytes / 8, 4096);
    if (hipMalloc(&b.p, want) != hipSuccess)
    {
        (void)hipGetLastError();
        b.p = nullptr;
        fail(ctx, SPHX_ERR_NOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
        return nullptr;
    }
    b.bytes = want;
    return b.p;
}

int hFactors(sphx_ctx* ctx, uint32_t ng0, HFactors& out)
{
    const uint32_t cap = kHFactorCap;
    char*          buf = static_cast<char*>(scratch(ctx, S_HFAC, size_t(cap) * (sizeof(double) + sizeof(float))));
    if (!buf) return SPHX_ERR_NOMEM;
    if (ctx->hfacNg0 != (long long)ng0)
    {
        std::vector<double> f64(cap);
        std::vector<float>  f32(cap);
        sphx_smoothing_length_factors(ng0, cap, f64.data(), f32.data());
        SPHX_HIP(ctx, hipMemcpyAsync(buf, f64.data(), cap * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        SPHX_HIP(ctx, hipMemcpyAsync(buf + cap * sizeof(double), f32.data(), cap * sizeof(float),
                                     hipMemcpyHostToDevice, ctx->stream));
        SPHX_HIP(ctx, hipStreamSynchronize(ctx->stream)); // the host vectors end here
        ctx->hfacNg0 = ng0;
    }
    out.f64 = reinterpret_cast<const double*>(buf);
    out.f32 = reinterpret_cast<const float*>(buf + cap * sizeof(double));
    out.cap = cap;
    return SPHX_OK;
}

hipEvent_t timingEvent(sphx_ctx* ctx)
{
    if (ctx->eventsUsed == ctx->eventPool.size())
    {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess)
        {
            (void)hipGetLastError();
            return nullptr;
        }
        ctx->eventPool.push_back(e);
    }
    return ctx->eventPool[ctx->eventsUsed++];
}

bool validBox(const sphx_box* b)
{
    if (!b) return false;
    if (!(b->xmax > b->xmin) || !(b->ymax > b->ymin) || !(b->zmax > b->zmin)) return false;
    int bc[3] = {b->bx, b->by, b->bz};
    for (int v : bc)
        if (v < 0 || v > 2) return false;
    return true;
}

// ===========================================================================
// exclusive scan (uint32), three-phase reduce-then-scan with recursion
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS   = 8;
constexpr int SCAN_TILE    = SCAN_THREADS * SCAN_ITEMS; // 2048

// block-wide exclusive scan of one value per thread (256 threads = 4 waves)
__device__ inline uint32_t blockExclusiveScan256(uint32_t v, uint32_t* sWave, uint32_t& total)
{
    unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1)
    {
        uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= unsigned(o)) incl += t;
    }
    if (lane == 63) sWave[wave] = incl;
    __syncthreads();
    uint32_t waveOfs = 0;
    for (unsigned w = 0; w < wave; ++w)
        waveOfs += sWave[w];
    total = sWave[0] + sWave[1] + sWave[2] + sWave[3];
    __syncthreads();
    return waveOfs + incl - v;
}

// logical input has n + 1 entries, in[n] := 0, so out[n] receives the total
__global__ __launch_bounds__(SCAN_THREADS) void scanTiles(const uint32_t* in, uint32_t* out, uint64_t n,
                                                          uint32_t* tileSums)
{
    __shared__ uint32_t sWave[4];
    uint64_t base = uint64_t(blockIdx.x) * SCAN_TILE + uint64_t(threadIdx.x) * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
    {
        uint64_t i = base + k;
        v[k]       = (i < n) ? in[i] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = blockExclusiveScan256(sum, sWave, total);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
    {
        uint64_t i = base + k;
        if (i <= n) out[i] = run;
        run += v[k];
    }
    if (threadIdx.x == 0 && tileSums) tileSums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void scanAddTileOffsets(uint32_t* out, uint64_t n,
                                                                   const uint32_t* tileOfs)
{
    uint32_t add = tileOfs[blockIdx.x];
    uint64_t base = uint64_t(blockIdx.x) * SCAN_TILE;
    for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_THREADS)
    {
        uint64_t i = base + k;
        if (i <= n) out[i] += add;
    }
}

// tiles of one scan level over n values (n + 1 outputs)
constexpr uint64_t scanTilesOf(uint64_t n) { return (n + 1 + SCAN_TILE - 1) / SCAN_TILE; }

// the tile sums of each level in the context's scratch slots
static int scanLevel(sphx_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, int level)
{
    uint64_t numTiles = scanTilesOf(n);
    if (numTiles == 1)
    {
        scanTiles<<<1, SCAN_THREADS, 0, ctx->stream>>>(in, out, n, nullptr);
        SPHX_LAUNCH_CHECK(ctx, "scanTiles");
        return SPHX_OK;
    }
    if (level > 2) return fail(ctx, SPHX_ERR_ARG, "scan too large");
    uint32_t* sums = scratchT<uint32_t>(ctx, Slot(S_SCAN0 + level), numTiles + 1);
    if (!sums) return SPHX_ERR_NOMEM;
    scanTiles<<<unsigned(numTiles), SCAN_THREADS, 0, ctx->stream>>>(in, out, n, sums);
    SPHX_LAUNCH_CHECK(ctx, "scanTiles");
    // exclusive scan of the tile sums in place (numTiles entries, total discarded into sums[numTiles])
    SPHX_TRY(scanLevel(ctx, sums, sums, numTiles, level + 1));
    scanAddTileOffsets<<<unsigned(numTiles), SCAN_THREADS, 0, ctx->stream>>>(out, n, sums);
    SPHX_LAUNCH_CHECK(ctx, "scanAddTileOffsets");
    return SPHX_OK;
}

int exclusiveScanU32(sphx_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n)
{
    return scanLevel(ctx, in, out, n, 0);
}

// ===========================================================================
// LSD radix sort, 8-bit digits.  Per pass: tile histograms -> per-digit scan over the
// tiles -> stable scatter through an LDS-staged local sort so that global writes are
// contiguous runs per digit.  Three launches per pass: the digit-major histogram is
// scanned one column (digit) per block, and each scatter tile adds the exclusive scan
// of the 256 digit totals itself.
constexpr int SORT_THREADS = 256;
#ifndef SPHX_SORT_ITEMS
#define SPHX_SORT_ITEMS 6 // measured: 12 -> 6 cut the sort kernels 0.83 -> 0.78 ms per 256^3 step, 0.21 -> 0.18 at 128^3
#endif
constexpr int SORT_ITEMS   = SPHX_SORT_ITEMS; // keys per thread of a tile
constexpr int SORT_TILE    = SORT_THREADS * SORT_ITEMS; // 1536
constexpr int RADIX        = 256;

// tileVar != nullptr: the tile's OR and AND of its keys go to tileVar[2 tile], [2 tile + 1]
__global__ __launch_bounds__(SORT_THREADS) void radixUpsweep(const uint64_t* __restrict__ keys, uint64_t n,
                                                             int shift, uint32_t* __restrict__ tileHist,
                                                             uint32_t numTiles, uint64_t* __restrict__ tileVar)
{
    __shared__ uint32_t h[4][RADIX];
    __shared__ uint64_t part[2][4];
    unsigned wave = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < 4 * RADIX; d += SORT_THREADS)
        (&h[0][0])[d] = 0;
    __syncthreads();
    uint64_t base = uint64_t(blockIdx.x) * SORT_TILE;
    uint64_t o = 0, a = ~uint64_t(0);
#pragma unroll 4
    for (int k = 0; k < SORT_ITEMS; ++k)
    {
        uint64_t i = base + uint64_t(k) * SORT_THREADS + threadIdx.x;
        if (i < n)
        {
            const uint64_t key = keys[i];
            o |= key, a &= key;
            atomicAdd(&h[wave][unsigned(key >> shift) & 255u], 1u);
        }
    }
    if (tileVar)
    {
        for (int s = 32; s > 0; s >>= 1)
            o |= __shfl_xor(o, s, 64), a &= __shfl_xor(a, s, 64);
        if ((threadIdx.x & 63u) == 0) part[0][wave] = o, part[1][wave] = a;
    }
    __syncthreads();
    unsigned d = threadIdx.x;
    tileHist[uint64_t(d) * numTiles + blockIdx.x] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
    if (tileVar && threadIdx.x == 0)
    {
        tileVar[2 * uint64_t(blockIdx.x)]     = part[0][0] | part[0][1] | part[0][2] | part[0][3];
        tileVar[2 * uint64_t(blockIdx.x) + 1] = part[1][0] & part[1][1] & part[1][2] & part[1][3];
    }
}

// Exclusive scan of each digit's column of tile counts, in place (hist[d * numTiles + t]),
// and the digit's total into hist[RADIX * numTiles + d]; one block per digit.
constexpr int COL_ITEMS = 8;
__global__ __launch_bounds__(256) void radixColumnScan(uint32_t* __restrict__ hist, uint32_t numTiles)
{
    __shared__ uint32_t sWave[4];
    uint32_t*           col   = hist + uint64_t(blockIdx.x) * numTiles;
    uint32_t            carry = 0;
    for (uint32_t base = 0; base < numTiles; base += 256u * COL_ITEMS)
    {
        const uint32_t b = base + threadIdx.x * COL_ITEMS;
        uint32_t       v[COL_ITEMS], sum = 0;
#pragma unroll
        for (int k = 0; k < COL_ITEMS; ++k)
        {
            v[k] = b + k < numTiles ? col[b + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t run = carry + blockExclusiveScan256(sum, sWave, total);
#pragma unroll
        for (int k = 0; k < COL_ITEMS; ++k)
        {
            if (b + k < numTiles) col[b + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) hist[uint64_t(RADIX) * numTiles + blockIdx.x] = carry;
}

// bits that vary across all keys: OR of the tiles' ORs ^ AND of their ANDs (one block)
__global__ __launch_bounds__(256) void radixVaryingBits(const uint64_t* __restrict__ tileVar, uint32_t numTiles,
                                                         uint64_t* __restrict__ varies)
{
    __shared__ uint64_t part[2][4];
    uint64_t            o = 0, a = ~uint64_t(0);
    for (uint32_t t = threadIdx.x; t < numTiles; t += 256)
        o |= tileVar[2 * uint64_t(t)], a &= tileVar[2 * uint64_t(t) + 1];
    for (int s = 32; s > 0; s >>= 1)
        o |= __shfl_xor(o, s, 64), a &= __shfl_xor(a, s, 64);
    if ((threadIdx.x & 63u) == 0) part[0][threadIdx.x >> 6] = o, part[1][threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0)
        varies[0] = (part[0][0] | part[0][1] | part[0][2] | part[0][3]) ^
                    (part[1][0] & part[1][1] & part[1][2] & part[1][3]);
}

template<bool HasVals>
__global__ __launch_bounds__(SORT_THREADS) void radixScatter(const uint64_t* __restrict__ kin,
                                                             const uint32_t* __restrict__ vin,
                                                             uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                             uint64_t n, int shift,
                                                             const uint32_t* __restrict__ tileOfs, uint32_t numTiles)
{
    __shared__ uint64_t sk[SORT_TILE];
    __shared__ uint32_t sv[HasVals ? SORT_TILE : 1];
    __shared__ uint32_t wcount[4][RADIX];
    __shared__ uint32_t dStart[RADIX];
    __shared__ uint32_t gOfs[RADIX];
    __shared__ uint32_t sWave[4];

    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < 4 * RADIX; d += SORT_THREADS)
        (&wcount[0][0])[d] = 0;
    __syncthreads();

    const uint64_t tileBase = uint64_t(blockIdx.x) * SORT_TILE;
    const uint64_t waveBase = tileBase + uint64_t(wave) * 64 * SORT_ITEMS;
    const uint64_t ltMask   = (uint64_t(1) << lane) - 1;

    uint64_t key[SORT_ITEMS];
    uint32_t val[SORT_ITEMS];
    uint32_t rank[SORT_ITEMS];

    // ranks within the wave, in original order (item k, lane l) <-> element waveBase + 64 k + l
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k)
    {
        uint64_t i     = waveBase + uint64_t(k) * 64 + lane;
        bool     valid = i < n;
        key[k]         = valid ? kin[i] : 0;
        if constexpr (HasVals) val[k] = valid ? vin[i] : 0;
        unsigned d     = unsigned(key[k] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            bool     bit = (d >> b) & 1u;
            uint64_t bb  = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        unsigned below = unsigned(__popcll(peers & ltMask));
        unsigned cnt   = unsigned(__popcll(peers));
        uint32_t old   = valid ? wcount[wave][d] : 0u;
        rank[k]        = old + below;
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) wcount[wave][d] = old + cnt;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // per digit: exclusive offsets across waves, tile totals, digit starts within the tile
    {
        unsigned d   = threadIdx.x;
        uint32_t c0  = wcount[0][d], c1 = wcount[1][d], c2 = wcount[2][d], c3 = wcount[3][d];
        uint32_t tot = c0 + c1 + c2 + c3;
        __syncthreads();
        wcount[0][d] = 0;
        wcount[1][d] = c0;
        wcount[2][d] = c0 + c1;
        wcount[3][d] = c0 + c1 + c2;
        uint32_t total, all;
        uint32_t start = blockExclusiveScan256(tot, sWave, total);
        // the digit's first output position: the totals of the smaller digits plus the
        // counts of this digit in the tiles before this one
        uint32_t below = blockExclusiveScan256(tileOfs[uint64_t(RADIX) * numTiles + d], sWave, all);
        dStart[d]      = start;
        gOfs[d]        = below + tileOfs[uint64_t(d) * numTiles + blockIdx.x];
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k)
    {
        uint64_t i = waveBase + uint64_t(k) * 64 + lane;
        if (i < n)
        {
            unsigned d   = unsigned(key[k] >> shift) & 255u;
            uint32_t pos = dStart[d] + wcount[wave][d] + rank[k];
            sk[pos]      = key[k];
            if constexpr (HasVals) sv[pos] = val[k];
        }
    }
    __syncthreads();

    uint64_t tileN = (n - tileBase) < uint64_t(SORT_TILE) ? (n - tileBase) : uint64_t(SORT_TILE);
    for (unsigned i = threadIdx.x; i < tileN; i += SORT_THREADS)
    {
        uint64_t k   = sk[i];
        unsigned d   = unsigned(k >> shift) & 255u;
        uint64_t out = uint64_t(gOfs[d]) + (i - dStart[d]);
        kout[out]    = k;
        if constexpr (HasVals) vout[out] = sv[i];
    }
}

__global__ void copyKV(const uint64_t* ks, const uint32_t* vs, uint64_t* kd, uint32_t* vd, uint64_t n)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    {
        kd[i] = ks[i];
        if (vs) vd[i] = vs[i];
    }
}

// Work buffers of one sort: ping-pong keys and values (n each) and the per-tile digit
// histograms followed by the digit totals (RADIX * (tiles + 1)).
struct SortWork
{
    uint64_t* kAlt;
    uint32_t* vAlt;
    uint32_t* hist;
};

// One pass over the 8-bit digit at `shift`: (kSrc, vSrc) -> (kDst, vDst), stable.
// tileVar != nullptr: the histogram kernel also writes each tile's OR and AND of the keys.
static int radixPass(sphx_ctx* ctx, sphx_ctx* tctx, const uint64_t* kSrc, const uint32_t* vSrc, uint64_t* kDst,
                     uint32_t* vDst, uint64_t n, int shift, uint32_t* hist, uint32_t numTiles, uint64_t* tileVar)
{
    {
        KTimer kt_(tctx, "radixUpsweep");
        radixUpsweep<<<numTiles, SORT_THREADS, 0, ctx->stream>>>(kSrc, n, shift, hist, numTiles, tileVar);
    }
    SPHX_LAUNCH_CHECK(ctx, "radixUpsweep");
    radixColumnScan<<<RADIX, 256, 0, ctx->stream>>>(hist, numTiles);
    SPHX_LAUNCH_CHECK(ctx, "radixColumnScan");
    {
        KTimer kt_(tctx, "radixScatter");
        if (vSrc)
            radixScatter<true><<<numTiles, SORT_THREADS, 0, ctx->stream>>>(kSrc, vSrc, kDst, vDst, n, shift, hist,
                                                                          numTiles);
        else
            radixScatter<false><<<numTiles, SORT_THREADS, 0, ctx->stream>>>(kSrc, nullptr, kDst, nullptr, n, shift,
                                                                           hist, numTiles);
    }
    SPHX_LAUNCH_CHECK(ctx, "radixScatter");
    return SPHX_OK;
}

// own == nullptr: buffers from the context's scratch, and the passes over digits that are
// constant across all keys are skipped: the first pass's histogram kernel also forms each
// tile's OR and AND of the keys, one block reduces them to the bits that vary, and the
// host reads those 8 bytes while the first pass's scan and scatter run (measured: skipping
// no pass costs the ~2M-entry tree sorts more in extra passes than the read).  own !=
// nullptr: the caller's buffers, every pass of keyBits runs, and the sort is stream-ordered
// throughout (no host synchronisation, no memset, no timing events).  Both give the same
// result: a pass over a constant digit is the identity permutation.
static int radixSortWith(sphx_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n, int keyBits,
                         const SortWork* own, int lowBits = 0, bool checkDigits = true)
{
    if (n <= 1) return SPHX_OK;
    if (n >= (uint64_t(1) << 32)) return argError(ctx, "sort: n must be < 2^32");
    uint32_t numTiles = uint32_t((n + SORT_TILE - 1) / SORT_TILE);

    uint64_t* kAlt;
    uint32_t *vAlt, *hist;
    uint64_t  varies = ~uint64_t(0);
    if (own)
    {
        kAlt = own->kAlt, vAlt = vals ? own->vAlt : nullptr, hist = own->hist;
    }
    else
    {
        kAlt = scratchT<uint64_t>(ctx, S_SORT_KEYS, n);
        vAlt = vals ? scratchT<uint32_t>(ctx, S_SORT_VALS, n) : nullptr;
        hist = scratchT<uint32_t>(ctx, S_SORT_HIST, uint64_t(RADIX) * (numTiles + 1));
        if (!kAlt || (vals && !vAlt) || !hist) return SPHX_ERR_NOMEM;
    }
    uint64_t *tileVar = nullptr, *vary = nullptr;
    if (!own && checkDigits && keyBits - 8 * (lowBits / 8) > 8)
    {
        tileVar = scratchT<uint64_t>(ctx, S_SORT_VAR, 2 * uint64_t(numTiles));
        vary    = scratchT<uint64_t>(ctx, S_SMALL, 8);
        if (!tileVar || !vary) return SPHX_ERR_NOMEM;
    }
    if (!kAlt || (vals && !vAlt) || !hist) return argError(ctx, "sort: missing work buffer");
    sphx_ctx* tctx = own ? nullptr : ctx; // KTimer records nothing for the caller-buffer variant

    uint64_t* kSrc = keys;
    uint32_t* vSrc = vals;
    uint64_t* kDst = kAlt;
    uint32_t* vDst = vAlt;
    bool pending = false; // the varying bits are being read
    for (int shift = 8 * (lowBits / 8); shift < keyBits; shift += 8)
    {
        if (pending)
        {
            SPHX_TRY(ctx->hostStage ? readBackFinish(ctx, &varies, sizeof(varies))
                                    : readBack(ctx, vary, &varies, sizeof(varies)));
            pending = false;
        }
        if (((varies >> shift) & 0xFFull) == 0) continue; // digit constant across all keys
        const bool first = shift == 8 * (lowBits / 8) && tileVar;
        SPHX_TRY(radixPass(ctx, tctx, kSrc, vSrc, kDst, vDst, n, shift, hist, numTiles, first ? tileVar : nullptr));
        if (first)
        {
            radixVaryingBits<<<1, 256, 0, ctx->stream>>>(tileVar, numTiles, vary);
            SPHX_LAUNCH_CHECK(ctx, "radixVaryingBits");
            if (ctx->hostStage) SPHX_TRY(readBackStart(ctx, vary, sizeof(uint64_t)));
            pending = true;
        }
        std::swap(kSrc, kDst);
        std::swap(vSrc, vDst);
    }
    if (kSrc != keys)
    {
        unsigned g = std::min<unsigned>(gridFor(n, 256), unsigned(ctx->numCU) * 16);
        copyKV<<<g, 256, 0, ctx->stream>>>(kSrc, vSrc, keys, vals, n);
        SPHX_LAUNCH_CHECK(ctx, "copyKV");
    }
    return SPHX_OK;
}

int radixSortPairs(sphx_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n, int keyBits, int lowBits,
                   bool checkDigits)
{
    return radixSortWith(ctx, keys, vals, n, keyBits, nullptr, std::max(0, std::min(lowBits, keyBits)), checkDigits);
}

// ---------------------------------------------------------------------------
// Key sort by a top prefix and a run fix-up.  Four passes sort by bits 32..63 (or, when the
// context's previous prefix sort found every run of equal top 24 bits short, three passes
// by bits 40..63 through two pairs of work arrays), ending in the caller's arrays; keys
// with an equal prefix then form runs that hold their input order.  When no run is
// longer than kRunMax, each run is sorted by the bits below the prefix, stably: the
// result is the stable sort by the whole key, as eight passes give it.  Particle keys of
// a space-filling curve share their top 32 bits (a cell 2^-10.67 of the box wide), or top
// 24 bits (2^-7.67 of the box), only with a few neighbours, so the fix-up is short; when
// some run is longer (keys packed far below that resolution, small integer keys), the
// eight passes run on the arrays as the prefix passes left them, which gives the same
// result (a stable sort of a permutation whose equal keys keep their input order): after
// four prefix passes the next kPrefixRetryCalls sorts of the context run them directly,
// after three the next sort takes four.
constexpr uint32_t kRunMax           = 64;
constexpr uint64_t kPrefixSortMinN   = uint64_t(1) << 18;
constexpr int      kPrefixRetryCalls = 16;
constexpr uint32_t kFixOwned = 192; // positions written per wave
constexpr uint32_t kFixWin   = 256; // the wave's window: 4 keys per lane


// flag bit 0: a run of keys with equal bits sh..63 is longer than kRunMax; bit 1: the same
// for bits 40..63 (keys ordered by bits sh..63, sh <= 40: a run that long has equal
// prefixes kRunMax positions apart)
__global__ __launch_bounds__(256) void runTooLong(const uint64_t* __restrict__ k, uint32_t n, uint32_t sh,
                                                  uint32_t* __restrict__ flag)
{
    __shared__ uint32_t any[4];
    uint32_t            f = 0;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i + kRunMax < n; i += gridDim.x * 256u)
    {
        const uint64_t a = k[i], b = k[i + kRunMax];
        f |= ((a >> sh) == (b >> sh) ? 1u : 0u) | ((a >> 40) == (b >> 40) ? 2u : 0u);
    }
    const uint32_t w = (__ballot(f & 1u) != 0 ? 1u : 0u) | (__ballot(f & 2u) != 0 ? 2u : 0u);
    if ((threadIdx.x & 63u) == 0) any[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0 && (any[0] | any[1] | any[2] | any[3])) atomicOr(flag, any[0] | any[1] | any[2] | any[3]);
}

// exchanges (ka, va) and (kb, vb) when both are in one run and ka is larger (within a run
// the prefixes are equal, so the bits below decide; strictly: equal keys keep their order)
__device__ inline bool runSwap(bool same, uint64_t& ka, uint32_t& va, uint64_t& kb, uint32_t& vb)
{
    if (!(same && ka > kb)) return false;
    const uint64_t tk = ka;
    const uint32_t tv = va;
    ka = kb, va = vb, kb = tk, vb = tv;
    return true;
}

// Runs are keys with equal bits sh..63.  In place.  One wave per kFixOwned positions
// [t0, t0 + kFixOwned): its window [t0, t0 + kFixWin) holds every run that starts among
// them whole (runs are at most kRunMax long).  Element e = 4 lane + s of the window sits
// at t0 + e.  Odd-even transposition rounds inside the runs repeat until a pair of rounds
// swaps nothing (a run of length L is sorted after L rounds); the wave then writes the
// keys that moved in the runs that start in [t0, t0 + kFixOwned).  The run that covers t0
// from an earlier start belongs to the previous wave, whose window reaches
// t0 + kFixWin - kFixOwned - 1 >= its end.  A wave writes only its own runs' positions;
// of its neighbours' positions it uses only the prefixes, which the fix-up leaves
// unchanged at every position.
__global__ __launch_bounds__(256) void runFix(uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t sh)
{
    static_assert(kFixWin == 4 * 64 && kFixWin - kFixOwned >= kRunMax, "a window holds its owned runs whole");
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t t0   = uint64_t(blockIdx.x * 4u + (threadIdx.x >> 6)) * kFixOwned;
    if (t0 >= n) return;
    uint64_t k[4], k0[4];
    uint32_t v[4], v0[4];
    bool     st[4]; // the element starts a run (padding beyond n: one run each)
#pragma unroll
    for (int s = 0; s < 4; ++s)
    {
        const uint64_t p = t0 + 4u * lane + uint32_t(s);
        const bool     in = p < n;
        k[s] = k0[s]      = in ? keys[p] : ~uint64_t(0);
        v[s] = v0[s]      = in ? vals[p] : 0u;
        st[s]             = !in;
    }
    // the predecessor of element 0 is t0 - 1 (none for t0 = 0), of 4 lane the previous lane's 3
    const uint64_t before = t0 > 0 ? keys[t0 - 1] : 0;
    const uint64_t prev3  = __shfl_up(k[3], 1, 64);
    st[0] = st[0] || (lane == 0 ? (t0 == 0 || (before >> sh) != (k[0] >> sh)) : (prev3 >> sh) != (k[0] >> sh));
#pragma unroll
    for (int s = 1; s < 4; ++s)
        st[s] = st[s] || (k[s - 1] >> sh) != (k[s] >> sh);
    // pairs inside one run: (4l, 4l+1), (4l+2, 4l+3), (4l+1, 4l+2), (4l+3, 4l+4)
    const bool s01 = !st[1], s23 = !st[3], s12 = !st[2];
    const bool nx0 = __shfl_down(st[0] ? 1 : 0, 1, 64) != 0;
    const bool s30 = lane < 63 && !nx0;
    for (;;)
    {
        bool sw = runSwap(s01, k[0], v[0], k[1], v[1]);
        sw      = runSwap(s23, k[2], v[2], k[3], v[3]) || sw;
        sw      = runSwap(s12, k[1], v[1], k[2], v[2]) || sw;
        // across lanes: lane l's element 3 with lane l + 1's element 0 (both decide alike)
        const uint64_t kn = __shfl_down(k[0], 1, 64);
        const uint32_t vn = __shfl_down(v[0], 1, 64);
        const uint64_t kp = __shfl_up(k[3], 1, 64);
        const uint32_t vp = __shfl_up(v[3], 1, 64);
        const bool     sp = __shfl_up(s30 ? 1 : 0, 1, 64) != 0 && lane > 0;
        const bool     down = s30 && k[3] > kn;
        const bool     up   = sp && kp > k[0];
        if (down) k[3] = kn, v[3] = vn;
        if (up) k[0] = kp, v[0] = vp;
        sw = sw || down;
        if (__ballot(sw) == 0) break;
    }
    // owned: runs starting at window positions [f, g), f = the first start, g = the first
    // start at or after kFixOwned
    uint32_t f = kFixWin, g = kFixWin;
#pragma unroll
    for (int s = 0; s < 4; ++s)
    {
        const uint64_t b  = __ballot(st[s]);
        const uint64_t bg = b & (~uint64_t(0) << (kFixOwned / 4));
        if (b) f = min(f, 4u * uint32_t(__builtin_ctzll(b)) + uint32_t(s));
        if (bg) g = min(g, 4u * uint32_t(__builtin_ctzll(bg)) + uint32_t(s));
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
    {
        const uint32_t e = 4u * lane + uint32_t(s);
        const uint64_t p = t0 + e;
        if (e >= f && e < g && p < n && (k[s] != k0[s] || v[s] != v0[s])) keys[p] = k[s], vals[p] = v[s];
    }
}

static int prefixSortPairs(sphx_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n)
{
    const uint32_t numTiles = uint32_t((n + SORT_TILE - 1) / SORT_TILE);
    uint64_t*      kA       = scratchT<uint64_t>(ctx, S_SORT_KEYS, n);
    uint32_t*      vA       = scratchT<uint32_t>(ctx, S_SORT_VALS, n);
    uint32_t*      hist     = scratchT<uint32_t>(ctx, S_SORT_HIST, uint64_t(RADIX) * (numTiles + 1));
    uint32_t*      flag     = scratchT<uint32_t>(ctx, S_SMALL, 4);
    if (!kA || !vA || !hist || !flag) return SPHX_ERR_NOMEM;
    const bool     p24 = ctx->sortPrefix24;
    const uint32_t sh  = p24 ? 40u : 32u;
    if (p24)
    {
        uint64_t* kB = scratchT<uint64_t>(ctx, S_SORT_KEYS2, n);
        uint32_t* vB = scratchT<uint32_t>(ctx, S_SORT_VALS2, n);
        if (!kB || !vB) return SPHX_ERR_NOMEM;
        SPHX_TRY(radixPass(ctx, ctx, keys, vals, kA, vA, n, 40, hist, numTiles, nullptr));
        SPHX_TRY(radixPass(ctx, ctx, kA, vA, kB, vB, n, 48, hist, numTiles, nullptr));
        SPHX_TRY(radixPass(ctx, ctx, kB, vB, keys, vals, n, 56, hist, numTiles, nullptr));
    }
    else
    {
        SPHX_TRY(radixPass(ctx, ctx, keys, vals, kA, vA, n, 32, hist, numTiles, nullptr));
        SPHX_TRY(radixPass(ctx, ctx, kA, vA, keys, vals, n, 40, hist, numTiles, nullptr));
        SPHX_TRY(radixPass(ctx, ctx, keys, vals, kA, vA, n, 48, hist, numTiles, nullptr));
        SPHX_TRY(radixPass(ctx, ctx, kA, vA, keys, vals, n, 56, hist, numTiles, nullptr));
    }
    SPHX_HIP(ctx, hipMemsetAsync(flag, 0, sizeof(uint32_t), ctx->stream));
    runTooLong<<<std::min<unsigned>(gridFor(n, 256), unsigned(ctx->numCU) * 8), 256, 0, ctx->stream>>>(
        keys, uint32_t(n), sh, flag);
    SPHX_LAUNCH_CHECK(ctx, "runTooLong");
    uint32_t tooLong = 0;
    SPHX_TRY(readBack(ctx, flag, &tooLong, sizeof(tooLong)));
    // the next prefix sort takes three passes when every run of equal top 24 bits is short
    ctx->sortPrefix24 = (tooLong & 2u) == 0;
    if (tooLong & 1u)
    {
        if (!p24) ctx->sortPrefixSkip = kPrefixRetryCalls;
        return radixSortWith(ctx, keys, vals, n, 64, nullptr);
    }
    const uint64_t waves = (n + kFixOwned - 1) / kFixOwned;
    {
        KTimer kt_(ctx, "runFix");
        runFix<<<unsigned((waves + 3) / 4), 256, 0, ctx->stream>>>(keys, vals, uint32_t(n), sh);
    }
    SPHX_LAUNCH_CHECK(ctx, "runFix");
    return SPHX_OK;
}

// sphx_sort_by_key: the prefix sort for large arrays (unless a recent one fell back)
static int sortKeysValues(sphx_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n)
{
    if (n < kPrefixSortMinN || n >= (uint64_t(1) << 31)) return radixSortPairs(ctx, keys, vals, n, 64);
    if (ctx->sortPrefixSkip > 0)
    {
        --ctx->sortPrefixSkip;
        return radixSortPairs(ctx, keys, vals, n, 64);
    }
    return prefixSortPairs(ctx, keys, vals, n);
}

// ---------------------------------------------------------------------------
// Allocation-free sort (sphx_sort_by_key_noalloc).  Its work buffers are one file-scope
// device array, `scratch`, of which every GPU holds its own copy for the life of the
// process.  The first sphx_ctx_create on a device looks up that device's address with
// hipGetSymbolAddress (the device current) and every context on the device keeps it;
// host code uses only that address.  Nothing in the array needs to start at zero: each
// pass writes every histogram entry and every ping-pong element it later reads.  Calls
// that overlap on one GPU share the array and race.
constexpr uint64_t kNaMaxN  = uint64_t(1) << 21; // 2,097,152 keys (a 128^3 share)
constexpr uint64_t kNaTiles = (kNaMaxN + SORT_TILE - 1) / SORT_TILE;
constexpr uint64_t kNaHistN = uint64_t(RADIX) * (kNaTiles + 1); // tile histograms + digit totals
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kNaKeysOff = 0;
constexpr size_t kNaValsOff = kNaKeysOff + align256(kNaMaxN * sizeof(uint64_t));
constexpr size_t kNaHistOff = kNaValsOff + align256(kNaMaxN * sizeof(uint32_t));
constexpr size_t kNaBytes   = kNaHistOff + align256(kNaHistN * sizeof(uint32_t)); // ~25.3 MiB

namespace noalloc
{
static __device__ __align__(256) unsigned char scratch[kNaBytes];
} // namespace noalloc

constexpr int      kNaMaxDevices = 64;
static std::mutex  naMutex;
static bool        naLooked[kNaMaxDevices];
static void*       naAddr[kNaMaxDevices];
static hipError_t  naStatus[kNaMaxDevices];

// the device's address of `scratch`, looked up once per device (device current)
static void* noallocScratchOf(int device)
{
    if (device < 0 || device >= kNaMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(naMutex);
    if (!naLooked[device])
    {
        void*      ptr = nullptr;
        hipError_t e   = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(noalloc::scratch));
        if (e != hipSuccess) (void)hipGetLastError();
        naStatus[device] = e;
        naAddr[device]   = e == hipSuccess ? ptr : nullptr;
        naLooked[device] = true;
    }
    return naAddr[device];
}

// ===========================================================================
// gather / sequence
template<class T>
__global__ __launch_bounds__(256) void gatherKernel(const uint32_t* __restrict__ order, const T* __restrict__ in,
                                                    T* __restrict__ out, uint64_t n)
{
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[order[i]];
}

// several arrays permuted by one ordering: order[i] is read once, every array's element
// is loaded before any is stored (the loads overlap in flight)
struct GatherSet
{
    const void* in[SPHX_GATHER_MAX_ARRAYS];
    void*       out[SPHX_GATHER_MAX_ARRAYS];
    uint32_t    wide; // bit k: array k has 8-byte elements, else 4-byte
    int         count;
};

__global__ __launch_bounds__(256) void gatherArraysKernel(const uint32_t* __restrict__ order, GatherSet s, uint64_t n)
{
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = order[i];
    uint64_t       v[SPHX_GATHER_MAX_ARRAYS];
#pragma unroll
    for (int k = 0; k < SPHX_GATHER_MAX_ARRAYS; ++k)
        if (k < s.count)
            v[k] = ((s.wide >> k) & 1u) ? static_cast<const uint64_t*>(s.in[k])[j]
                                        : uint64_t(static_cast<const uint32_t*>(s.in[k])[j]);
#pragma unroll
    for (int k = 0; k < SPHX_GATHER_MAX_ARRAYS; ++k)
        if (k < s.count)
        {
            if ((s.wide >> k) & 1u) static_cast<uint64_t*>(s.out[k])[i] = v[k];
            else static_cast<uint32_t*>(s.out[k])[i] = uint32_t(v[k]);
        }
}

__global__ __launch_bounds__(256) void sequenceKernel(uint32_t* v, uint64_t n, uint32_t first)
{
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = first + uint32_t(i);
}

// ===========================================================================
// Hilbert keys: sfc3D (domain/include/cstone/sfc/sfc.hpp:141-178) with separately
// rounded products, as in the reference expression int(floor(x*mx) - xmin*mx)
// The keys are encoded two levels per lookup: every workgroup first tabulates, in LDS,
// hilbertStep applied twice for each of the 48 states and each pair of levels' six raw
// bits (digits and next state), and the top level from the initial state; a key is then
// one top-level lookup and ten pair lookups (grid-stride over the particles).
__global__ __launch_bounds__(256) void sfcKeysKernel(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, uint64_t* __restrict__ keys,
                                                     uint64_t n, DBox box)
{
#pragma clang fp contract(off)
    static_assert(kMaxLevel == 21, "one top level and ten pairs of levels");
    __shared__ uint16_t pairTab[48 * 64]; // [state][x1x0 y1y0 z1z0] = digits (6 bits) | next state << 6
    __shared__ uint16_t topTab[8];        // [x y z] from state 0 = digit | next state << 3
    for (uint32_t e = threadIdx.x; e < 48u * 64u; e += 256u)
    {
        uint32_t       s  = e >> 6;
        const uint32_t xb = (e >> 4) & 3u, yb = (e >> 2) & 3u, zb = e & 3u;
        const uint32_t d1 = hilbertStep(s, ((xb >> 1) << 2) | ((yb >> 1) << 1) | (zb >> 1));
        const uint32_t d0 = hilbertStep(s, ((xb & 1u) << 2) | ((yb & 1u) << 1) | (zb & 1u));
        pairTab[e]        = uint16_t((d1 << 3) | d0 | (s << 6));
    }
    if (threadIdx.x < 8)
    {
        uint32_t       s = 0;
        const uint32_t d = hilbertStep(s, threadIdx.x);
        topTab[threadIdx.x] = uint16_t(d | (s << 3));
    }
    __syncthreads();

    const double cube   = double(1u << kMaxLevel);
    const int    mcoord = (1 << kMaxLevel) - 1;
    const double mx = cube * box.ilen[0], my = cube * box.ilen[1], mz = cube * box.ilen[2];
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256)
    {
        if (keys[i] == kRemoveKey) continue;
        int ix = int(floor(x[i] * mx) - box.lo[0] * mx);
        int iy = int(floor(y[i] * my) - box.lo[1] * my);
        int iz = int(floor(z[i] * mz) - box.lo[2] * mz);
        ix = min(ix, mcoord);
        iy = min(iy, mcoord);
        iz = min(iz, mcoord);
        ix = max(ix, 0); // coordinates below the box floor would be undefined in the reference
        iy = max(iy, 0);
        iz = max(iz, 0);
        const uint32_t ux = uint32_t(ix), uy = uint32_t(iy), uz = uint32_t(iz);
        const uint32_t t  = topTab[((ux >> 20) << 2) | ((uy >> 20) << 1) | (uz >> 20)];
        uint64_t       key = t & 7u;
        uint32_t       s   = t >> 3;
#pragma unroll
        for (int l = 18; l >= 0; l -= 2)
        {
            const uint32_t idx = (((ux >> l) & 3u) << 4) | (((uy >> l) & 3u) << 2) | ((uz >> l) & 3u);
            const uint32_t e   = pairTab[(s << 6) | idx];
            key                = (key << 6) | (e & 63u);
            s                  = e >> 6;
        }
        keys[i] = key;
    }
}

template<class T>
int gatherT(sphx_ctx* ctx, const uint32_t* order, const T* in, T* out, uint64_t n)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (n == 0) return SPHX_OK;
    if (!order || !in || !out) return argError(ctx, "gather: null pointer");
    if (static_cast<const void*>(in) == static_cast<const void*>(out)) return argError(ctx, "gather: in == out");
    {
        KTimer kt_(ctx, "gatherKernel");
        gatherKernel<T><<<gridFor(n, 256), 256, 0, ctx->stream>>>(order, in, out, n);
    }
    SPHX_LAUNCH_CHECK(ctx, "gatherKernel");
    return SPHX_OK;
}

static int gatherArrays(sphx_ctx* ctx, const uint32_t* order, uint64_t n, int32_t count, const void* const* in,
                        void* const* out, const uint32_t* elemBytes)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (count < 0 || count > SPHX_GATHER_MAX_ARRAYS)
        return fail(ctx, SPHX_ERR_ARG, "gather_arrays: count must be in [0, " + std::to_string(SPHX_GATHER_MAX_ARRAYS) + "]");
    if (n == 0 || count == 0) return SPHX_OK;
    if (!order || !in || !out || !elemBytes) return argError(ctx, "gather_arrays: null pointer");
    GatherSet s{};
    s.count = count;
    for (int k = 0; k < count; ++k)
    {
        if (!in[k] || !out[k]) return fail(ctx, SPHX_ERR_ARG, "gather_arrays: null array " + std::to_string(k));
        if (elemBytes[k] != 4 && elemBytes[k] != 8)
            return fail(ctx, SPHX_ERR_ARG, "gather_arrays: element size of array " + std::to_string(k) + " must be 4 or 8");
        s.in[k]  = in[k];
        s.out[k] = out[k];
        if (elemBytes[k] == 8) s.wide |= 1u << k;
    }
    // every destination is written while all sources are read: no destination may overlap
    // a source or another destination
    auto span = [&](const void* p, int k) {
        auto b = reinterpret_cast<uintptr_t>(p);
        return std::pair<uintptr_t, uintptr_t>(b, b + n * elemBytes[k]);
    };
    for (int a = 0; a < count; ++a)
        for (int b = 0; b < count; ++b)
        {
            auto o = span(out[a], a), i = span(in[b], b), o2 = span(out[b], b);
            if (o.first < i.second && i.first < o.second)
                return fail(ctx, SPHX_ERR_ARG, "gather_arrays: output " + std::to_string(a) + " overlaps input " +
                                         std::to_string(b));
            if (a != b && o.first < o2.second && o2.first < o.second)
                return fail(ctx, SPHX_ERR_ARG, "gather_arrays: outputs " + std::to_string(a) + " and " + std::to_string(b) +
                                         " overlap");
        }
    {
        KTimer kt_(ctx, "gatherArraysKernel");
        gatherArraysKernel<<<gridFor(n, 256), 256, 0, ctx->stream>>>(order, s, n);
    }
    SPHX_LAUNCH_CHECK(ctx, "gatherArraysKernel");
    return SPHX_OK;
}

// coordinate extrema for makeGlobalBox: per block min and max of x, y, z (6 values)
__global__ __launch_bounds__(256) void boxExtremaKernel(const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z, uint64_t n,
                                                        double* __restrict__ partial)
{
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    {
        double v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            lo[d] = v[d] < lo[d] ? v[d] : lo[d]; // std::min / std::max order (MinMax, box_mpi.hpp:34-52)
            hi[d] = hi[d] < v[d] ? v[d] : hi[d];
        }
    }
    __shared__ double red[6][4];
    const unsigned    lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        double a = waveMin(lo[d]), b = waveMax(hi[d]);
        if (lane == 0) red[d][w] = a, red[3 + d][w] = b;
    }
    __syncthreads();
    if (threadIdx.x < 6)
    {
        int    k = threadIdx.x;
        double v = red[k][0];
        for (int q = 1; q < 4; ++q)
            v = k < 3 ? (red[k][q] < v ? red[k][q] : v) : (v < red[k][q] ? red[k][q] : v);
        partial[size_t(blockIdx.x) * 6 + k] = v;
    }
}

__global__ __launch_bounds__(64) void boxExtremaFinal(const double* __restrict__ partial, uint32_t blocks,
                                                      double* __restrict__ out)
{
    const unsigned k = threadIdx.x;
    if (k >= 6) return;
    double v = k < 3 ? INFINITY : -INFINITY;
    for (uint32_t b = 0; b < blocks; ++b)
    {
        double p = partial[size_t(b) * 6 + k];
        v        = k < 3 ? (p < v ? p : v) : (v < p ? p : v);
    }
    out[k] = v;
}

} // namespace sphx

using namespace sphx;

// ===========================================================================
// C ABI
extern "C"
{

int sphx_smoothing_length_factors(uint32_t ng0, uint32_t count, double* f64, float* f32)
{
#pragma clang fp contract(off)
    // updateH<T> (sph/kernels.hpp:27-33): T(1) + c0 * ng0 / T(nc) with c0 = 1023 and
    // exponent T(1.0 / 10.0), evaluated in T; std::pow(float, float) is powf
    const double c0 = 1023.0, e = 1.0 / 10.0;
    const float  c0f = 1023.0f, ef = float(1.0 / 10.0);
    for (uint32_t nc = 0; nc < count; ++nc)
    {
        if (f64) f64[nc] = std::pow(1.0 + c0 * double(ng0) / double(nc), e);
        if (f32) f32[nc] = std::pow(1.0f + c0f * float(ng0) / float(nc), ef);
    }
    return SPHX_OK;
}


const char* sphx_version(void) { return "sphx 0.1 gfx950"; }

const char* sphx_last_error(const sphx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sphx_ctx_create(int device, void* stream, sphx_ctx** out)
{
    if (!out) return SPHX_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    {
        (void)hipGetLastError();
        return SPHX_ERR_NODEV;
    }
    if (device < 0 || device >= n) return SPHX_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return SPHX_ERR_HIP;
    auto* ctx   = new sphx_ctx;
    ctx->device = device;
    ctx->stream = static_cast<hipStream_t>(stream);
    ctx->noallocScratch = noallocScratchOf(device);
    // pinned staging for small reads; without it the reads go to pageable memory
    if (hipHostMalloc(&ctx->hostStage, kHostStageBytes, hipHostMallocDefault) != hipSuccess)
    {
        (void)hipGetLastError();
        ctx->hostStage = nullptr;
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&ctx->minDtHost), 64, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->minDtEvent, hipEventDisableTiming) != hipSuccess)
    {
        (void)hipGetLastError(); // no deferred time step on this context
        if (ctx->minDtHost) (void)hipHostFree(ctx->minDtHost);
        ctx->minDtHost  = nullptr;
        ctx->minDtEvent = nullptr;
    }
    if (ctx->hostStage && hipEventCreateWithFlags(&ctx->readEvent, hipEventDisableTiming) != hipSuccess)
    {
        (void)hipGetLastError();
        (void)hipHostFree(ctx->hostStage);
        ctx->hostStage = nullptr;
        ctx->readEvent = nullptr;
    }
    int cu      = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0)
        ctx->numCU = cu;
    *out = ctx;
    return SPHX_OK;
}

int sphx_timing_summary(sphx_ctx* ctx, int cap, const char** names, int* launches, double* totalMs,
                        int* numEntries)
{
    if (!ctx || !numEntries) return SPHX_ERR_ARG;
    SPHX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<const char*> nm;
    std::vector<int>         cnt;
    std::vector<double>      ms;
    for (auto& t : ctx->timed)
    {
        float e = 0.f;
        SPHX_HIP(ctx, hipEventElapsedTime(&e, t.a, t.b));
        size_t k = 0;
        while (k < nm.size() && std::strcmp(nm[k], t.name) != 0)
            ++k;
        if (k == nm.size())
        {
            nm.push_back(t.name);
            cnt.push_back(0);
            ms.push_back(0.0);
        }
        cnt[k] += 1;
        ms[k] += double(e);
    }
    ctx->timed.clear();
    ctx->eventsUsed = 0;
    *numEntries     = int(nm.size());
    for (int k = 0; k < int(nm.size()) && k < cap; ++k)
    {
        if (names) names[k] = nm[k];
        if (launches) launches[k] = cnt[k];
        if (totalMs) totalMs[k] = ms[k];
    }
    return SPHX_OK;
}

int sphx_ctx_destroy(sphx_ctx* ctx)
{
    if (!ctx) return SPHX_OK;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto e : ctx->eventPool)
        (void)hipEventDestroy(e);
    for (auto& b : ctx->buf)
        if (b.p) (void)hipFree(b.p);
    if (ctx->hostStage) (void)hipHostFree(ctx->hostStage);
    if (ctx->readEvent) (void)hipEventDestroy(ctx->readEvent);
    if (ctx->minDtHost) (void)hipHostFree(ctx->minDtHost);
    if (ctx->minDtEvent) (void)hipEventDestroy(ctx->minDtEvent);
    delete ctx;
    return SPHX_OK;
}

int sphx_ctx_set_stream(sphx_ctx* ctx, void* stream)
{
    if (!ctx) return SPHX_ERR_ARG;
    double dt; // a deferred time step is read before its stream is left
    SPHX_TRY(sphx_min_dt_courant(ctx, &dt));
    ctx->stream = static_cast<hipStream_t>(stream);
    return SPHX_OK;
}

int sphx_ctx_synchronize(sphx_ctx* ctx)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SPHX_OK;
}

int sphx_compute_sfc_keys(sphx_ctx* ctx, const double* x, const double* y, const double* z, uint64_t* keys,
                          uint64_t n, const sphx_box* box)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (n == 0) return SPHX_OK;
    if (!x || !y || !z || !keys) return argError(ctx, "compute_sfc_keys: null pointer");
    if (!validBox(box)) return argError(ctx, "compute_sfc_keys: invalid box");
    {
        KTimer kt_(ctx, "sfcKeysKernel");
        const unsigned g = std::min<unsigned>(gridFor(n, 256), unsigned(ctx->numCU) * 4);
        sfcKeysKernel<<<g, 256, 0, ctx->stream>>>(x, y, z, keys, n, makeDBox(box));
    }
    SPHX_LAUNCH_CHECK(ctx, "sfcKeysKernel");
    return SPHX_OK;
}

int sphx_sort_by_key(sphx_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (n == 0) return SPHX_OK;
    if (!keys || !values) return argError(ctx, "sort_by_key: null pointer");
    return sortKeysValues(ctx, keys, values, n);
}

int sphx_sort_by_key_noalloc(sphx_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (n > kNaMaxN)
        return fail(ctx, SPHX_ERR_ARG, "sort_by_key_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
                                           std::to_string(kNaMaxN) + " keys");
    if (n == 0) return SPHX_OK;
    if (!keys || !values) return argError(ctx, "sort_by_key_noalloc: null pointer");
    if (!ctx->noallocScratch)
        return fail(ctx, SPHX_ERR_HIP, "sort_by_key_noalloc: no device address for the work array "
                                       "(hipGetSymbolAddress failed when the context was created)");
    char*    base = static_cast<char*>(ctx->noallocScratch);
    SortWork w;
    w.kAlt        = reinterpret_cast<uint64_t*>(base + kNaKeysOff);
    w.vAlt        = reinterpret_cast<uint32_t*>(base + kNaValsOff);
    w.hist        = reinterpret_cast<uint32_t*>(base + kNaHistOff);
    return radixSortWith(ctx, keys, values, n, 64, &w);
}

int sphx_sequence(sphx_ctx* ctx, uint32_t* values, uint64_t n, uint32_t first)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (n == 0) return SPHX_OK;
    if (!values) return argError(ctx, "sequence: null pointer");
    sequenceKernel<<<gridFor(n, 256), 256, 0, ctx->stream>>>(values, n, first);
    SPHX_LAUNCH_CHECK(ctx, "sequenceKernel");
    return SPHX_OK;
}

int sphx_gather_f64(sphx_ctx* ctx, const uint32_t* order, const double* in, double* out, uint64_t n)
{
    return gatherT(ctx, order, in, out, n);
}
int sphx_gather_f32(sphx_ctx* ctx, const uint32_t* order, const float* in, float* out, uint64_t n)
{
    return gatherT(ctx, order, in, out, n);
}
int sphx_gather_arrays(sphx_ctx* ctx, const uint32_t* order, uint64_t n, int32_t count, const void* const* in,
                       void* const* out, const uint32_t* elemBytes)
{
    return gatherArrays(ctx, order, n, count, in, out, elemBytes);
}

int sphx_gather_u32(sphx_ctx* ctx, const uint32_t* order, const uint32_t* in, uint32_t* out, uint64_t n)
{
    return gatherT(ctx, order, in, out, n);
}

int sphx_make_global_box(sphx_ctx* ctx, const double* x, const double* y, const double* z, uint64_t n,
                         const sphx_box* previous, sphx_box* out)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!previous || !out) return argError(ctx, "make_global_box: null box");
    const int32_t bc[3] = {previous->bx, previous->by, previous->bz};
    bool          keep[3];
    for (int d = 0; d < 3; ++d)
        keep[d] = bc[d] == SPHX_BC_PERIODIC || bc[d] == SPHX_BC_FIXED;
    sphx_box r = *previous;
    if (n > 0 && !(keep[0] && keep[1] && keep[2]))
    {
        if (!x || !y || !z) return argError(ctx, "make_global_box: null coordinates");
        uint32_t blocks  = uint32_t(std::min<uint64_t>(gridFor(n, 256), 1024));
        double*  partial = scratchT<double>(ctx, S_SMALL, size_t(blocks) * 6 + 6);
        if (!partial) return SPHX_ERR_NOMEM;
        boxExtremaKernel<<<blocks, 256, 0, ctx->stream>>>(x, y, z, n, partial);
        SPHX_LAUNCH_CHECK(ctx, "boxExtremaKernel");
        boxExtremaFinal<<<1, 64, 0, ctx->stream>>>(partial, blocks, partial + size_t(blocks) * 6);
        SPHX_LAUNCH_CHECK(ctx, "boxExtremaFinal");
        double e[6];
        SPHX_TRY(readBack(ctx, partial + size_t(blocks) * 6, e, sizeof(e)));
        if (!keep[0]) r.xmin = e[0], r.xmax = e[3];
        if (!keep[1]) r.ymin = e[1], r.ymax = e[4];
        if (!keep[2]) r.zmin = e[2], r.zmax = e[5];
    }
    *out = r;
    return SPHX_OK;
}

} // extern "C"
