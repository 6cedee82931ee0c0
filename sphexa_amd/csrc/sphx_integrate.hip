// sphx_integrate.hip - the integration step after the force evaluation: global time
// step (computeTimestep), 2nd-order position and energy update (computePositions,
// CPU semantics) and the smoothing-length update (updateSmoothingLength).
//
// All three are elementwise over the particles and bound by HBM bandwidth; one
// thread per particle, 256-thread blocks, coalesced SoA loads and stores.  The
// position/energy kernel is compiled without FMA contraction so that, in FP64, it
// rounds every operation exactly as the reference CPU loop does.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "sphx_internal.hpp"

namespace sphx
{

int groupsFromArgs(sphx_ctx* ctx, const sphx_groups* groups, uint32_t first, uint32_t last, GroupsDev& G);

// box limits and boundary kinds as the position update reads them
struct IntBox
{
    double lo[3], hi[3], len[3];
    int    pbc[3], fixed[3];
    int    anyFixed;
};

static IntBox makeIntBox(const sphx_box* b)
{
    DBox    d = makeDBox(b);
    IntBox  r;
    int     bc[3] = {b->bx, b->by, b->bz};
    r.anyFixed    = 0;
    for (int k = 0; k < 3; ++k)
    {
        r.lo[k] = d.lo[k], r.hi[k] = d.hi[k], r.len[k] = d.len[k];
        r.pbc[k]   = d.pbc[k];
        r.fixed[k] = (bc[k] == SPHX_BC_FIXED);
        r.anyFixed |= r.fixed[k];
    }
    return r;
}

// fbcAdjustFactors (sph/include/sph/positions.hpp:47-74): smooth 0..1 factor per
// dimension within 5 smoothing lengths of a fixed boundary
__device__ inline void fbcFactors(const double X[3], const IntBox& box, double hi, double f[3])
{
#pragma clang fp contract(off)
    const double width = 2.0, invWidth = 1.0 / width, offset = 1.0;
    for (int k = 0; k < 3; ++k)
    {
        f[k]      = 1.0;
        double md = fmin(fabs(box.hi[k] - X[k]), fabs(box.lo[k] - X[k])) * (1.0 / hi);
        if (box.fixed[k] && md < 2 * width + offset) f[k] = 0.5 * (tanh(4 * (md - offset) * invWidth - 4) + 1);
    }
}

// energyUpdate (positions.hpp:77-84): Adams-Bashforth 2, exponential decay instead of u < 0
__device__ inline double energyUpdate(double u_old, double dt, double dt_m1, double du, double du_m1)
{
#pragma clang fp contract(off)
    double u_new = u_old + du * dt + 0.5 * (du - du_m1) / dt_m1 * fabs(dt) * dt;
    if (u_new < 0.) u_new = u_old * exp(u_new * dt / u_old);
    return u_new;
}

// updatePositionsHost + updateIntEnergyHost (positions.hpp:113-179) for one particle per
// thread: positionUpdate (:100-111) with putInBox (cstone/sfc/box.hpp:193-214), then the
// energy update with the fixed-boundary factor of the new position
template<class T>
__global__ __launch_bounds__(256) void positionsKernel(uint32_t first, uint32_t last, double* __restrict__ x,
                                                       double* __restrict__ y, double* __restrict__ z,
                                                       T* __restrict__ xm1, T* __restrict__ ym1, T* __restrict__ zm1,
                                                       T* __restrict__ vx, T* __restrict__ vy, T* __restrict__ vz,
                                                       const T* __restrict__ ax, const T* __restrict__ ay,
                                                       const T* __restrict__ az, const T* __restrict__ h,
                                                       T* __restrict__ u, const T* __restrict__ du,
                                                       T* __restrict__ dum1, double dt, double dt_m1, IntBox box)
{
#pragma clang fp contract(off)
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) return;
    double X[3] = {x[i], y[i], z[i]};
    double f[3] = {1.0, 1.0, 1.0};
    double hi   = box.anyFixed ? double(h[i]) : 1.0;
    if (box.anyFixed) fbcFactors(X, box, hi, f);
    const double A[3]   = {double(ax[i]) * f[0], double(ay[i]) * f[1], double(az[i]) * f[2]};
    const double dXn[3] = {double(xm1[i]) * f[0], double(ym1[i]) * f[1], double(zm1[i]) * f[2]};
    const double idt_m1 = 1.0 / dt_m1, half_dt_m1 = 0.5 * dt_m1, adt = fabs(dt);
    double       V[3], dX[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        double Vn = dXn[k] * idt_m1 + half_dt_m1 * A[k];
        V[k]      = Vn + A[k] * dt;
        dX[k]     = (Vn + 0.5 * A[k] * adt) * dt;
        double Xk = X[k] + dX[k];
        if (box.pbc[k] && Xk > box.hi[k]) Xk -= box.len[k];
        else if (box.pbc[k] && Xk < box.lo[k]) Xk += box.len[k];
        X[k] = Xk;
    }
    x[i] = X[0], y[i] = X[1], z[i] = X[2];
    xm1[i] = T(dX[0]), ym1[i] = T(dX[1]), zm1[i] = T(dX[2]);
    vx[i] = T(V[0]), vy[i] = T(V[1]), vz[i] = T(V[2]);
    double fe = 1.0;
    if (box.anyFixed)
    {
        fbcFactors(X, box, hi, f);
        fe = fmin(fmin(f[0], f[1]), f[2]);
    }
    T duI   = du[i];
    u[i]    = T(energyUpdate(double(u[i]), dt * fe, dt_m1 * fe, double(duI), double(dum1[i])));
    dum1[i] = duI;
}

// accelerationTimestep (sph/include/sph/ts_global.hpp:48-71): min_i h_i^2 / |a_i|^2,
// h_i^2 formed in the field type as the reference's HydroType product; NaN terms are
// skipped.  Non-negative doubles order like their bit patterns, so the block's
// minimum enters *minBits with a 64-bit atomic min.
template<class T>
__global__ __launch_bounds__(256) void minH2A2Kernel(uint32_t first, uint32_t last, const T* __restrict__ h,
                                                     const T* __restrict__ ax, const T* __restrict__ ay,
                                                     const T* __restrict__ az, unsigned long long* minBits)
{
#pragma clang fp contract(off)
    double m = INFINITY;
    for (uint64_t i = uint64_t(first) + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < last;
         i += uint64_t(gridDim.x) * blockDim.x)
    {
        double a0 = ax[i], a1 = ay[i], a2 = az[i];
        // norm2 = dot(a, a), a right fold: a0 a0 + (a1 a1 + a2 a2) (cstone/util/array.hpp:236-240)
        double v  = double(T(h[i] * h[i])) / (a0 * a0 + (a1 * a1 + a2 * a2));
        if (v < m) m = v;
    }
    for (int o = 32; o > 0; o >>= 1)
    {
        double other = __shfl_xor(m, o, 64);
        if (other < m) m = other;
    }
    // one atomic per block (a device-scope access per wave to one address serializes)
    __shared__ double wmin[4];
    if ((threadIdx.x & 63u) == 0) wmin[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < 4; ++w)
            if (wmin[w] < m) m = wmin[w];
        if (m < INFINITY) atomicMin(minBits, (unsigned long long)__double_as_longlong(m));
    }
}

__global__ void fillU64(unsigned long long* p, unsigned long long v, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// updateSmoothingLengthGpu semantics (sph/include/sph/update_h.hpp:13-32): new h from
// the neighbor count (updateH, sph/kernels.hpp:27-33, in the precision of h with the
// tabulated factor), removeKey for particles with at most themselves as neighbor.
// status bit 0: a key was removed; bit 1: a count beyond the factor table (h unchanged)
template<class T>
__global__ __launch_bounds__(256) void updateHKernel(uint32_t first, uint32_t last, HFactors hfac,
                                                     const uint32_t* __restrict__ nc, T* __restrict__ h,
                                                     uint64_t* __restrict__ keys, uint32_t* status)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) return;
    uint32_t c = nc[i];
    if (c <= 1)
    {
        if (keys) keys[i] = kRemoveKey;
        atomicOr(status, 1u);
    }
    T hv = h[i];
    if (updateHTab(hfac, c, hv)) h[i] = hv;
    else atomicOr(status, 2u);
}

} // namespace sphx

using namespace sphx;

extern "C"
{

int sphx_compute_timestep(sphx_ctx* ctx, uint32_t first, uint32_t last, const sphx_particles* d, sphx_timestep* ts)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!ts) return argError(ctx, "compute_timestep: null timestep");
    double minDtAcc = INFINITY;
    if (ts->g != 0.0)
    {
        if (!d || !d->h || !d->ax || !d->ay || !d->az) return argError(ctx, "compute_timestep: null field");
        if (d->precision != SPHX_FP64 && d->precision != SPHX_MIXED)
            return argError(ctx, "compute_timestep: unknown precision");
        if (last < first || last > d->n) return argError(ctx, "compute_timestep: bad range");
        if (last > first)
        {
            auto* mb = scratchT<unsigned long long>(ctx, S_SMALL, 1);
            if (!mb) return SPHX_ERR_NOMEM;
            const unsigned long long infBits = (unsigned long long)0x7ff0000000000000ull;
            fillU64<<<1, 64, 0, ctx->stream>>>(mb, infBits, 1);
            SPHX_LAUNCH_CHECK(ctx, "fillU64");
            unsigned blocks = std::min(gridFor(last - first, 256), unsigned(ctx->numCU) * 8u);
            {
                KTimer kt_(ctx, "minH2A2Kernel");
                if (d->precision == SPHX_FP64)
                    minH2A2Kernel<double><<<blocks, 256, 0, ctx->stream>>>(
                        first, last, static_cast<const double*>(d->h), static_cast<const double*>(d->ax),
                        static_cast<const double*>(d->ay), static_cast<const double*>(d->az), mb);
                else
                    minH2A2Kernel<float><<<blocks, 256, 0, ctx->stream>>>(
                        first, last, static_cast<const float*>(d->h), static_cast<const float*>(d->ax),
                        static_cast<const float*>(d->ay), static_cast<const float*>(d->az), mb);
            }
            SPHX_LAUNCH_CHECK(ctx, "minH2A2Kernel");
            unsigned long long bits = 0;
            SPHX_TRY(readBack(ctx, mb, &bits, sizeof(bits)));
            double m;
            std::memcpy(&m, &bits, sizeof(m));
            minDtAcc = ts->etaAcc * std::pow(m, 0.25);
        }
    }
    const double maxDtIncrease = 1.1; // ParticlesData::maxDtIncrease, particles_data.hpp:138
    double       minDt = std::min(std::min(minDtAcc, ts->minDtCourant), std::min(ts->minDtRho, maxDtIncrease * ts->minDt));
    ts->minDtAcc = minDtAcc;
    ts->ttot += minDt;
    ts->minDt_m1 = ts->minDt;
    ts->minDt    = minDt;
    return SPHX_OK;
}

int sphx_compute_positions(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box,
                           const sphx_timestep* ts)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!d || !ts) return argError(ctx, "compute_positions: null argument");
    if (d->n == 0) return SPHX_OK; // nothing to move (empty fields may be null)
    if (!d->x || !d->y || !d->z || !d->x_m1 || !d->y_m1 || !d->z_m1 || !d->vx || !d->vy || !d->vz || !d->ax ||
        !d->ay || !d->az || !d->h || !d->u || !d->du || !d->du_m1)
        return argError(ctx, "compute_positions: null field");
    if (!validBox(box)) return argError(ctx, "compute_positions: invalid box");
    GroupsDev G;
    SPHX_TRY(groupsFromArgs(ctx, groups, 0, uint32_t(d->n), G));
    if (G.last > d->n) return argError(ctx, "compute_positions: groups exceed n");
    if (G.last == G.first) return SPHX_OK;
    IntBox   ib = makeIntBox(box);
    unsigned blocks = gridFor(G.last - G.first, 256);
    KTimer   kt_(ctx, "positionsKernel");
    if (d->precision == SPHX_FP64)
    {
        using T = double;
        positionsKernel<T><<<blocks, 256, 0, ctx->stream>>>(
            G.first, G.last, d->x, d->y, d->z, (T*)d->x_m1, (T*)d->y_m1, (T*)d->z_m1, (T*)d->vx, (T*)d->vy,
            (T*)d->vz, (const T*)d->ax, (const T*)d->ay, (const T*)d->az, (const T*)d->h, (T*)d->u,
            (const T*)d->du, (T*)d->du_m1, ts->minDt, ts->minDt_m1, ib);
    }
    else if (d->precision == SPHX_MIXED)
    {
        using T = float;
        positionsKernel<T><<<blocks, 256, 0, ctx->stream>>>(
            G.first, G.last, d->x, d->y, d->z, (T*)d->x_m1, (T*)d->y_m1, (T*)d->z_m1, (T*)d->vx, (T*)d->vy,
            (T*)d->vz, (const T*)d->ax, (const T*)d->ay, (const T*)d->az, (const T*)d->h, (T*)d->u,
            (const T*)d->du, (T*)d->du_m1, ts->minDt, ts->minDt_m1, ib);
    }
    else
        return argError(ctx, "compute_positions: unknown precision");
    SPHX_LAUNCH_CHECK(ctx, "positionsKernel");
    return SPHX_OK;
}

int sphx_update_smoothing_length(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, uint64_t* keys,
                                 int* keysRemoved)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (keysRemoved) *keysRemoved = 0;
    if (d && d->n == 0) return SPHX_OK; // nothing to update (empty fields may be null)
    if (!d || !d->h || !d->nc) return argError(ctx, "update_smoothing_length: null field");
    GroupsDev G;
    SPHX_TRY(groupsFromArgs(ctx, groups, 0, uint32_t(d->n), G));
    if (G.last > d->n) return argError(ctx, "update_smoothing_length: groups exceed n");
    if (keysRemoved) *keysRemoved = 0;
    if (G.last == G.first) return SPHX_OK;
    auto* flag = scratchT<unsigned long long>(ctx, S_SMALL, 1);
    if (!flag) return SPHX_ERR_NOMEM;
    fillU64<<<1, 64, 0, ctx->stream>>>(flag, 0ull, 1);
    SPHX_LAUNCH_CHECK(ctx, "fillU64");
    if (d->precision != SPHX_FP64 && d->precision != SPHX_MIXED)
        return argError(ctx, "update_smoothing_length: unknown precision");
    HFactors hfac;
    SPHX_TRY(hFactors(ctx, d->ng0, hfac));
    uint32_t* status = reinterpret_cast<uint32_t*>(flag);
    unsigned  blocks = gridFor(G.last - G.first, 256);
    {
        KTimer kt_(ctx, "updateHKernel");
        if (d->precision == SPHX_FP64)
            updateHKernel<double><<<blocks, 256, 0, ctx->stream>>>(G.first, G.last, hfac, d->nc,
                                                                    static_cast<double*>(d->h), keys, status);
        else
            updateHKernel<float><<<blocks, 256, 0, ctx->stream>>>(G.first, G.last, hfac, d->nc,
                                                                   static_cast<float*>(d->h), keys, status);
    }
    SPHX_LAUNCH_CHECK(ctx, "updateHKernel");
    uint32_t r = 0;
    SPHX_TRY(readBack(ctx, status, &r, sizeof(r)));
    if (keysRemoved) *keysRemoved = (r & 1u) ? 1 : 0;
    if (r & 2u)
        return fail(ctx, SPHX_ERR_ARG, "update_smoothing_length: a neighbor count exceeds the factor table (" +
                                           std::to_string(kHFactorCap) + "); h left unchanged there");
    return SPHX_OK;
}

} // extern "C"
