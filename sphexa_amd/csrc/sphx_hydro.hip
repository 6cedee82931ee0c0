// sphx_hydro.hip - C ABI of the std-SPH loops (density, EOS, IAD,
// momentum-energy), the ideal-gas EOS kernel and the host-side kernel tables.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "sphx_internal.hpp"

#define SPHX_HYDRO_NS fast
#define SPHX_HYDRO_EXACT 0
#include "sphx_hydro_kernels.inc"
#include "sphx_ve_kernels.inc"
#include "sphx_adjoint.inc"

namespace sphx
{
namespace exact
{
int launchDensity(sphx_ctx*, const GroupsDev&, const SphLists&, sphx_particles*, const DBox&, bool);
int launchIAD(sphx_ctx*, const GroupsDev&, const SphLists&, sphx_particles*, const DBox&, bool, bool);
int launchMomentumEnergy(sphx_ctx*, const GroupsDev&, const SphLists&, sphx_particles*, const DBox&,
                         unsigned long long*, bool);
int launchXmass(sphx_ctx*, const GroupsDev&, const SphLists&, sphx_particles*, sphx_ve_fields*, const DBox&);
int launchVeDefGradh(sphx_ctx*, const GroupsDev&, const SphLists&, sphx_particles*, sphx_ve_fields*, const DBox&);
int launchIadDivvCurlv(sphx_ctx*, const GroupsDev&, const SphLists&, sphx_particles*, sphx_ve_fields*,
                       const DBox&);
int launchAvSwitches(sphx_ctx*, const GroupsDev&, const SphLists&, sphx_particles*, sphx_ve_fields*, const DBox&,
                     double);
int launchMomentumEnergyVe(sphx_ctx*, const GroupsDev&, const SphLists&, sphx_particles*, sphx_ve_fields*,
                           const DBox&, unsigned long long*);
} // namespace exact

int groupsFromArgs(sphx_ctx* ctx, const sphx_groups* groups, uint32_t first, uint32_t last, GroupsDev& G);

// idealGasEOS_u (sph/include/sph/eos.hpp:34-44) evaluated in double like the
// reference's common_type with a double gamma
template<class T>
__global__ __launch_bounds__(256) void eosKernel(uint32_t first, uint32_t last, const T* __restrict__ u,
                                                 const T* __restrict__ rho, double gamma, T* __restrict__ p,
                                                 T* __restrict__ c)
{
    uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) return;
    double tmp = double(u[i]) * (gamma - 1.0);
    p[i]       = T(double(rho[i]) * tmp);
    c[i]       = T(sqrt(gamma * tmp));
}

__global__ void initMinDt(unsigned long long* v)
{
    if (threadIdx.x == 0) *v = (unsigned long long)__double_as_longlong(INFINITY);
}

static bool useExact(const sphx_ctx* ctx, const sphx_particles* d)
{
    return ctx->fp64Exact && d->precision == SPHX_FP64;
}

// the SPH loops run on the lists of the last sphx_find_neighbors_sph call and
// require the same groups
static int listsFor(sphx_ctx* ctx, const sphx_groups* groups, const sphx_particles* d, GroupsDev& G,
                    SphLists& lists)
{
    if (!ctx->nsValid) return fail(ctx, SPHX_ERR_STATE, "no neighbor lists: call sphx_find_neighbors_sph first");
    SPHX_TRY(groupsFromArgs(ctx, groups, 0, uint32_t(d->n), G));
    const GroupsDev& N = ctx->nsGroups;
    if (G.first != N.first || G.last != N.last || G.numGroups != N.numGroups || G.start != N.start ||
        G.end != N.end || d->ngmax != ctx->nsNgmax || d->nc != ctx->nsNc)
        return fail(ctx, SPHX_ERR_STATE, "groups, ngmax or nc differ from the neighbor search");
    lists = ctx->nsLists;
    return SPHX_OK;
}

static int checkParticles(sphx_ctx* ctx, const sphx_particles* d, const char* what)
{
    if (!d) return argError(ctx, what);
    if (d->precision != SPHX_FP64 && d->precision != SPHX_MIXED) return argError(ctx, "unknown precision");
    if (!d->x || !d->y || !d->z || !d->h || !d->m || !d->nc || !d->wh) return argError(ctx, what);
    return SPHX_OK;
}


} // namespace sphx

using namespace sphx;

extern "C"
{

int sphx_ctx_set_option(sphx_ctx* ctx, int option, int value)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (option == SPHX_OPT_FP64_EXACT)
    {
        ctx->fp64Exact = value ? 1 : 0;
        return SPHX_OK;
    }
    if (option == SPHX_OPT_TIMING)
    {
        if (value < 0 || value > 2) return argError(ctx, "ctx_set_option: timing must be 0, 1 or 2");
        ctx->timing = value;
        return SPHX_OK;
    }
    if (option == SPHX_OPT_LOOP_SHAPE_FP64 || option == SPHX_OPT_LOOP_SHAPE_MIXED)
    {
        if (value != 0 && value != 1) return argError(ctx, "ctx_set_option: loop shape must be 0 or 1");
        ctx->loopShape[option == SPHX_OPT_LOOP_SHAPE_FP64 ? 0 : 1] = value;
        return SPHX_OK;
    }
    if (option == SPHX_OPT_DEFER_MIN_DT)
    {
        if (value != 0 && value != 1) return argError(ctx, "ctx_set_option: defer_min_dt must be 0 or 1");
        if (value && !ctx->minDtHost) return fail(ctx, SPHX_ERR_HIP, "ctx_set_option: no pinned memory to defer into");
        if (!value)
        {
            double dt; // a pending time step stays readable
            SPHX_TRY(sphx_min_dt_courant(ctx, &dt));
        }
        ctx->deferMinDt = value;
        return SPHX_OK;
    }
    return argError(ctx, "ctx_set_option: unknown option");
}

int sphx_get_stat(sphx_ctx* ctx, int stat, double* value)
{
    if (!ctx || !value) return SPHX_ERR_ARG;
    if (stat == SPHX_STAT_CANDIDATE_TESTS)
    {
        *value = ctx->nsTests;
        return SPHX_OK;
    }
    if (stat == SPHX_STAT_SCRATCH_BYTES)
    {
        double b = 0;
        for (const auto& s : ctx->buf)
            b += double(s.bytes);
        *value = b;
        return SPHX_OK;
    }
    return argError(ctx, "get_stat: unknown statistic");
}

int sphx_kernel_tables(double sincIndex, double* wh64, float* wh32, double* K)
{
#pragma clang fp contract(off)
    auto W = [sincIndex](double v) {
        double s = (v == 0.0) ? 1.0 : std::sin(M_PI_2 * v) / (M_PI_2 * v); // wharmonic_std, kernels.hpp:36-43
        return std::pow(s, sincIndex);
    };
    if (K)
    {
        // kernel_3D_k: 1 / Simpson(4 pi x^2 W(x), [0, 2], 2000 intervals) with sorted partial sums
        const uint64_t n = 2000;
        const double   a = 0.0, b = 2.0, hs = (b - a) / double(n);
        auto f = [&](double x) { return 4.0 * M_PI * x * x * W(x); };
        std::vector<double> odd(n / 2), even(n / 2 - 1);
        for (uint64_t i = 0; i < odd.size(); ++i)
            odd[i] = f(a + double(2 * (i + 1) - 1) * hs);
        for (uint64_t i = 0; i < even.size(); ++i)
            even[i] = f(a + double(2 * (i + 1)) * hs);
        std::sort(odd.begin(), odd.end());
        std::sort(even.begin(), even.end());
        double so = 0.0, se = 0.0;
        for (double v : odd)
            so += v;
        for (double v : even)
            se += v;
        *K = 1.0 / (hs / 3.0 * (f(a) + f(b) + 4.0 * so + 2.0 * se));
    }
    const int numIntervals = SPHX_TABLE_SIZE - 1;
    if (wh64)
    {
        const double dx = 2.0 / numIntervals;
        for (int i = 0; i < SPHX_TABLE_SIZE; ++i)
            wh64[i] = W(0.0 + double(size_t(i)) * dx);
    }
    if (wh32)
    {
        const float dx = float(2.0 / numIntervals);
        for (int i = 0; i < SPHX_TABLE_SIZE; ++i)
        {
            float v = float(0.0 + double(float(size_t(i)) * dx));
            wh32[i] = float(W(double(v)));
        }
    }
    return SPHX_OK;
}

int sphx_compute_density(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "compute_density: null field"));
    if (!d->rho) return argError(ctx, "compute_density: null rho");
    if (!validBox(box)) return argError(ctx, "compute_density: invalid box");
    GroupsDev       G;
    SphLists        lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    if (G.numGroups == 0) return SPHX_OK;
    DBox db = makeDBox(box);
    return useExact(ctx, d) ? exact::launchDensity(ctx, G, lists, d, db, false)
                            : fast::launchDensity(ctx, G, lists, d, db, false);
}

static int launchEos(sphx_ctx* ctx, uint32_t first, uint32_t last, sphx_particles* d)
{
    if (last == first) return SPHX_OK;
    if (d->precision == SPHX_FP64)
        {
            KTimer kt_(ctx, "eosKernel");
            eosKernel<double><<<gridFor(last - first, 256), 256, 0, ctx->stream>>>(
            first, last, static_cast<const double*>(d->u), static_cast<const double*>(d->rho), d->gamma,
            static_cast<double*>(d->p), static_cast<double*>(d->c));
        }
    else if (d->precision == SPHX_MIXED)
        {
            KTimer kt_(ctx, "eosKernel");
            eosKernel<float><<<gridFor(last - first, 256), 256, 0, ctx->stream>>>(
            first, last, static_cast<const float*>(d->u), static_cast<const float*>(d->rho), d->gamma,
            static_cast<float*>(d->p), static_cast<float*>(d->c));
        }
    else
        return argError(ctx, "compute_eos: unknown precision");
    SPHX_LAUNCH_CHECK(ctx, "eosKernel");
    return SPHX_OK;
}

int sphx_compute_eos(sphx_ctx* ctx, uint32_t first, uint32_t last, sphx_particles* d)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!d || !d->u || !d->rho || !d->p || !d->c) return argError(ctx, "compute_eos: null field");
    if (last < first || last > d->n) return argError(ctx, "compute_eos: bad range");
    SPHX_TRY(launchEos(ctx, first, last, d));
    return SPHX_OK;
}

int sphx_compute_iad(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "compute_iad: null field"));
    if (!d->rho || !d->c11 || !d->c12 || !d->c13 || !d->c22 || !d->c23 || !d->c33)
        return argError(ctx, "compute_iad: null field");
    if (!validBox(box)) return argError(ctx, "compute_iad: invalid box");
    GroupsDev       G;
    SphLists        lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    if (G.numGroups == 0) return SPHX_OK;
    DBox db = makeDBox(box);
    return useExact(ctx, d) ? exact::launchIAD(ctx, G, lists, d, db, false, false)
                            : fast::launchIAD(ctx, G, lists, d, db, false, false);
}

// The Courant time step of a momentum-energy call: read back (the stream synchronised), or
// with SPHX_OPT_DEFER_MIN_DT copied to pinned memory behind the loop and d->minDtCourant
// set to NaN until sphx_min_dt_courant waits for it.  The device minimum is bits of a
// non-negative double (ordered as integers).
static int finishMinDt(sphx_ctx* ctx, const unsigned long long* dt, sphx_particles* d)
{
    if (ctx->deferMinDt)
    {
        SPHX_HIP(ctx, hipMemcpyAsync(ctx->minDtHost, dt, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        SPHX_HIP(ctx, hipEventRecord(ctx->minDtEvent, ctx->stream));
        ctx->minDtPending = true;
        d->minDtCourant   = NAN;
        return SPHX_OK;
    }
    unsigned long long bits = 0;
    SPHX_TRY(readBack(ctx, dt, &bits, sizeof(bits)));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    d->minDtCourant   = v;
    ctx->minDtPending = false;
    ctx->minDtLast    = v;
    return SPHX_OK;
}

// a call without targets: no pairs, no limit
static int noMinDt(sphx_ctx* ctx, sphx_particles* d)
{
    d->minDtCourant   = INFINITY;
    ctx->minDtPending = false;
    ctx->minDtLast    = INFINITY;
    return SPHX_OK;
}

int sphx_min_dt_courant(sphx_ctx* ctx, double* out)
{
    if (!ctx || !out) return SPHX_ERR_ARG;
    if (ctx->minDtPending)
    {
        SPHX_HIP(ctx, hipEventSynchronize(ctx->minDtEvent));
        std::memcpy(&ctx->minDtLast, ctx->minDtHost, sizeof(double));
        ctx->minDtPending = false;
    }
    *out = ctx->minDtLast;
    return SPHX_OK;
}

int sphx_compute_momentum_energy(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "compute_momentum_energy: null field"));
    if (!d->vx || !d->vy || !d->vz || !d->rho || !d->p || !d->c || !d->c11 || !d->c12 || !d->c13 || !d->c22 ||
        !d->c23 || !d->c33 || !d->ax || !d->ay || !d->az || !d->du)
        return argError(ctx, "compute_momentum_energy: null field");
    if (!validBox(box)) return argError(ctx, "compute_momentum_energy: invalid box");
    GroupsDev       G;
    SphLists        lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    d->minDtCourant = INFINITY;
    if (G.numGroups == 0) return noMinDt(ctx, d);
    auto* dt = scratchT<unsigned long long>(ctx, S_SMALL, 4);
    if (!dt) return SPHX_ERR_NOMEM;
    DBox db = makeDBox(box);
    initMinDt<<<1, 64, 0, ctx->stream>>>(dt);
    SPHX_LAUNCH_CHECK(ctx, "initMinDt");
    SPHX_TRY(useExact(ctx, d) ? exact::launchMomentumEnergy(ctx, G, lists, d, db, dt, false)
                              : fast::launchMomentumEnergy(ctx, G, lists, d, db, dt, false));
    return finishMinDt(ctx, dt, d);
}

int sphx_compute_hydro_std(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, const sphx_box* box)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "compute_hydro_std: null field"));
    if (!d->u || !d->vx || !d->vy || !d->vz || !d->rho || !d->p || !d->c || !d->c11 || !d->c12 || !d->c13 ||
        !d->c22 || !d->c23 || !d->c33 || !d->ax || !d->ay || !d->az || !d->du)
        return argError(ctx, "compute_hydro_std: null field");
    if (d->precision != SPHX_FP64 && d->precision != SPHX_MIXED)
        return argError(ctx, "compute_hydro_std: unknown precision");
    if (!validBox(box)) return argError(ctx, "compute_hydro_std: invalid box");
    GroupsDev G;
    SphLists  lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    d->minDtCourant = INFINITY;
    if (G.numGroups == 0) return noMinDt(ctx, d);
    auto* dt = scratchT<unsigned long long>(ctx, S_SMALL, 4);
    if (!dt) return SPHX_ERR_NOMEM;
    // every particle a target: the density loop's epilogue writes the IAD records and, in
    // mixed precision, the IAD loop's epilogue the momentum-energy records (the 144-byte
    // FP64 records cost the FP64 IAD epilogue more than their packing launch: measured)
    const bool handOff = G.first == 0 && G.last == d->n;
    const bool meOff   = handOff && d->precision == SPHX_MIXED;
    const bool ex      = useExact(ctx, d);
    DBox       db      = makeDBox(box);
    SPHX_TRY(ex ? exact::launchDensity(ctx, G, lists, d, db, handOff)
                : fast::launchDensity(ctx, G, lists, d, db, handOff));
    SPHX_TRY(launchEos(ctx, G.first, G.last, d));
    SPHX_TRY(ex ? exact::launchIAD(ctx, G, lists, d, db, handOff, meOff)
                : fast::launchIAD(ctx, G, lists, d, db, handOff, meOff));
    initMinDt<<<1, 64, 0, ctx->stream>>>(dt);
    SPHX_LAUNCH_CHECK(ctx, "initMinDt");
    SPHX_TRY(ex ? exact::launchMomentumEnergy(ctx, G, lists, d, db, dt, meOff)
                : fast::launchMomentumEnergy(ctx, G, lists, d, db, dt, meOff));
    return finishMinDt(ctx, dt, d);
}

int sphx_kernel_derivative_tables(double sincIndex, double* whd64, float* whd32)
{
#pragma clang fp contract(off)
    // powSincDerivative (sph_kernel_tables.hpp:104-108) with wharmonic_derivative_std (kernels.hpp:45-58).
    // The libm calls are those of the reference CPU path as g++ -O2 builds it: the sin and cos
    // of wharmonic_derivative_std come from one sincos call (g++ merges them), the sin of
    // wharmonic_std from a separate sin call.  glibc's sincos and sin round differently for a
    // few arguments, so this choice decides a handful of the table's last bits.
    auto dW = [sincIndex](double v) {
        double s = 1.0, d = 0.0;
        if (v != 0.0)
        {
            const double Pv = M_PI_2 * v;
            double       sinPv, cosPv;
            ::sincos(Pv, &sinPv, &cosPv);
            const double sincv = sinPv / Pv;
            d                  = sincv * M_PI_2 * ((cosPv / sinPv) - 1.0 / Pv);
            s                  = std::sin(Pv) / Pv;
        }
        return sincIndex * std::pow(s, sincIndex - 1) * d;
    };
    const int numIntervals = SPHX_TABLE_SIZE - 1;
    if (whd64)
    {
        const double dx = 2.0 / numIntervals;
        for (int i = 0; i < SPHX_TABLE_SIZE; ++i)
            whd64[i] = dW(0.0 + double(size_t(i)) * dx);
    }
    if (whd32)
    {
        const float dx = float(2.0 / numIntervals);
        for (int i = 0; i < SPHX_TABLE_SIZE; ++i)
            whd32[i] = float(dW(double(float(0.0 + double(float(size_t(i)) * dx)))));
    }
    return SPHX_OK;
}

static int checkVe(sphx_ctx* ctx, const sphx_ve_fields* ve, const char* what)
{
    if (!ve) return argError(ctx, what);
    return SPHX_OK;
}

int sphx_compute_xmass(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, sphx_ve_fields* ve,
                       const sphx_box* box)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "compute_xmass: null field"));
    SPHX_TRY(checkVe(ctx, ve, "compute_xmass: null ve fields"));
    if (!ve->xm) return argError(ctx, "compute_xmass: null xm");
    if (!validBox(box)) return argError(ctx, "compute_xmass: invalid box");
    GroupsDev G;
    SphLists  lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    if (G.numGroups == 0) return SPHX_OK;
    DBox db = makeDBox(box);
    return useExact(ctx, d) ? exact::launchXmass(ctx, G, lists, d, ve, db) : fast::launchXmass(ctx, G, lists, d, ve, db);
}

int sphx_compute_ve_def_gradh(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, sphx_ve_fields* ve,
                              const sphx_box* box)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "compute_ve_def_gradh: null field"));
    SPHX_TRY(checkVe(ctx, ve, "compute_ve_def_gradh: null ve fields"));
    if (!ve->xm || !ve->kx || !ve->gradh || (d->precision == SPHX_FP64 && !ve->whd))
        return argError(ctx, "compute_ve_def_gradh: null xm, kx, gradh or whd");
    if (!validBox(box)) return argError(ctx, "compute_ve_def_gradh: invalid box");
    GroupsDev G;
    SphLists  lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    if (G.numGroups == 0) return SPHX_OK;
    DBox db = makeDBox(box);
    return useExact(ctx, d) ? exact::launchVeDefGradh(ctx, G, lists, d, ve, db)
                            : fast::launchVeDefGradh(ctx, G, lists, d, ve, db);
}

int sphx_compute_eos_ve(sphx_ctx* ctx, uint32_t first, uint32_t last, sphx_particles* d, sphx_ve_fields* ve)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!d || !d->u || !d->m || !d->c) return argError(ctx, "compute_eos_ve: null field");
    if (d->precision != SPHX_FP64 && d->precision != SPHX_MIXED)
        return argError(ctx, "compute_eos_ve: unknown precision");
    SPHX_TRY(checkVe(ctx, ve, "compute_eos_ve: null ve fields"));
    if (!ve->xm || !ve->kx || !ve->gradh || !ve->prho) return argError(ctx, "compute_eos_ve: null ve field");
    if (last < first || last > d->n) return argError(ctx, "compute_eos_ve: bad range");
    if (last == first) return SPHX_OK;
    return fast::launchEosVe(ctx, first, last, d, ve);
}

int sphx_compute_iad_divv_curlv(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, sphx_ve_fields* ve,
                                const sphx_box* box)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "compute_iad_divv_curlv: null field"));
    SPHX_TRY(checkVe(ctx, ve, "compute_iad_divv_curlv: null ve fields"));
    if (!d->vx || !d->vy || !d->vz || !d->c11 || !d->c12 || !d->c13 || !d->c22 || !d->c23 || !d->c33 || !ve->xm ||
        !ve->kx || !ve->divv)
        return argError(ctx, "compute_iad_divv_curlv: null field");
    if (!validBox(box)) return argError(ctx, "compute_iad_divv_curlv: invalid box");
    GroupsDev G;
    SphLists  lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    if (G.numGroups == 0) return SPHX_OK;
    DBox db = makeDBox(box);
    return useExact(ctx, d) ? exact::launchIadDivvCurlv(ctx, G, lists, d, ve, db)
                            : fast::launchIadDivvCurlv(ctx, G, lists, d, ve, db);
}

int sphx_compute_av_switches(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d, sphx_ve_fields* ve,
                             const sphx_box* box, double dt)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "compute_av_switches: null field"));
    SPHX_TRY(checkVe(ctx, ve, "compute_av_switches: null ve fields"));
    if (!d->vx || !d->vy || !d->vz || !d->c || !d->c11 || !d->c12 || !d->c13 || !d->c22 || !d->c23 || !d->c33 ||
        !ve->xm || !ve->kx || !ve->divv || !ve->alpha)
        return argError(ctx, "compute_av_switches: null field");
    if (!validBox(box)) return argError(ctx, "compute_av_switches: invalid box");
    GroupsDev G;
    SphLists  lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    if (G.numGroups == 0) return SPHX_OK;
    DBox db = makeDBox(box);
    return useExact(ctx, d) ? exact::launchAvSwitches(ctx, G, lists, d, ve, db, dt)
                            : fast::launchAvSwitches(ctx, G, lists, d, ve, db, dt);
}

int sphx_compute_momentum_energy_ve(sphx_ctx* ctx, const sphx_groups* groups, sphx_particles* d,
                                    sphx_ve_fields* ve, const sphx_box* box)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "compute_momentum_energy_ve: null field"));
    SPHX_TRY(checkVe(ctx, ve, "compute_momentum_energy_ve: null ve fields"));
    if (!d->vx || !d->vy || !d->vz || !d->c || !d->c11 || !d->c12 || !d->c13 || !d->c22 || !d->c23 || !d->c33 ||
        !d->ax || !d->ay || !d->az || !d->du || !ve->xm || !ve->kx || !ve->prho || !ve->alpha)
        return argError(ctx, "compute_momentum_energy_ve: null field");
    if (ve->avClean && (!ve->dV11 || !ve->dV12 || !ve->dV13 || !ve->dV22 || !ve->dV23 || !ve->dV33))
        return argError(ctx, "compute_momentum_energy_ve: AV cleaning needs dV11..dV33");
    if (!validBox(box)) return argError(ctx, "compute_momentum_energy_ve: invalid box");
    GroupsDev G;
    SphLists  lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    d->minDtCourant = INFINITY;
    if (G.numGroups == 0) return noMinDt(ctx, d);
    auto* dt = scratchT<unsigned long long>(ctx, S_SMALL, 4);
    if (!dt) return SPHX_ERR_NOMEM;
    DBox db = makeDBox(box);
    initMinDt<<<1, 64, 0, ctx->stream>>>(dt);
    SPHX_LAUNCH_CHECK(ctx, "initMinDt");
    SPHX_TRY(useExact(ctx, d) ? exact::launchMomentumEnergyVe(ctx, G, lists, d, ve, db, dt)
                              : fast::launchMomentumEnergyVe(ctx, G, lists, d, ve, db, dt));
    return finishMinDt(ctx, dt, d);
}

int sphx_rho_timestep(sphx_ctx* ctx, uint32_t first, uint32_t last, const sphx_particles* d,
                      const sphx_ve_fields* ve, double Krho, double* minDtRho)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!d || !ve || !ve->divv || !minDtRho) return argError(ctx, "rho_timestep: null argument");
    if (d->precision != SPHX_FP64 && d->precision != SPHX_MIXED) return argError(ctx, "rho_timestep: unknown precision");
    if (last < first || last > d->n) return argError(ctx, "rho_timestep: bad range");
    if (last == first)
    {
        *minDtRho = Krho / INFINITY; // max over nothing is -inf
        return SPHX_OK;
    }
    return fast::launchRhoTimestep(ctx, first, last, d, ve, Krho, minDtRho);
}

int sphx_density_adjoint(sphx_ctx* ctx, const sphx_groups* groups, const sphx_particles* d, const sphx_box* box,
                         const void* gRho, double* gx, double* gy, double* gz, void* gh)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "density_adjoint: null field"));
    if (!gRho || !gx || !gy || !gz) return argError(ctx, "density_adjoint: null gradient pointer");
    if (!validBox(box)) return argError(ctx, "density_adjoint: invalid box");
    GroupsDev G;
    SphLists  lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    if (G.numGroups == 0) return SPHX_OK;
    if (G.first != 0 || G.last != d->n) return argError(ctx, "density_adjoint: groups must cover [0, n)");
    DBox db = makeDBox(box);
    return d->precision == SPHX_FP64 ? fast::densityAdjointT<double>(ctx, G, lists, d, db, gRho, gx, gy, gz, gh)
                                     : fast::densityAdjointT<float>(ctx, G, lists, d, db, gRho, gx, gy, gz, gh);
}

int sphx_momentum_energy_adjoint(sphx_ctx* ctx, const sphx_groups* groups, const sphx_particles* d,
                                 const sphx_box* box, const void* gax, const void* gay, const void* gaz,
                                 const void* gdu, double* gx, double* gy, double* gz)
{
    if (!ctx) return SPHX_ERR_ARG;
    SPHX_TRY(checkParticles(ctx, d, "momentum_energy_adjoint: null field"));
    if (!d->vx || !d->vy || !d->vz || !d->rho || !d->p || !d->c || !d->c11 || !d->c12 || !d->c13 || !d->c22 ||
        !d->c23 || !d->c33)
        return argError(ctx, "momentum_energy_adjoint: null field");
    if (!gx || !gy || !gz) return argError(ctx, "momentum_energy_adjoint: null gradient pointer");
    if (!validBox(box)) return argError(ctx, "momentum_energy_adjoint: invalid box");
    GroupsDev G;
    SphLists  lists;
    SPHX_TRY(listsFor(ctx, groups, d, G, lists));
    if (G.numGroups == 0) return SPHX_OK;
    if (G.first != 0 || G.last != d->n) return argError(ctx, "momentum_energy_adjoint: groups must cover [0, n)");
    DBox db = makeDBox(box);
    return d->precision == SPHX_FP64
               ? fast::meAdjointT<double>(ctx, G, lists, d, db, gax, gay, gaz, gdu, gx, gy, gz)
               : fast::meAdjointT<float>(ctx, G, lists, d, db, gax, gay, gaz, gdu, gx, gy, gz);
}

} // extern "C"
