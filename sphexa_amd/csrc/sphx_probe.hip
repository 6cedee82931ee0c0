// sphx_probe.hip - measurement helper: the sustained VALU FMA rate of the device, to check
// the published vector peaks that the roofline figures are quoted against.  No counterpart
// in the reference.
#include <algorithm>

#include "sphx_internal.hpp"

namespace sphx
{

// 8 independent FMA chains per lane, `iters` rounds; the result is stored so that nothing
// is optimised away.  FP32 here is scalar (non-packed) FMA: the build disables packed FP32.
template<class T>
__global__ __launch_bounds__(256) void fmaChains(uint32_t iters, T b, T c, T* __restrict__ out)
{
    T a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        a[k] = T(threadIdx.x + k) * T(1e-3);
    for (uint32_t it = 0; it < iters; ++it)
    {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            a[k] = a[k] * b + c;
    }
    T s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        s += a[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

} // namespace sphx

using namespace sphx;

extern "C"
{

int sphx_valu_peak(sphx_ctx* ctx, int precision, double* tflops)
{
    if (!ctx) return SPHX_ERR_ARG;
    if (!tflops || (precision != SPHX_FP64 && precision != SPHX_MIXED)) return argError(ctx, "valu_peak: bad argument");
    const unsigned blocks = unsigned(ctx->numCU) * 16u, threads = 256u;
    const uint32_t iters  = 2048;
    void*          out    = scratch(ctx, S_SMALL, size_t(blocks) * threads * sizeof(double) + 256);
    if (!out) return SPHX_ERR_NOMEM;
    hipEvent_t a, b;
    SPHX_HIP(ctx, hipEventCreate(&a));
    SPHX_HIP(ctx, hipEventCreate(&b));
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) // the first launch warms up
    {
        SPHX_HIP(ctx, hipEventRecord(a, ctx->stream));
        if (precision == SPHX_FP64)
            fmaChains<double><<<blocks, threads, 0, ctx->stream>>>(iters, 0.999999, 1e-7, static_cast<double*>(out));
        else
            fmaChains<float><<<blocks, threads, 0, ctx->stream>>>(iters, 0.999999f, 1e-7f, static_cast<float*>(out));
        SPHX_HIP(ctx, hipEventRecord(b, ctx->stream));
        SPHX_LAUNCH_CHECK(ctx, "fmaChains");
        SPHX_HIP(ctx, hipEventSynchronize(b));
        SPHX_HIP(ctx, hipEventElapsedTime(&ms, a, b));
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    const double flops = double(blocks) * threads * iters * 8.0 * 2.0;
    *tflops            = flops / (double(ms) * 1e-3) / 1e12;
    return SPHX_OK;
}

} // extern "C"
