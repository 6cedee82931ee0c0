// bench_valu_probe.hip - bench.py's measurement helper (not part of libsphx): the sustained
// VALU FMA rate of the device, measured beside the published vector peaks that the roofline
// figures are quoted against.  8 independent FMA chains per lane on 16 blocks per CU; FP32 is
// scalar (non-packed) FMA, as the library is built without packed FP32.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace
{
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
    out[blockIdx.x * blockDim.x + threadIdx.x] = s; // keeps the chains live
}
} // namespace

// TFLOP/s (an FMA counts 2 flops) of FP64 (fp64 != 0) or scalar FP32 FMA chains on `device`;
// 0 on success, else the HIP error code
extern "C" int bench_valu_peak(int device, int fp64, double* tflops)
{
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return int(e);
    int cu = 0;
    if ((e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) return int(e);
    const unsigned blocks = unsigned(cu) * 16u, threads = 256u;
    const uint32_t iters  = 2048;
    void*          out    = nullptr;
    if ((e = hipMalloc(&out, size_t(blocks) * threads * sizeof(double))) != hipSuccess) return int(e);
    hipEvent_t a = nullptr, b = nullptr;
    float      ms = 0.f;
    if ((e = hipEventCreate(&a)) == hipSuccess && (e = hipEventCreate(&b)) == hipSuccess)
    {
        for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) // the first launch warms up
        {
            (void)hipEventRecord(a, nullptr);
            if (fp64)
                fmaChains<double><<<blocks, threads>>>(iters, 0.999999, 1e-7, static_cast<double*>(out));
            else
                fmaChains<float><<<blocks, threads>>>(iters, 0.999999f, 1e-7f, static_cast<float*>(out));
            (void)hipEventRecord(b, nullptr);
            if ((e = hipGetLastError()) == hipSuccess && (e = hipEventSynchronize(b)) == hipSuccess)
                e = hipEventElapsedTime(&ms, a, b);
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(out);
    if (e != hipSuccess) return int(e);
    *tflops = double(blocks) * threads * iters * 8.0 * 2.0 / (double(ms) * 1e-3) / 1e12;
    return 0;
}
