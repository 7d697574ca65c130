// calibrate.hip -- a plain device-to-device copy kernel of known size.  Its
// rocprofv3 FETCH_SIZE / WRITE_SIZE readings calibrate the HBM-traffic
// counters that the profiles of the hot-path kernels are scaled by.
#include "common.hpp"

namespace
{
    __global__ __launch_bounds__(256) void CopyKernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                      int64_t n)
    {
        int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
        for (; i < n; i += stride)
        {
            dst[i] = src[i];
        }
    }
}

extern "C" int halhip_calibration_copy(halhip_context *ctx, int64_t bytes, int32_t repeats, float *msPerCopy)
{
    using namespace halhip;
    return Guard(
        [&]
        {
            Require(ctx != nullptr, "halhip_calibration_copy: null context");
            Require(bytes > 0 && bytes % 16 == 0 && repeats > 0, "halhip_calibration_copy: bad size");
            Bind(ctx);
            DeviceArray<float4> a;
            DeviceArray<float4> b;
            const int64_t n = bytes / 16;
            a.Allocate(ctx, static_cast<size_t>(n));
            b.Allocate(ctx, static_cast<size_t>(n));
            Check(hipMemset(a.get(), 0, static_cast<size_t>(bytes)), "hipMemset");
            hipEvent_t e0;
            hipEvent_t e1;
            Check(hipEventCreate(&e0), "hipEventCreate");
            Check(hipEventCreate(&e1), "hipEventCreate");
            const unsigned blocks = 256 * 8 * 4;
            // one untimed launch first: the first touch of fresh allocations
            // is slower than steady state
            hipLaunchKernelGGL(CopyKernel, dim3(blocks), dim3(256), 0, nullptr, a.get(), b.get(), n);
            Check(hipGetLastError(), "CopyKernel launch");
            Check(hipEventRecord(e0, nullptr), "hipEventRecord");
            for (int32_t r = 0; r < repeats; r++)
            {
                hipLaunchKernelGGL(CopyKernel, dim3(blocks), dim3(256), 0, nullptr, a.get(), b.get(), n);
                Check(hipGetLastError(), "CopyKernel launch");
            }
            Check(hipEventRecord(e1, nullptr), "hipEventRecord");
            Check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.0f;
            Check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (msPerCopy != nullptr)
            {
                *msPerCopy = ms / static_cast<float>(repeats);
            }
        });
}
