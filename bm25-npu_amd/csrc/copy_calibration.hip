// Plain device copy of a known size, used to calibrate the rocprofv3 HBM counters
// (FETCH_SIZE / WRITE_SIZE, KiB per dispatch) that the Search kernels are measured with:
// each launch of KCopy reads exactly `bytes` and writes exactly `bytes`.
//
// Usage: copy_calibration [MiB] [launches]   (defaults 256 MiB, 3 launches)
// Prints one JSON line with the per-launch times (HIP events) and achieved GB/s.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void KCopy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

#define CHECK(expr)                                                                                                    \
    do {                                                                                                               \
        const hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess) {                                                                                        \
            std::fprintf(stderr, "%s: %s\n", #expr, hipGetErrorString(e_));                                            \
            return 1;                                                                                                  \
        }                                                                                                              \
    } while (0)

int main(int argc, char **argv)
{
    const long mib = argc > 1 ? std::atol(argv[1]) : 256;
    const int launches = argc > 2 ? std::atoi(argv[2]) : 3;
    if (mib < 1 || mib > 16384 || launches < 1 || launches > 100) {
        std::fprintf(stderr, "usage: copy_calibration [MiB in 1..16384] [launches in 1..100]\n");
        return 2;
    }
    const size_t bytes = static_cast<size_t>(mib) << 20;
    const size_t n = bytes / sizeof(uint4);
    uint4 *src = nullptr, *dst = nullptr;
    CHECK(hipMalloc(&src, bytes));
    CHECK(hipMalloc(&dst, bytes));
    CHECK(hipMemset(src, 1, bytes));
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int grid = prop.multiProcessorCount * 8;
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    std::printf("{\"kernel\": \"KCopy\", \"read_kib\": %ld, \"write_kib\": %ld, \"ms\": [", mib * 1024, mib * 1024);
    float best = 1e30F;
    for (int i = 0; i < launches; ++i) {
        CHECK(hipEventRecord(a, nullptr));
        hipLaunchKernelGGL(KCopy, dim3(grid), dim3(256), 0, nullptr, src, dst, n);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(b, nullptr));
        CHECK(hipEventSynchronize(b));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, a, b));
        best = ms < best ? ms : best;
        std::printf("%s%.6f", i ? ", " : "", ms);
    }
    std::printf("], \"gbs\": %.1f}\n", 2.0 * static_cast<double>(bytes) / (best * 1e-3) / 1e9);
    CHECK(hipFree(src));
    CHECK(hipFree(dst));
    return 0;
}
