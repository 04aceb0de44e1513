// hbm_copy -- a plain device-to-device copy of a known size, for calibrating
// HBM traffic counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): the
// counters read for this kernel, divided by the bytes it must move, give the
// factor that turns the same counters read for an engine kernel into bytes.
//
//   hbm_copy [bytes] [repeats]      (defaults: 1 GiB, 3)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_fill(float4* __restrict__ p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4((float)i, 1.f, 2.f, 3.f);
}

#define CHECK(x)                                                              \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) {                                               \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));      \
            return 1;                                                         \
        }                                                                     \
    } while (0)

int main(int argc, char** argv)
{
    const size_t bytes = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : ((size_t)1 << 30);
    const int reps = argc > 2 ? std::atoi(argv[2]) : 3;
    const size_t n = bytes / sizeof(float4);
    if (n == 0 || reps <= 0) {
        std::fprintf(stderr, "usage: hbm_copy [bytes >= 16] [repeats > 0]\n");
        return 2;
    }
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, n * sizeof(float4)));
    CHECK(hipMalloc(&b, n * sizeof(float4)));
    hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, 0, a, n);
    CHECK(hipGetLastError());
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int r = 0; r < reps; ++r) {
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, 0, a, b, n);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("copy %zu bytes: %.3f ms, %.1f GB/s (read + write)\n", n * sizeof(float4),
                    ms, 2.0 * n * sizeof(float4) / (ms * 1e-3) / 1e9);
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
