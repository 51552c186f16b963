// hbm_copy.hip -- calibration program for HBM traffic counters.
//
// Copies a buffer of a known size with a plain streaming kernel, so that the FETCH_SIZE and
// WRITE_SIZE counters of a profiler run can be compared with the bytes the kernel must move
// (size read + size written per launch).  The ratio is the scale factor applied to the same
// counters read on the sampler's kernels.
//
//   hbm_copy [MiB] [launches]      default 1024 MiB, 4 launches
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

#define CHECK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
            return 1;                                                                \
        }                                                                            \
    } while (0)

int main(int argc, char** argv) {
    size_t mib = argc > 1 ? std::strtoul(argv[1], nullptr, 10) : 1024;
    int launches = argc > 2 ? std::atoi(argv[2]) : 4;
    if (mib < 1 || mib > 16384 || launches < 1 || launches > 100) {
        std::fprintf(stderr, "size 1..16384 MiB, launches 1..100\n");
        return 2;
    }
    size_t bytes = mib << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    CHECK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    int blocks = 256 * 8 * 4;
    k_copy<<<blocks, 256>>>(a, b, n);  // warm-up
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e0));
    for (int i = 0; i < launches; ++i) k_copy<<<blocks, 256>>>(a, b, n);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    double per = ms / launches;
    std::printf("{\"kernel\": \"k_copy\", \"bytes_read\": %zu, \"bytes_written\": %zu, \"launches\": %d, "
                "\"ms_per_launch\": %.4f, \"GBps\": %.1f}\n",
                bytes, bytes, launches + 1, per, 2.0 * bytes / (per * 1e-3) / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
