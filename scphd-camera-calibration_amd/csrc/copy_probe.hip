/*
 * copy_probe -- a plain device-to-device copy kernel of known size.
 *
 * Two uses: its achieved bandwidth is the practical HBM ceiling that the
 * filter kernels are compared with, and under `rocprofv3 --pmc FETCH_SIZE`
 * / `--pmc WRITE_SIZE` its counter readings, divided into the bytes it is
 * known to move, give the factor that turns those counters into HBM bytes
 * for the filter kernels.
 *
 * usage: copy_probe [MiB per buffer, default 1024] [repetitions, default 5]
 * Prints one line per repetition and a summary line.
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                      \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

int main(int argc, char** argv) {
    const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    if (mib <= 0 || mib > 65536 || reps <= 0 || reps > 1000) {
        std::fprintf(stderr, "usage: copy_probe [MiB 1..65536] [reps 1..1000]\n");
        return 2;
    }
    const size_t bytes = (size_t)mib << 20;
    const long n = (long)(bytes / sizeof(float4));
    float4 *src = nullptr, *dst = nullptr;
    CHECK(hipMalloc(&src, bytes));
    CHECK(hipMalloc(&dst, bytes));
    CHECK(hipMemset(src, 0x3c, bytes));
    CHECK(hipMemset(dst, 0, bytes));
    const long grid = (n + 255) / 256; /* one float4 per thread */
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    double best = 0, sum = 0;
    for (int r = 0; r < reps; ++r) {
        CHECK(hipEventRecord(a, 0));
        k_copy<<<(unsigned)grid, 256, 0, 0>>>(src, dst, n);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(b, 0));
        CHECK(hipEventSynchronize(b));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, a, b));
        const double gbs = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
        std::printf("copy rep %d: %zu B read + %zu B written in %.3f ms = %.1f GB/s\n", r, bytes, bytes, ms, gbs);
        best = gbs > best ? gbs : best;
        sum += gbs;
    }
    std::printf("copy summary: bytes_read_per_launch=%zu bytes_written_per_launch=%zu best_GBs=%.1f mean_GBs=%.1f\n",
                bytes, bytes, best, sum / reps);
    CHECK(hipFree(src));
    CHECK(hipFree(dst));
    return 0;
}
