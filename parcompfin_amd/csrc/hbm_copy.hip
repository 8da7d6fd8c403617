// hbm_copy.hip -- calibration program for HBM traffic counters.
//
// Copies a buffer of known size (default 1 GiB) with a plain coalesced
// 16-byte-per-lane kernel, so a counter run over it (FETCH_SIZE, WRITE_SIZE)
// reads what the counter reports for exactly `bytes` read and `bytes`
// written per dispatch.  The ratio counter / bytes is the scale applied to
// the same counter on the pricing kernels.  Prints one JSON line with the
// size, the dispatch count and the measured copy bandwidth.
//
//   hbm_copy [MiB] [reps]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_fill(float4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = make_float4((float)(i & 1023), 1.f, 2.f, 3.f);
}

#define CHECK(x)                                                                           \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "hbm_copy: %s: %s\n", #x, hipGetErrorString(e_));         \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    if (mib < 1 || mib > 65536 || reps < 1 || reps > 1000) {
        std::fprintf(stderr, "usage: hbm_copy [MiB 1..65536] [reps 1..1000]\n");
        return 2;
    }
    const size_t bytes = mib << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    int cus = 0;
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const unsigned grid = (unsigned)cus * 16;
    k_fill<<<grid, 256>>>(a, n);
    CHECK(hipGetLastError());
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    k_copy<<<grid, 256>>>(a, b, n);  // warm-up
    CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) k_copy<<<grid, 256>>>(a, b, n);
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipGetLastError());
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / reps;
    std::printf("{\"kernel\": \"k_copy\", \"bytes_read\": %zu, \"bytes_written\": %zu, \"dispatches\": %d, "
                "\"ms_per_copy\": %.6f, \"copy_gbs\": %.1f}\n",
                bytes, bytes, reps + 1, per, 2.0 * bytes / (per * 1e-3) / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
