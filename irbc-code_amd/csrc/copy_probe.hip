/*
 * copy_probe.hip — plain device-to-device copy of a known size.
 *
 * Measures the achievable HBM copy bandwidth on the box and serves as the
 * calibration kernel for the HBM-traffic counters: a copy of B bytes must
 * read B and write B, so FETCH_SIZE and WRITE_SIZE of this kernel give the
 * counter-to-bytes factor used for the streaming kernel.
 *
 *   copy_probe [bytes=1073741824] [launches=20]
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHK(x)                                                                        \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));              \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

__global__ __launch_bounds__(256) void copy_f4(const float4 *__restrict__ a, float4 *__restrict__ b,
                                               size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char **argv)
{
    const size_t bytes = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : (1ull << 30);
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    const size_t n = bytes / sizeof(float4);
    if (n == 0 || reps <= 0) return 2;
    float4 *a = nullptr, *b = nullptr;
    CHK(hipMalloc(&a, n * sizeof(float4)));
    CHK(hipMalloc(&b, n * sizeof(float4)));
    CHK(hipMemset(a, 1, n * sizeof(float4)));
    CHK(hipMemset(b, 0, n * sizeof(float4)));
    const unsigned blocks = 256u * 16u;
    copy_f4<<<blocks, 256>>>(a, b, n);
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    CHK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) copy_f4<<<blocks, 256>>>(a, b, n);
    CHK(hipGetLastError());
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / reps;
    std::printf("{\"copy_bytes\": %zu, \"launches\": %d, \"ms_per_launch\": %.6f, "
                "\"GBps_read_plus_write\": %.1f}\n",
                n * sizeof(float4), reps, per, 2.0 * n * sizeof(float4) / (per * 1e-3) / 1e9);
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
