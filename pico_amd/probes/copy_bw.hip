// Calibration kernel for the HBM traffic counters.  One launch reads exactly
// 1 GiB and writes exactly 1 GiB (one 16-byte load and store per thread, grid
// sized to the buffer), so FETCH_SIZE / WRITE_SIZE of this dispatch divided
// by 1 GiB is the factor by which each counter reports true bytes on this
// device.  Also prints the copy's achieved bandwidth.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            std::exit(1);                                                         \
        }                                                                         \
    } while (0)

__global__ void copy_kernel(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

int main() {
    const size_t bytes = 1ull << 30, n = bytes / 16;
    const unsigned grid = (unsigned)((n + 255) / 256);
    uint4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    CHECK(hipDeviceSynchronize());
    copy_kernel<<<grid, 256>>>(a, b, n);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    CHECK(hipEventRecord(e0));
    copy_kernel<<<grid, 256>>>(a, b, n);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("copy 1 GiB -> 1 GiB: %.3f ms, %.1f GB/s (read + write)\n", ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
