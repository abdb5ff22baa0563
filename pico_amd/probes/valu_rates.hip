// VALU issue-rate probe for the opcodes the Poseidon2 and NTT kernels are
// made of.  Each thread runs 8 independent chains of one opcode (inline asm,
// so the instruction is exactly the one named), 4096 iterations, no memory
// traffic but one store per thread into a buffer sized for the grid.
// Prints lane-operations per second for each opcode; with 256 CUs x 4 SIMDs x
// 16 lanes x 2.4 GHz = 39.3 T lane-slots/s as the single-issue ceiling, the
// ratio tells which opcodes issue at half, single or double rate.
#include <hip/hip_runtime.h>

#include <cstdint>
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

constexpr int CHAINS = 8, ITERS = 4096;

template <int OP>
__global__ void __launch_bounds__(256) chains_kernel(uint32_t* out, uint32_t seed) {
    uint32_t x[CHAINS];
    uint64_t y[CHAINS];
    for (int c = 0; c < CHAINS; c++) {
        x[c] = seed + threadIdx.x * 7 + c;
        y[c] = x[c];
    }
    for (int it = 0; it < ITERS; it++) {
#pragma unroll
        for (int c = 0; c < CHAINS; c++) {
            if constexpr (OP == 0) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x[c]) : "v"(seed));
            if constexpr (OP == 1) asm volatile("v_sub_u32 %0, %0, %1" : "+v"(x[c]) : "v"(seed));
            if constexpr (OP == 2) asm volatile("v_mov_b32 %0, %1" : "=v"(x[c]) : "v"(x[(c + 1) % CHAINS]));
            if constexpr (OP == 3) asm volatile("v_min_u32 %0, %0, %1" : "+v"(x[c]) : "v"(seed));
            if constexpr (OP == 4) asm volatile("v_alignbit_b32 %0, %0, %1, 31" : "+v"(x[c]) : "v"(seed));
            if constexpr (OP == 5) asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(x[c]));
            if constexpr (OP == 6) asm volatile("v_and_b32 %0, %0, %1" : "+v"(x[c]) : "v"(seed));
            if constexpr (OP == 7) asm volatile("v_add3_u32 %0, %0, %1, %1" : "+v"(x[c]) : "v"(seed));
            if constexpr (OP == 8) asm volatile("v_mad_u32_u24 %0, %0, %1, %0" : "+v"(x[c]) : "v"(seed));
            if constexpr (OP == 9) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(x[c]) : "v"(seed));
            if constexpr (OP == 10) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(x[c]) : "v"(seed));
            if constexpr (OP == 11)
                asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(y[c]) : "v"(x[c]), "v"(seed) : "vcc");
            if constexpr (OP == 12) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(y[c]) : "v"(y[(c + 1) % CHAINS]));
            if constexpr (OP == 13) asm volatile("v_lshlrev_b64 %0, 1, %0" : "+v"(y[c]));
        }
    }
    uint32_t acc = 0;
    for (int c = 0; c < CHAINS; c++) acc += x[c] + (uint32_t)y[c];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

template <int OP>
static double run(uint32_t* d, int blocks) {
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    chains_kernel<OP><<<blocks, 256>>>(d, 3);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(a));
    chains_kernel<OP><<<blocks, 256>>>(d, 3);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
    return (double)blocks * 256 * ITERS * CHAINS / (ms * 1e-3) / 1e12;
}

int main() {
    const int blocks = 256 * 32;
    uint32_t* d = nullptr;
    CHECK(hipMalloc(&d, (size_t)blocks * 256 * sizeof(uint32_t)));
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, 0));
    std::printf("device %s CUs %d clock %d kHz\n", p.gcnArchName, p.multiProcessorCount, p.clockRate);
    const char* names[] = {"v_add_u32", "v_sub_u32", "v_mov_b32", "v_min_u32", "v_alignbit_b32",
                           "v_lshlrev_b32", "v_and_b32", "v_add3_u32", "v_mad_u32_u24", "v_mul_lo_u32",
                           "v_mul_hi_u32", "v_mad_u64_u32", "v_lshl_add_u64", "v_lshlrev_b64"};
    const double r[] = {run<0>(d, blocks), run<1>(d, blocks), run<2>(d, blocks),  run<3>(d, blocks),
                        run<4>(d, blocks), run<5>(d, blocks), run<6>(d, blocks),  run<7>(d, blocks),
                        run<8>(d, blocks), run<9>(d, blocks), run<10>(d, blocks), run<11>(d, blocks),
                        run<12>(d, blocks), run<13>(d, blocks)};
    for (int i = 0; i < 14; i++) std::printf("%-15s %6.2f T lane-ops/s\n", names[i], r[i]);
    CHECK(hipFree(d));
    return 0;
}
