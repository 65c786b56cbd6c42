// ars_probe.hip -- VALU issue-rate probes for the bench's VALU roofline
// (built into libars_probe.so, separate from the product library).
//
// gfx950 issues simple 32-bit VALU ops (v_xor_b32, v_and_b32, shifts,
// v_bitop3_b32) at about twice the rate of v_perm_b32, so the attainable
// instruction rate of a kernel depends on its share of v_perm.  These probes
// measure, on the device at hand:
//   out[0]  wave64 instructions / s of independent v_xor_b32 chains
//   out[1]  wave64 instructions / s of independent v_perm_b32 chains
//   out[2]  wave64 instructions / s of the fused kernels' butterfly
//           (a ^= c * b; b ^= a with the product's gmadd4) on
//           register-resident data, as the compiler schedules it
//   out[3]  the same butterflies in straight-line code: a loop body of
//           21 steps x 8 butterflies (~2k VALU instructions, the size of a
//           fused tile body) with a different table per step, as in a real
//           transform level
// Every kernel is pure ALU with fixed trip counts and one store per thread.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_device.hpp"

namespace {

constexpr int kBlocks = 256 * 16, kThreads = 256, kIters = 4000;

__global__ void __launch_bounds__(kThreads) k_xor(uint32_t *out) {
    uint32_t a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
    const uint32_t k = 0x07070707u ^ blockIdx.x;
    for (int i = 0; i < kIters; ++i)
        asm volatile(
            "v_xor_b32 %0, %0, %8\n v_xor_b32 %1, %1, %8\n v_xor_b32 %2, %2, %8\n v_xor_b32 %3, %3, %8\n"
            "v_xor_b32 %4, %4, %8\n v_xor_b32 %5, %5, %8\n v_xor_b32 %6, %6, %8\n v_xor_b32 %7, %7, %8\n"
            : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)
            : "v"(k));
    out[blockIdx.x * blockDim.x + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}

__global__ void __launch_bounds__(kThreads) k_perm(uint32_t *out) {
    uint32_t a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
    const uint32_t t0 = 0x03020100u, t1 = 0x07060504u;
    for (int i = 0; i < kIters; ++i)
        asm volatile(
            "v_perm_b32 %0, %8, %9, %0\n v_perm_b32 %1, %8, %9, %1\n v_perm_b32 %2, %8, %9, %2\n"
            "v_perm_b32 %3, %8, %9, %3\n v_perm_b32 %4, %8, %9, %4\n v_perm_b32 %5, %8, %9, %5\n"
            "v_perm_b32 %6, %8, %9, %6\n v_perm_b32 %7, %8, %9, %7\n"
            : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)
            : "v"(t1), "v"(t0));
    out[blockIdx.x * blockDim.x + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}

// 8 independent butterflies per iteration; the table is lane-varying, as in
// the tiles' first phase.
__global__ void __launch_bounds__(kThreads) k_butterfly(uint32_t *out) {
    uint32_t a[8], b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        a[j] = threadIdx.x * 0x01010101u + (uint32_t)j;
        b[j] = blockIdx.x * 0x00010001u + 3u * (uint32_t)j;
    }
    ars::MulTab m;
    m.w[0] = 0x03020100u + threadIdx.x;
    m.w[1] = 0x07060504u;
    m.w[2] = 0x11223344u;
    m.w[3] = 0x55667788u;
    m.w[4] = 0x99aabbccu ^ threadIdx.x;
    m.w[5] = 0u;
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            a[j] = ars::gmadd4(a[j], b[j], m);
            b[j] ^= a[j];
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) r ^= a[j] ^ b[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

constexpr int kSteps = 21, kOuter = 190;

__global__ void __launch_bounds__(kThreads) k_butterfly_long(uint32_t *out) {
    uint32_t a[8], b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        a[j] = threadIdx.x * 0x01010101u + (uint32_t)j;
        b[j] = blockIdx.x * 0x00010001u + 3u * (uint32_t)j;
    }
    ars::MulTab m;
    m.w[0] = 0x03020100u + threadIdx.x;
    m.w[1] = 0x07060504u;
    m.w[2] = 0x11223344u;
    m.w[3] = 0x55667788u;
    m.w[4] = 0x99aabbccu ^ threadIdx.x;
    m.w[5] = 0u;
    for (int o = 0; o < kOuter; ++o) {
#pragma unroll
        for (int u = 0; u < kSteps; ++u) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                a[j] = ars::gmadd4(a[j], b[j], m);
                b[j] ^= a[j];
            }
            m.w[0] += 0x01010101u;
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) r ^= a[j] ^ b[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

}  // namespace

// Wave64 VALU instructions per butterfly of k_butterfly's loop as compiled
// for gfx950 (3 v_perm_b32 and 9 other VALU ops).
constexpr double kButterflyInstrs = 12.0;
// k_butterfly_long as compiled for gfx950: 1869 VALU instructions (504
// v_perm_b32) per outer iteration of 21 steps x 8 butterflies.
constexpr double kLongStepInstrs = 1869.0 / 21.0;

// out must hold 4 doubles (see the list at the top of this file).
extern "C" int ars_probe_valu_rates(double *out) {
    if (!out) return -1;
    uint32_t *buf = nullptr;
    if (hipMalloc(&buf, (size_t)kBlocks * kThreads * 4) != hipSuccess) return -5;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = 0;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) rc = -5;
    const double waves = (double)kBlocks * kThreads / 64.0;
    for (int k = 0; k < 4 && rc == 0; ++k) {
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && rc == 0; ++rep) {  // the first launch warms up
            if (hipEventRecord(a) != hipSuccess) rc = -5;
            if (k == 0) hipLaunchKernelGGL(k_xor, dim3(kBlocks), dim3(kThreads), 0, 0, buf);
            if (k == 1) hipLaunchKernelGGL(k_perm, dim3(kBlocks), dim3(kThreads), 0, 0, buf);
            if (k == 2) hipLaunchKernelGGL(k_butterfly, dim3(kBlocks), dim3(kThreads), 0, 0, buf);
            if (k == 3) hipLaunchKernelGGL(k_butterfly_long, dim3(kBlocks), dim3(kThreads), 0, 0, buf);
            if (hipGetLastError() != hipSuccess || hipEventRecord(b) != hipSuccess ||
                hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess)
                rc = -5;
        }
        if (rc == 0) {
            const double instrs = k == 3 ? waves * kOuter * kSteps * kLongStepInstrs
                                         : waves * kIters * 8.0 * (k == 2 ? kButterflyInstrs : 1.0);
            out[k] = instrs / (ms * 1e-3);
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(buf);
    return rc;
}
