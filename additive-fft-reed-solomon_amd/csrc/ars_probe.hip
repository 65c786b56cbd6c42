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
//   out[4]  out[3]'s butterflies with the lookups in bursts at raised wave
//           priority (perm_burst), as the tiles issue them
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

// k_butterfly_long's butterflies with the table lookups in bursts at raised
// wave priority, as the tiles issue them (perm_burst: per group of four
// multiplies, index prep, the twelve v_perm_b32 at priority 3, the combines).
__global__ void __launch_bounds__(kThreads) k_butterfly_burst(uint32_t *out) {
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
            for (int g = 0; g < 8; g += 4) {
                uint32_t sel[12];
                const ars::MulTab mt[4] = {m, m, m, m};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    sel[3 * k] = b[g + k] & 0x07070707u;
                    sel[3 * k + 1] = (b[g + k] >> 3) & 0x07070707u;
                    sel[3 * k + 2] = (b[g + k] >> 6) & 0x03030303u;
                }
                ars::perm_burst<4>(sel, mt);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    a[g + k] = ars::xor3(a[g + k], sel[3 * k], sel[3 * k + 1]) ^ sel[3 * k + 2];
                    b[g + k] ^= a[g + k];
                }
            }
            m.w[0] += 0x01010101u;
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) r ^= a[j] ^ b[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

// ---------------------------------------------------------------------------
// Phase B of the 256-point recover in two representations, for comparison.
//
// Phase B is the part of the recover tile whose twiddles are compile-time
// constants: on the 16 rows h*16 + q of one shard group (beta = 0) it runs
// IFFT levels 4..7, the high-bit derivative terms (Dhi), then FFT levels
// 7..4.  The block offsets of levels >= 4 have zero low 4 bits, so the map is
// the same for every q.
//   * k_phaseb_bytes: the product's form.  A lane holds one column quad of the
//     16 rows (16 dwords) and multiplies with three v_perm_b32 lookups into
//     LDS-staged tables (level_pass).
//   * k_phaseb_bits: bit-sliced.  A lane holds 32 columns of the 16 rows
//     (128 dwords), turns each row's 8 dwords into 8 bit planes (an 8x8 bit
//     transpose per byte lane), multiplies by XOR networks fixed at compile
//     time, and transposes back.  No v_perm at all.
// Each launch applies the map kPhaseBReps times in registers (the bit-sliced
// form transposes in and out every time, as a fused kernel would), so the
// arithmetic, not HBM, sets the time.  Both read 16 rows of L bytes and
// write 16 rows; the outputs must be identical.
constexpr int kPhaseBReps = 4;

// Compile-time field and Cantor-basis tables (the definitions of ars_gf.hpp).
struct CGf {
    uint8_t basis[8];
    uint8_t sub[256];
    uint8_t spv[9][256];
};
constexpr uint8_t cmul(uint32_t a, uint32_t b) {
    uint32_t x = a, r = 0;
    for (int i = 0; i < 8; ++i) {
        if (b & (1u << i)) r ^= x;
        x <<= 1;
        if (x & 0x100u) x ^= 0x11du;
    }
    return (uint8_t)r;
}
constexpr int ctrace(uint8_t x) {
    uint8_t t = x, s = x;
    for (int i = 0; i < 7; ++i) {
        t = cmul(t, t);
        s ^= t;
    }
    return s & 1;
}
constexpr CGf make_cgf() {
    CGf g{};
    g.basis[0] = 1;
    for (int i = 1; i < 8; ++i) {
        uint8_t root = 0;
        for (int r = 0; r < 256; ++r)
            if ((uint8_t)(cmul((uint32_t)r, (uint32_t)r) ^ r) == g.basis[i - 1]) {
                root = (uint8_t)r;
                break;
            }
        g.basis[i] = ctrace(root) ? (uint8_t)(root ^ 1) : root;
    }
    for (int i = 0; i < 256; ++i) {
        uint8_t p = 0;
        for (int b = 0; b < 8; ++b)
            if ((i >> b) & 1) p ^= g.basis[b];
        g.sub[i] = p;
    }
    for (int i = 0; i < 256; ++i) g.spv[0][i] = (uint8_t)i;
    for (int j = 0; j < 8; ++j) {
        const uint8_t bj = g.spv[j][g.basis[j]];
        for (int i = 0; i < 256; ++i) g.spv[j + 1][i] = cmul(g.spv[j][i], (uint8_t)(g.spv[j][i] ^ bj));
    }
    return g;
}
constexpr CGf kCgf = make_cgf();

// Twiddle of level LV (4..7) for the block whose first register is BLK.
template <int LV, int BLK>
constexpr uint8_t phaseb_twiddle() {
    return kCgf.spv[LV][kCgf.sub[BLK * 16]];
}

// a ^= c * b on 8 bit planes: plane p of c*b is the XOR of the planes j of b
// with bit p of c * 2^j set.  The terms are folded in pairs with 3-input XORs.
template <uint8_t C, int P, int J = 0>
__device__ __forceinline__ uint32_t bs_plane(uint32_t acc, const uint32_t (&b)[8]) {
    if constexpr (J >= 8) {
        return acc;
    } else if constexpr (!((cmul(C, 1u << J) >> P) & 1u)) {
        return bs_plane<C, P, J + 1>(acc, b);
    } else {
        constexpr int K = [] {
            int k = J + 1;
            while (k < 8 && !((cmul(C, 1u << k) >> P) & 1u)) ++k;
            return k;
        }();
        if constexpr (K >= 8)
            return acc ^ b[J];
        else
            return bs_plane<C, P, K + 1>(ars::xor3(acc, b[J], b[K]), b);
    }
}
template <uint8_t C>
__device__ __forceinline__ void bs_madd(uint32_t (&a)[8], const uint32_t (&b)[8]) {
    a[0] = bs_plane<C, 0>(a[0], b);
    a[1] = bs_plane<C, 1>(a[1], b);
    a[2] = bs_plane<C, 2>(a[2], b);
    a[3] = bs_plane<C, 3>(a[3], b);
    a[4] = bs_plane<C, 4>(a[4], b);
    a[5] = bs_plane<C, 5>(a[5], b);
    a[6] = bs_plane<C, 6>(a[6], b);
    a[7] = bs_plane<C, 7>(a[7], b);
}

template <int LV, bool INV, int BLK>
__device__ __forceinline__ void bs_blocks(uint32_t (&y)[16][8]) {
    constexpr int half = 1 << (LV - 4);
    if constexpr (BLK < 16) {
        constexpr uint8_t t = phaseb_twiddle<LV, BLK>();
#pragma unroll
        for (int i = 0; i < half; ++i) {
            uint32_t(&a)[8] = y[BLK + i];
            uint32_t(&b)[8] = y[BLK + i + half];
            if constexpr (INV) {
#pragma unroll
                for (int p = 0; p < 8; ++p) b[p] ^= a[p];
                if constexpr (t != 0) bs_madd<t>(a, b);
            } else {
                if constexpr (t != 0) bs_madd<t>(a, b);
#pragma unroll
                for (int p = 0; p < 8; ++p) b[p] ^= a[p];
            }
        }
        bs_blocks<LV, INV, BLK + 2 * half>(y);
    }
}

// 8x8 bit transpose inside each byte lane of 8 words: afterwards word b holds,
// in byte k, bit b of byte k of words 0..7 (bit i from word i).  An involution.
// Each exchange of bit blocks between two words is two shifts and two
// bit selects.
__host__ __device__ __forceinline__ uint32_t bsel(uint32_t m, uint32_t a, uint32_t b) { return (a & m) | (b & ~m); }
__host__ __device__ __forceinline__ void swap_blocks(uint32_t &x, uint32_t &y, int s, uint32_t m) {
    const uint32_t ny = bsel(m, x >> s, y), nx = bsel(m << s, y << s, x);
    x = nx;
    y = ny;
}
__host__ __device__ __forceinline__ void transpose8(uint32_t (&w)[8]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) swap_blocks(w[i], w[i + 4], 4, 0x0f0f0f0fu);
#pragma unroll
    for (int i = 0; i < 8; i += 4)
#pragma unroll
        for (int j = 0; j < 2; ++j) swap_blocks(w[i + j], w[i + j + 2], 2, 0x33333333u);
#pragma unroll
    for (int i = 0; i < 8; i += 2) swap_blocks(w[i], w[i + 1], 1, 0x55555555u);
}

// Dhi in place: y[h] = xor over bits b < 4 clear in h of y[h | 2^b]
// (ascending h reads only higher, still untouched, entries)
template <class V>
__device__ __forceinline__ void dhi16(V (&y)[16]) {
#pragma unroll
    for (int h = 0; h < 16; ++h) {
        V a{};
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (!((h >> b) & 1)) {
                if constexpr (sizeof(V) == 4) {
                    a ^= y[h | (1 << b)];
                } else {
#pragma unroll
                    for (int p = 0; p < 8; ++p) a[p] ^= y[h | (1 << b)][p];
                }
            }
        if constexpr (sizeof(V) == 4) {
            y[h] = a;
        } else {
#pragma unroll
            for (int p = 0; p < 8; ++p) y[h][p] = a[p];
        }
    }
}

struct Planes {
    uint32_t p[8];
    __device__ uint32_t &operator[](int i) { return p[i]; }
    __device__ const uint32_t &operator[](int i) const { return p[i]; }
};

__global__ void __launch_bounds__(256) k_phaseb_bytes(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                       uint32_t L, const ars::MulTab *__restrict__ g_tw) {
    __shared__ ars::MulTab tw[ars::kTwiddleEntries];
    for (uint32_t i = threadIdx.x; i < (uint32_t)ars::kTwiddleEntries * 6u; i += blockDim.x)
        reinterpret_cast<uint32_t *>(tw)[i] = reinterpret_cast<const uint32_t *>(g_tw)[i];
    __syncthreads();
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;  // column quad
    if (c * 4u >= L) return;
    uint32_t y[16];
#pragma unroll
    for (int h = 0; h < 16; ++h) y[h] = *reinterpret_cast<const uint32_t *>(in + (size_t)h * L + c * 4u);
    for (int r = 0; r < kPhaseBReps; ++r) {
        ars::run_levels<4, 4, 4, 8, true, true>(y, 0u, tw);
        dhi16(y);
        ars::run_levels<4, 4, 4, 8, false, true>(y, 0u, tw);
    }
#pragma unroll
    for (int h = 0; h < 16; ++h) *reinterpret_cast<uint32_t *>(out + (size_t)h * L + c * 4u) = y[h];
}

__global__ void __launch_bounds__(256) k_phaseb_bits(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                      uint32_t L) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;  // 32-column group
    if (c * 32u >= L) return;
    uint32_t y[16][8];
#pragma unroll
    for (int h = 0; h < 16; ++h) {
        const uint4 *p = reinterpret_cast<const uint4 *>(in + (size_t)h * L + c * 32u);
        const uint4 a = p[0], b = p[1];
        y[h][0] = a.x; y[h][1] = a.y; y[h][2] = a.z; y[h][3] = a.w;
        y[h][4] = b.x; y[h][5] = b.y; y[h][6] = b.z; y[h][7] = b.w;
    }
    for (int r = 0; r < kPhaseBReps; ++r) {
#pragma unroll
        for (int h = 0; h < 16; ++h) transpose8(y[h]);
        bs_blocks<4, true, 0>(y);
        bs_blocks<5, true, 0>(y);
        bs_blocks<6, true, 0>(y);
        bs_blocks<7, true, 0>(y);
        dhi16(reinterpret_cast<Planes(&)[16]>(y));
        bs_blocks<7, false, 0>(y);
        bs_blocks<6, false, 0>(y);
        bs_blocks<5, false, 0>(y);
        bs_blocks<4, false, 0>(y);
#pragma unroll
        for (int h = 0; h < 16; ++h) transpose8(y[h]);
    }
#pragma unroll
    for (int h = 0; h < 16; ++h) {
        uint4 *p = reinterpret_cast<uint4 *>(out + (size_t)h * L + c * 32u);
        p[0] = make_uint4(y[h][0], y[h][1], y[h][2], y[h][3]);
        p[1] = make_uint4(y[h][4], y[h][5], y[h][6], y[h][7]);
    }
}

}  // namespace

// Wave64 VALU instructions per butterfly of k_butterfly's loop as compiled
// for gfx950 (3 v_perm_b32 and 9 other VALU ops).
constexpr double kButterflyInstrs = 12.0;
// k_butterfly_long as compiled for gfx950: 1869 VALU instructions (504
// v_perm_b32) per outer iteration of 21 steps x 8 butterflies; the same for
// k_butterfly_burst.
constexpr double kLongStepInstrs = 1869.0 / 21.0;

// out must hold 5 doubles (see the list at the top of this file).
extern "C" int ars_probe_valu_rates(double *out) {
    if (!out) return -1;
    uint32_t *buf = nullptr;
    if (hipMalloc(&buf, (size_t)kBlocks * kThreads * 4) != hipSuccess) return -5;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = 0;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) rc = -5;
    const double waves = (double)kBlocks * kThreads / 64.0;
    for (int k = 0; k < 5 && rc == 0; ++k) {
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && rc == 0; ++rep) {  // the first launch warms up
            if (hipEventRecord(a) != hipSuccess) rc = -5;
            if (k == 0) hipLaunchKernelGGL(k_xor, dim3(kBlocks), dim3(kThreads), 0, 0, buf);
            if (k == 1) hipLaunchKernelGGL(k_perm, dim3(kBlocks), dim3(kThreads), 0, 0, buf);
            if (k == 2) hipLaunchKernelGGL(k_butterfly, dim3(kBlocks), dim3(kThreads), 0, 0, buf);
            if (k == 3) hipLaunchKernelGGL(k_butterfly_long, dim3(kBlocks), dim3(kThreads), 0, 0, buf);
            if (k == 4) hipLaunchKernelGGL(k_butterfly_burst, dim3(kBlocks), dim3(kThreads), 0, 0, buf);
            if (hipGetLastError() != hipSuccess || hipEventRecord(b) != hipSuccess ||
                hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess)
                rc = -5;
        }
        if (rc == 0) {
            const double instrs = k >= 3 ? waves * kOuter * kSteps * kLongStepInstrs
                                         : waves * kIters * 8.0 * (k == 2 ? kButterflyInstrs : 1.0);
            out[k] = instrs / (ms * 1e-3);
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(buf);
    return rc;
}

// Maps of phase B applied per launch of the phase-B probes.
extern "C" int ars_probe_phaseb_maps(void) { return kPhaseBReps; }

// Phase-B comparison (see k_phaseb_bytes / k_phaseb_bits): variant 0 = bytes
// (v_perm), 1 = bit-sliced.  d_in and d_out hold 16 rows of L bytes each
// (L a multiple of 32, 16 * L < 2^32, 16-byte aligned).  Runs `launches`
// back-to-back launches after one warm-up and returns the mean time of one
// launch in *us; the output is that of one launch (kPhaseBReps maps applied).
extern "C" int ars_probe_phaseb(int variant, const uint8_t *d_in, uint8_t *d_out, uint32_t L, int launches,
                                double *us) {
    if (!d_in || !d_out || !us || L == 0 || L % 32 || (uint64_t)L * 16 >= (1ull << 32) || launches < 1 ||
        (variant != 0 && variant != 1))
        return -1;
    static ars::MulTab *d_tw = nullptr;
    if (!d_tw) {
        const ars::GfTables &g = ars::gf();
        ars::MulTab tw[ars::kTwiddleEntries];
        for (int l = 0; l < 8; ++l)
            for (int j = 0; j < (128 >> l); ++j) tw[ars::lvl_off(l) + j] = g.multab(g.spv[l][g.sub[j << (l + 1)]]);
        if (hipMalloc(&d_tw, sizeof tw) != hipSuccess) return -5;
        if (hipMemcpy(d_tw, tw, sizeof tw, hipMemcpyHostToDevice) != hipSuccess) return -5;
    }
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -5;
    auto launch = [&] {
        if (variant == 0)
            hipLaunchKernelGGL(k_phaseb_bytes, dim3((L / 4 + 255) / 256), dim3(256), 0, 0, d_in, d_out, L, d_tw);
        else
            hipLaunchKernelGGL(k_phaseb_bits, dim3((L / 32 + 255) / 256), dim3(256), 0, 0, d_in, d_out, L);
    };
    int rc = 0;
    launch();
    float ms = 0.0f;
    if (hipGetLastError() != hipSuccess || hipEventRecord(a) != hipSuccess) rc = -5;
    for (int i = 0; i < launches && rc == 0; ++i) launch();
    if (rc == 0 && (hipGetLastError() != hipSuccess || hipEventRecord(b) != hipSuccess ||
                    hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess))
        rc = -5;
    if (rc == 0) *us = ms * 1e3 / launches;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}
