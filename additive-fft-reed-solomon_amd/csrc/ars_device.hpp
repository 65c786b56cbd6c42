// ars_device.hpp -- gfx950 device primitives for the sharded additive FFT.
//
// Data representation: one lane owns a "column quad" -- 4 consecutive byte
// columns -- of every shard it works on, packed little-endian in one dword.
// Each byte column is an independent length-n transform (the reference's
// butterflies act element-wise on whole shards, rs/src/kernel/lut_kernel.rs:11-26),
// so all four bytes of a dword always see the same operation.
//
// GF(2^8) multiplication by a constant c uses three v_perm_b32 lookups into
// 8-entry byte tables (c*x is GF(2)-linear in x, so
// c*x = c*(x & 7) ^ c*(x & 0x38) ^ c*(x & 0xc0)); see MulTab in ars_gf.hpp.
// That is 10 VALU ops per 4 byte-products, with the tables as plain data, so
// run-time twiddles cost the same as compile-time ones.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_gf.hpp"

namespace ars {

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// c * x for the 4 packed bytes of x, c given by its MulTab.
__device__ __forceinline__ uint32_t gmul4(uint32_t x, const MulTab &m) {
    const uint32_t i0 = x & 0x07070707u;
    const uint32_t i1 = (x >> 3) & 0x07070707u;
    const uint32_t i2 = (x >> 6) & 0x03030303u;
    const uint32_t p0 = __builtin_amdgcn_perm(m.w[1], m.w[0], i0);
    const uint32_t p1 = __builtin_amdgcn_perm(m.w[3], m.w[2], i1);
    const uint32_t p2 = __builtin_amdgcn_perm(0u, m.w[4], i2);
    return xor3(p0, p1, p2);
}

// a ^ c * x
__device__ __forceinline__ uint32_t gmadd4(uint32_t a, uint32_t x, const MulTab &m) {
    const uint32_t i0 = x & 0x07070707u;
    const uint32_t i1 = (x >> 3) & 0x07070707u;
    const uint32_t i2 = (x >> 6) & 0x03030303u;
    const uint32_t p0 = __builtin_amdgcn_perm(m.w[1], m.w[0], i0);
    const uint32_t p1 = __builtin_amdgcn_perm(m.w[3], m.w[2], i1);
    const uint32_t p2 = __builtin_amdgcn_perm(0u, m.w[4], i2);
    return xor3(a, p0, p1) ^ p2;
}

__device__ __forceinline__ MulTab lds_tab(const MulTab *t, uint32_t i) {
    const uint2 *p = reinterpret_cast<const uint2 *>(t + i);
    const uint2 a = p[0], b = p[1], c = p[2];
    MulTab m;
    m.w[0] = a.x; m.w[1] = a.y; m.w[2] = b.x; m.w[3] = b.y; m.w[4] = c.x; m.w[5] = c.y;
    return m;
}

// One butterfly level over R = 2^RL registers.  Register j holds the point
// with absolute index  base + (j << SH)  (base has zero bits in
// [SH, SH + RL)); LV is the absolute level.  The twiddle of the block whose
// first point has absolute index o is s_LV(omega_o) (LNH Alg. 1/2 unrolled:
// the recursion of rs/src/kernel/lut_kernel.rs:28-82 passes beta + v_{k-1}
// to the upper half, so a block at offset o sees beta + omega_o).
// FWD: a ^= t*b; b ^= a   (rs/src/kernel/lut_kernel.rs:11-18)
// INV: b ^= a; a ^= t*b   (rs/src/kernel/lut_kernel.rs:20-26)
template <int RL, int SH, int LV, bool INV>
__device__ __forceinline__ void level_pass(uint32_t *x, uint32_t base, const MulTab *tw) {
    constexpr int lb = LV - SH;
    static_assert(lb >= 0 && lb < RL, "level outside the register range");
    constexpr int half = 1 << lb;
#pragma unroll
    for (int blk = 0; blk < (1 << RL); blk += 2 * half) {
        const uint32_t o = (base | ((uint32_t)blk << SH)) & ~((2u << LV) - 1u);
        const uint32_t ti = o >> (LV + 1);
        if (ti != 0u) {
            const MulTab m = lds_tab(tw, (uint32_t)lvl_off(LV) + ti);
#pragma unroll
            for (int i = 0; i < half; ++i) {
                uint32_t &a = x[blk + i];
                uint32_t &b = x[blk + i + half];
                if constexpr (INV) {
                    b ^= a;
                    a = gmadd4(a, b, m);
                } else {
                    a = gmadd4(a, b, m);
                    b ^= a;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < half; ++i) x[blk + i + half] ^= x[blk + i];
        }
    }
}

// Levels [L0, L1): ascending inverse butterflies (IFFT) or descending
// forward butterflies (FFT).
template <int RL, int SH, int L0, int L1, bool INV>
__device__ __forceinline__ void run_levels(uint32_t *x, uint32_t base, const MulTab *tw) {
    if constexpr (L0 < L1) {
        if constexpr (INV) {
            level_pass<RL, SH, L0, true>(x, base, tw);
            run_levels<RL, SH, L0 + 1, L1, true>(x, base, tw);
        } else {
            level_pass<RL, SH, L1 - 1, false>(x, base, tw);
            run_levels<RL, SH, L0, L1 - 1, false>(x, base, tw);
        }
    }
}

// Load / store one column quad of one shard.  nb = valid bytes (0..4).
template <bool AL>
__device__ __forceinline__ uint32_t ldq(const uint8_t *__restrict__ p, uint32_t nb) {
    if constexpr (AL) {
        return nb ? *reinterpret_cast<const uint32_t *>(p) : 0u;
    } else {
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((uint32_t)i < nb) v |= (uint32_t)p[i] << (8 * i);
        return v;
    }
}

template <bool AL>
__device__ __forceinline__ void stq(uint8_t *__restrict__ p, uint32_t v, uint32_t nb) {
    if constexpr (AL) {
        if (nb) *reinterpret_cast<uint32_t *>(p) = v;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((uint32_t)i < nb) p[i] = (uint8_t)(v >> (8 * i));
    }
}

}  // namespace ars
