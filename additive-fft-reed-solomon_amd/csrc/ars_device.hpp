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

// The v_perm_b32 lookups of K multiplies (K <= 4) issued back to back at wave
// priority 3, then back to priority 0.  gfx950 issues one wave's v_perm_b32
// beside other waves' simple VALU ops (XOR, AND, shifts), but with equal
// priorities the waves of a SIMD, each at its own point of the butterfly,
// collide: the mix issues ~1.35x slower than its two kinds of instruction
// apart.  With the lookups in bursts at raised priority the others' simple
// ops fill in around them, and a register-only loop of these butterflies
// issues ~1.6x faster.  Inline assembly keeps the lookups together between
// the two priority changes (registers only, no memory access).
// s[3k .. 3k+2]: the table indices of multiply k on entry (gmul4's i0, i1,
// i2), its three partial products on exit.
template <int K>
__device__ __forceinline__ void perm_burst(uint32_t (&s)[3 * K], const MulTab (&m)[K]);
template <>
__device__ __forceinline__ void perm_burst<1>(uint32_t (&s)[3], const MulTab (&m)[1]) {
    asm volatile("s_setprio 3\n"
                 "v_perm_b32 %0, %4, %3, %0\n v_perm_b32 %1, %6, %5, %1\n v_perm_b32 %2, 0, %7, %2\n"
                 "s_setprio 0\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2])
                 : "v"(m[0].w[0]), "v"(m[0].w[1]), "v"(m[0].w[2]), "v"(m[0].w[3]), "v"(m[0].w[4]));
}
template <>
__device__ __forceinline__ void perm_burst<2>(uint32_t (&s)[6], const MulTab (&m)[2]) {
    asm volatile("s_setprio 3\n"
                 "v_perm_b32 %0, %7, %6, %0\n v_perm_b32 %1, %9, %8, %1\n v_perm_b32 %2, 0, %10, %2\n"
                 "v_perm_b32 %3, %12, %11, %3\n v_perm_b32 %4, %14, %13, %4\n v_perm_b32 %5, 0, %15, %5\n"
                 "s_setprio 0\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3]), "+v"(s[4]), "+v"(s[5])
                 : "v"(m[0].w[0]), "v"(m[0].w[1]), "v"(m[0].w[2]), "v"(m[0].w[3]), "v"(m[0].w[4]),
                   "v"(m[1].w[0]), "v"(m[1].w[1]), "v"(m[1].w[2]), "v"(m[1].w[3]), "v"(m[1].w[4]));
}
template <>
__device__ __forceinline__ void perm_burst<3>(uint32_t (&s)[9], const MulTab (&m)[3]) {
    asm volatile("s_setprio 3\n"
                 "v_perm_b32 %0, %10, %9, %0\n v_perm_b32 %1, %12, %11, %1\n v_perm_b32 %2, 0, %13, %2\n"
                 "v_perm_b32 %3, %15, %14, %3\n v_perm_b32 %4, %17, %16, %4\n v_perm_b32 %5, 0, %18, %5\n"
                 "v_perm_b32 %6, %20, %19, %6\n v_perm_b32 %7, %22, %21, %7\n v_perm_b32 %8, 0, %23, %8\n"
                 "s_setprio 0\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3]), "+v"(s[4]), "+v"(s[5]), "+v"(s[6]), "+v"(s[7]),
                   "+v"(s[8])
                 : "v"(m[0].w[0]), "v"(m[0].w[1]), "v"(m[0].w[2]), "v"(m[0].w[3]), "v"(m[0].w[4]),
                   "v"(m[1].w[0]), "v"(m[1].w[1]), "v"(m[1].w[2]), "v"(m[1].w[3]), "v"(m[1].w[4]),
                   "v"(m[2].w[0]), "v"(m[2].w[1]), "v"(m[2].w[2]), "v"(m[2].w[3]), "v"(m[2].w[4]));
}
template <>
__device__ __forceinline__ void perm_burst<4>(uint32_t (&s)[12], const MulTab (&m)[4]) {
    asm volatile("s_setprio 3\n"
                 "v_perm_b32 %0, %13, %12, %0\n v_perm_b32 %1, %15, %14, %1\n v_perm_b32 %2, 0, %16, %2\n"
                 "v_perm_b32 %3, %18, %17, %3\n v_perm_b32 %4, %20, %19, %4\n v_perm_b32 %5, 0, %21, %5\n"
                 "v_perm_b32 %6, %23, %22, %6\n v_perm_b32 %7, %25, %24, %7\n v_perm_b32 %8, 0, %26, %8\n"
                 "v_perm_b32 %9, %28, %27, %9\n v_perm_b32 %10, %30, %29, %10\n v_perm_b32 %11, 0, %31, %11\n"
                 "s_setprio 0\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3]), "+v"(s[4]), "+v"(s[5]), "+v"(s[6]), "+v"(s[7]),
                   "+v"(s[8]), "+v"(s[9]), "+v"(s[10]), "+v"(s[11])
                 : "v"(m[0].w[0]), "v"(m[0].w[1]), "v"(m[0].w[2]), "v"(m[0].w[3]), "v"(m[0].w[4]),
                   "v"(m[1].w[0]), "v"(m[1].w[1]), "v"(m[1].w[2]), "v"(m[1].w[3]), "v"(m[1].w[4]),
                   "v"(m[2].w[0]), "v"(m[2].w[1]), "v"(m[2].w[2]), "v"(m[2].w[3]), "v"(m[2].w[4]),
                   "v"(m[3].w[0]), "v"(m[3].w[1]), "v"(m[3].w[2]), "v"(m[3].w[3]), "v"(m[3].w[4]));
}

// The same lookups for tables read at a wave-uniform index (see UniTw): each
// table's high words and its T2 word are scalar operands (one SGPR per
// v_perm_b32 is allowed), only the two low words occupy VGPRs.
template <int K>
__device__ __forceinline__ void perm_burst_u(uint32_t (&s)[3 * K], const MulTab (&m)[K]);
template <>
__device__ __forceinline__ void perm_burst_u<1>(uint32_t (&s)[3], const MulTab (&m)[1]) {
    asm volatile("s_setprio 3\n"
                 "v_perm_b32 %0, %4, %3, %0\n v_perm_b32 %1, %6, %5, %1\n v_perm_b32 %2, 0, %7, %2\n"
                 "s_setprio 0\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2])
                 : "v"(m[0].w[0]), "s"(m[0].w[1]), "v"(m[0].w[2]), "s"(m[0].w[3]), "s"(m[0].w[4]));
}
template <>
__device__ __forceinline__ void perm_burst_u<2>(uint32_t (&s)[6], const MulTab (&m)[2]) {
    asm volatile("s_setprio 3\n"
                 "v_perm_b32 %0, %7, %6, %0\n v_perm_b32 %1, %9, %8, %1\n v_perm_b32 %2, 0, %10, %2\n"
                 "v_perm_b32 %3, %12, %11, %3\n v_perm_b32 %4, %14, %13, %4\n v_perm_b32 %5, 0, %15, %5\n"
                 "s_setprio 0\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3]), "+v"(s[4]), "+v"(s[5])
                 : "v"(m[0].w[0]), "s"(m[0].w[1]), "v"(m[0].w[2]), "s"(m[0].w[3]), "s"(m[0].w[4]),
                   "v"(m[1].w[0]), "s"(m[1].w[1]), "v"(m[1].w[2]), "s"(m[1].w[3]), "s"(m[1].w[4]));
}
template <>
__device__ __forceinline__ void perm_burst_u<3>(uint32_t (&s)[9], const MulTab (&m)[3]) {
    asm volatile("s_setprio 3\n"
                 "v_perm_b32 %0, %10, %9, %0\n v_perm_b32 %1, %12, %11, %1\n v_perm_b32 %2, 0, %13, %2\n"
                 "v_perm_b32 %3, %15, %14, %3\n v_perm_b32 %4, %17, %16, %4\n v_perm_b32 %5, 0, %18, %5\n"
                 "v_perm_b32 %6, %20, %19, %6\n v_perm_b32 %7, %22, %21, %7\n v_perm_b32 %8, 0, %23, %8\n"
                 "s_setprio 0\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3]), "+v"(s[4]), "+v"(s[5]), "+v"(s[6]), "+v"(s[7]),
                   "+v"(s[8])
                 : "v"(m[0].w[0]), "s"(m[0].w[1]), "v"(m[0].w[2]), "s"(m[0].w[3]), "s"(m[0].w[4]),
                   "v"(m[1].w[0]), "s"(m[1].w[1]), "v"(m[1].w[2]), "s"(m[1].w[3]), "s"(m[1].w[4]),
                   "v"(m[2].w[0]), "s"(m[2].w[1]), "v"(m[2].w[2]), "s"(m[2].w[3]), "s"(m[2].w[4]));
}
template <>
__device__ __forceinline__ void perm_burst_u<4>(uint32_t (&s)[12], const MulTab (&m)[4]) {
    asm volatile("s_setprio 3\n"
                 "v_perm_b32 %0, %13, %12, %0\n v_perm_b32 %1, %15, %14, %1\n v_perm_b32 %2, 0, %16, %2\n"
                 "v_perm_b32 %3, %18, %17, %3\n v_perm_b32 %4, %20, %19, %4\n v_perm_b32 %5, 0, %21, %5\n"
                 "v_perm_b32 %6, %23, %22, %6\n v_perm_b32 %7, %25, %24, %7\n v_perm_b32 %8, 0, %26, %8\n"
                 "v_perm_b32 %9, %28, %27, %9\n v_perm_b32 %10, %30, %29, %10\n v_perm_b32 %11, 0, %31, %11\n"
                 "s_setprio 0\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3]), "+v"(s[4]), "+v"(s[5]), "+v"(s[6]), "+v"(s[7]),
                   "+v"(s[8]), "+v"(s[9]), "+v"(s[10]), "+v"(s[11])
                 : "v"(m[0].w[0]), "s"(m[0].w[1]), "v"(m[0].w[2]), "s"(m[0].w[3]), "s"(m[0].w[4]),
                   "v"(m[1].w[0]), "s"(m[1].w[1]), "v"(m[1].w[2]), "s"(m[1].w[3]), "s"(m[1].w[4]),
                   "v"(m[2].w[0]), "s"(m[2].w[1]), "v"(m[2].w[2]), "s"(m[2].w[3]), "s"(m[2].w[4]),
                   "v"(m[3].w[0]), "s"(m[3].w[1]), "v"(m[3].w[2]), "s"(m[3].w[3]), "s"(m[3].w[4]));
}

// Where a butterfly level takes its twiddle tables from.
// LdsTw: the workgroup's LDS copy, at any (lane-varying) index.
struct LdsTw {
    const MulTab *t;
    static constexpr bool kUniform = false;
    __device__ __forceinline__ MulTab get(uint32_t i) const { return lds_tab(t, i); }
};
// UniTw: an index that is the same in every lane of the wave.  The two low
// words come from the LDS copy (8 bytes per lane instead of 20), the high
// words and T2 by scalar loads from the same table in global memory (t and g
// hold equal contents).  Only for wave-uniform indices: the high words enter
// perm_burst_u as scalar operands, so a lane-varying index would be wrong.
struct UniTw {
    const MulTab *t;
    const MulTab *__restrict__ g;
    static constexpr bool kUniform = true;
    __device__ __forceinline__ MulTab get(uint32_t i) const {
        const uint32_t *lp = reinterpret_cast<const uint32_t *>(t + i);
        const uint32_t *gp = reinterpret_cast<const uint32_t *>(g + i);
        MulTab m;
        m.w[0] = lp[0];
        m.w[2] = lp[2];
        m.w[1] = gp[1];
        m.w[3] = gp[3];
        m.w[4] = gp[4];
        m.w[5] = 0u;
        return m;
    }
};

// Multiplies per lookup burst in a butterfly level (the default; two where a
// kernel has no registers to spare).
constexpr int kBurstMuls = 4;

// The butterflies of a level that multiply (the rest have a compile-time zero
// twiddle), in register order; see level_pass for the indexing.
template <int NBF>
struct MulList {
    int count;
    int f[NBF];  // butterfly f: a = x[blk + i], b = x[blk + i + half], blk = (f / half) * 2 * half, i = f % half
};
template <int NBF>
constexpr MulList<NBF> level_muls(int half, int sh, int lv, bool bz) {
    MulList<NBF> r{};
    for (int f = 0; f < NBF; ++f) {
        const int blk = (f / half) * 2 * half;
        const uint32_t cpart = ((uint32_t)blk << sh) >> (lv + 1);
        if (!(bz && cpart == 0u)) r.f[r.count++] = f;
    }
    return r;
}

// Multiplies G0 .. G0 + K - 1 of a level (K = min(GRP, count - G0)):
// index prep (after INV's b ^= a), one lookup burst, the XOR combines (and
// FWD's b ^= a); then the next group.
template <int RL, int SH, int LV, bool INV, bool BZ, int GRP, int G0, class TS>
__device__ __forceinline__ void level_group(uint32_t *x, uint32_t tbase, const TS &tw) {
    constexpr int half = 1 << (LV - SH);
    constexpr int NBF = (1 << RL) / 2;
    constexpr MulList<NBF> ml = level_muls<NBF>(half, SH, LV, BZ);
    if constexpr (G0 < ml.count) {
        constexpr int K = (ml.count - G0) < GRP ? (ml.count - G0) : GRP;
        uint32_t s[3 * K];
        MulTab m[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int f = ml.f[G0 + k];
            const int blk = (f / half) * 2 * half, i = f % half;
            uint32_t &a = x[blk + i];
            uint32_t &b = x[blk + i + half];
            if constexpr (INV) b ^= a;
            m[k] = tw.get(tbase + ((((uint32_t)blk) << SH) >> (LV + 1)));
            s[3 * k] = b & 0x07070707u;
            s[3 * k + 1] = (b >> 3) & 0x07070707u;
            s[3 * k + 2] = (b >> 6) & 0x03030303u;
        }
        if constexpr (TS::kUniform)
            perm_burst_u<K>(s, m);
        else
            perm_burst<K>(s, m);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int f = ml.f[G0 + k];
            const int blk = (f / half) * 2 * half, i = f % half;
            uint32_t &a = x[blk + i];
            uint32_t &b = x[blk + i + half];
            a = xor3(a, s[3 * k], s[3 * k + 1]) ^ s[3 * k + 2];
            if constexpr (!INV) b ^= a;
        }
        level_group<RL, SH, LV, INV, BZ, GRP, G0 + K, TS>(x, tbase, tw);
    }
}

// One butterfly level over R = 2^RL registers.  Register j holds the point
// with absolute index  hb + (j << SH) + c,  where hb has zero bits below
// SH + RL and c < 2^SH never reaches a twiddle (LV >= SH).  The twiddle of
// the block whose first point has absolute index o is s_LV(omega_o): the
// recursion of rs/src/kernel/lut_kernel.rs:28-82 hands beta + v_{k-1} to the
// upper half, so a block at offset o sees beta + omega_o, and beta is itself
// omega_B for the shifts the codec uses (rs/src/codec.rs:257, 515).
// FWD: a ^= t*b; b ^= a   (rs/src/kernel/lut_kernel.rs:11-18)
// INV: b ^= a; a ^= t*b   (rs/src/kernel/lut_kernel.rs:20-26)
// BZ = hb is known to be zero: blocks at o = 0 have twiddle s_LV(0) = 0 and
// become XOR-only at compile time.  Otherwise every block multiplies (the
// zero twiddle's table is all-zero, so a lane whose block starts at 0 still
// computes the right value) and no lane-divergent branch is emitted.  The
// multiplies go in groups of GRP through perm_burst.
template <int RL, int SH, int LV, bool INV, bool BZ, int GRP = kBurstMuls, class TS = LdsTw>
__device__ __forceinline__ void level_pass(uint32_t *x, uint32_t hb, const TS &tw) {
    constexpr int lb = LV - SH;
    static_assert(lb >= 0 && lb < RL, "level outside the register range");
    constexpr int half = 1 << lb;
    const uint32_t tbase = (uint32_t)lvl_off(LV) + (BZ ? 0u : (hb >> (LV + 1)));
#pragma unroll
    for (int blk = 0; blk < (1 << RL); blk += 2 * half) {
        const uint32_t cpart = ((uint32_t)blk << SH) >> (LV + 1);
        if (BZ && cpart == 0u) {
#pragma unroll
            for (int i = 0; i < half; ++i) x[blk + i + half] ^= x[blk + i];
        }
    }
    level_group<RL, SH, LV, INV, BZ, GRP, 0, TS>(x, tbase, tw);
}

// Levels [L0, L1): ascending inverse butterflies (IFFT) or descending
// forward butterflies (FFT).
template <int RL, int SH, int L0, int L1, bool INV, bool BZ, int GRP = kBurstMuls, class TS = LdsTw>
__device__ __forceinline__ void run_levels_t(uint32_t *x, uint32_t hb, const TS &tw) {
    if constexpr (L0 < L1) {
        if constexpr (INV) {
            level_pass<RL, SH, L0, true, BZ, GRP, TS>(x, hb, tw);
            run_levels_t<RL, SH, L0 + 1, L1, true, BZ, GRP, TS>(x, hb, tw);
        } else {
            level_pass<RL, SH, L1 - 1, false, BZ, GRP, TS>(x, hb, tw);
            run_levels_t<RL, SH, L0, L1 - 1, false, BZ, GRP, TS>(x, hb, tw);
        }
    }
}
// ... with the tables from the workgroup's LDS copy at any index
template <int RL, int SH, int L0, int L1, bool INV, bool BZ, int GRP = kBurstMuls>
__device__ __forceinline__ void run_levels(uint32_t *x, uint32_t hb, const MulTab *tw) {
    run_levels_t<RL, SH, L0, L1, INV, BZ, GRP>(x, hb, LdsTw{tw});
}

// Shard rows of one buffer, addressed as (per-lane byte offset, uniform
// shard byte offset) through a buffer resource, so the per-shard offset is
// scalar.  AL: shard_len % 4 == 0 and a 4-byte aligned base, so a column quad
// is one dword load; lanes whose quad lies past the end of the shard read a
// clamped in-bounds quad (byte columns are independent, so the garbage never
// reaches a stored column) and store nothing.  !AL: byte loads and stores
// predicated on the nb valid bytes of the quad.
struct Rows {
    __amdgpu_buffer_rsrc_t r;
};

// mem ^= v for one dword through a buffer resource; the L2 performs the
// read-modify-write and nothing is returned (buffer_atomic_xor).  Bound to
// the LLVM intrinsic directly: clang has no builtin for the xor form.
__device__ int ars_buffer_atomic_xor(int v, __amdgpu_buffer_rsrc_t r, int voff, int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.atomic.xor.i32");

__device__ __forceinline__ Rows make_rows(const void *base, uint32_t bytes) {
    Rows w;
    w.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
    return w;
}

template <bool AL>
__device__ __forceinline__ uint32_t ldq(const Rows &w, uint32_t voff, uint32_t soff, uint32_t nb) {
    if constexpr (AL) {
        (void)nb;
        return __builtin_amdgcn_raw_buffer_load_b32(w.r, (int)voff, (int)soff, 0);
    } else {
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((uint32_t)i < nb)
                v |= (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(w.r, (int)(voff + i), (int)soff, 0) << (8 * i);
        return v;
    }
}

template <bool AL>
__device__ __forceinline__ void stq(const Rows &w, uint32_t voff, uint32_t soff, uint32_t v, uint32_t nb) {
    if constexpr (AL) {
        if (nb) __builtin_amdgcn_raw_buffer_store_b32(v, w.r, (int)voff, (int)soff, 0);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((uint32_t)i < nb)
                __builtin_amdgcn_raw_buffer_store_b8((uint8_t)(v >> (8 * i)), w.r, (int)(voff + i), (int)soff, 0);
    }
}

}  // namespace ars
