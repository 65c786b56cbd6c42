// ars_tiles.hpp -- the fused encode / recover / recompute tile bodies, shared
// by the one-codeword-per-launch kernels (ars_kernels.hip) and the batched
// persistent kernel (ars_batch.hip).
//
// Work decomposition.  A workgroup owns a tile of PQ column quads (4*PQ byte
// columns) of every shard of one transform.  For a transform of 2^TL points:
//   TL <= 4 : each lane keeps all 2^TL shards of its quad in registers and
//             runs every butterfly level there (no LDS traffic).
//   TL >= 5 : Q = 2^(TL-4) lanes share a quad.  Phase A: lane q holds the 16
//             consecutive shards q*16..q*16+15 and runs levels 0..3 in
//             registers.  One LDS exchange regroups the tile so lane q holds
//             shards {h*16 + c : c in its set of 16/Q}; phase B runs levels
//             4..TL-1 in registers.  The FFT walks the same two phases in the
//             opposite order, so a transform costs one LDS exchange.
// Point indices are absolute (they include the transform's shift beta =
// omega_B), so every transform of every code reads the same 255-entry
// twiddle table (see level_pass).
//
// The bodies are templated on an I/O policy:
//   io.ld_in(g, j)   quad of input row g + j   (g lane-varying, j compile-time)
//   io.ld_out(g, j)  quad of output row g + j  (encode accumulate only)
//   io.st_out(g, j, v)
// and, for recover, on a per-shard multiplier source:
//   tabs.tab(s) -> MulTab, tabs.erased(s) -> bool.
// Columns are independent, so lanes may even belong to different codewords
// as long as they share the code shape (the batched kernel uses that).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_device.hpp"
#include "ars_kernels.hpp"

namespace ars {

template <int TL, int PQ>
struct Tile {
    static constexpr int N = 1 << TL;
    static constexpr int Q = N / 16;
    static constexpr int RB = TL - 4;
    static constexpr int GB = 16 / Q;
    static constexpr int THREADS = Q * PQ;
    static constexpr int WORDS = N * PQ;

    __device__ static __forceinline__ void put_a(uint32_t *tb, const uint32_t (&x)[16], int q, int p) {
#pragma unroll
        for (int j = 0; j < 16; ++j) tb[(q * 16 + j) * PQ + p] = x[j];
    }
    __device__ static __forceinline__ void get_a(const uint32_t *tb, uint32_t (&x)[16], int q, int p) {
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = tb[(q * 16 + j) * PQ + p];
    }
    __device__ static __forceinline__ void put_b(uint32_t *tb, const uint32_t (&y)[GB][Q], int q, int p) {
#pragma unroll
        for (int m = 0; m < GB; ++m)
#pragma unroll
            for (int h = 0; h < Q; ++h) tb[(h * 16 + q * GB + m) * PQ + p] = y[m][h];
    }
    __device__ static __forceinline__ void get_b(const uint32_t *tb, uint32_t (&y)[GB][Q], int q, int p) {
#pragma unroll
        for (int m = 0; m < GB; ++m)
#pragma unroll
            for (int h = 0; h < Q; ++h) y[m][h] = tb[(h * 16 + q * GB + m) * PQ + p];
    }
};

// UT: the twiddle indices are wave-uniform (tiles of 64+ quads, where every
// lane of a wave holds the same shard group q, and the small recover, where
// every lane runs the same transform): the tables' high words come by scalar
// loads from gtw, the global table tw was staged from (UniTw), and each lane
// reads 8 of a table's 20 bytes from LDS.  The LDS traffic of the tiles is
// then mostly the two exchanges (measured 8.7% faster for the 128-point
// encode).  The small encode bodies measured 4-7% slower with it and keep
// the LDS tables.
template <int RL, int SH, int L0, int L1, bool INV, bool BZ, bool UT>
__device__ __forceinline__ void uni_levels(uint32_t *x, uint32_t hb, const MulTab *tw, const MulTab *gtw) {
    if constexpr (UT)
        run_levels_t<RL, SH, L0, L1, INV, BZ, kBurstMuls, UniTw>(x, hb, UniTw{tw, gtw});
    else
        run_levels<RL, SH, L0, L1, INV, BZ>(x, hb, tw);
}

// ---------------------------------------------------------------------------
// Encode, reference algorithm (rs/src/codec.rs:241-265): for each message
// chunk c, acc ^= IFFT_T(chunk c, omega_{(c+1)T}); parity = FFT_T(acc, 0).
// Input rows: message shards (chunk c row j = c*T + j); output: parity.
template <int TL, bool ACC, class IO>
__device__ __forceinline__ void body_encode_small(IO &io, uint32_t chunks, const MulTab *tw) {
    constexpr int T = 1 << TL;
    uint32_t acc[T], x[T];
#pragma unroll
    for (int j = 0; j < T; ++j) acc[j] = ACC ? io.ld_out(0u, j) : 0u;
    for (uint32_t c = 0; c < chunks; ++c) {
#pragma unroll
        for (int j = 0; j < T; ++j) x[j] = io.ld_in(c * T, j);
        run_levels<TL, 0, 0, TL, true, false>(x, (c + 1) * T, tw);
#pragma unroll
        for (int j = 0; j < T; ++j) acc[j] ^= x[j];
    }
    run_levels<TL, 0, 0, TL, false, true>(acc, 0u, tw);
#pragma unroll
    for (int j = 0; j < T; ++j) io.st_out(0u, j, acc[j]);
}

// NBUF = 2: ping-pong LDS buffers (one barrier per exchange); NBUF = 1: one
// buffer and an extra barrier before each reuse.
// x0 (optional): chunk 0's 16 input rows, already loaded by the caller (the
// one-codeword kernels issue those loads before staging their tables, so the
// two latencies overlap).
// Without UT every table comes from the LDS copy.  (Whole tables by scalar
// loads, copied to VGPRs for the bursts, measured 1.5% slower than LDS for
// phase B; UT copies only the two low words, and takes them from LDS.)
template <int TL, int PQ, bool ACC, int NBUF, class IO, bool UT = false>
__device__ __forceinline__ void body_encode_tile(IO &io, int q, int p, uint32_t *tile, uint32_t chunks,
                                                 const MulTab *tw, const uint32_t *x0 = nullptr,
                                                 const MulTab *gtw = nullptr) {
    static_assert(!UT || PQ % 64 == 0, "scalar table words need one shard group per wave");
    using TT = Tile<TL, PQ>;
    constexpr int T = TT::N, Q = TT::Q, GB = TT::GB, RB = TT::RB;
    uint32_t acc[GB][Q], y[GB][Q], x[16];
#pragma unroll
    for (int m = 0; m < GB; ++m)
#pragma unroll
        for (int h = 0; h < Q; ++h) acc[m][h] = ACC ? io.ld_out((uint32_t)(q * GB + m), h * 16) : 0u;
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t B = (c + 1) * T;
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = (x0 && c == 0) ? x0[j] : io.ld_in(c * T + q * 16, j);
        uni_levels<4, 0, 0, 4, true, false, UT>(x, B + q * 16, tw, gtw);
        uint32_t *tb = tile + (NBUF == 2 ? (c & 1) * TT::WORDS : 0);
        if (NBUF == 1 && c > 0) __syncthreads();
        TT::put_a(tb, x, q, p);
        __syncthreads();
        TT::get_b(tb, y, q, p);
#pragma unroll
        for (int m = 0; m < GB; ++m) {
            uni_levels<RB, 4, 4, TL, true, false, UT>(y[m], B, tw, gtw);
#pragma unroll
            for (int h = 0; h < Q; ++h) acc[m][h] ^= y[m][h];
        }
    }
#pragma unroll
    for (int m = 0; m < GB; ++m) uni_levels<RB, 4, 4, TL, false, true, UT>(acc[m], 0u, tw, gtw);
    // No barrier before put_b, even with one buffer: put_b writes exactly the
    // words this lane's last get_b read (same index set), and one lane's LDS
    // read and later write of a word are ordered.
    uint32_t *tb = tile + (NBUF == 2 ? (chunks & 1) * TT::WORDS : 0);
    TT::put_b(tb, acc, q, p);
    __syncthreads();
    TT::get_a(tb, x, q, p);
    uni_levels<4, 0, 0, 4, false, false, UT>(x, q * 16, tw, gtw);
#pragma unroll
    for (int j = 0; j < 16; ++j) io.st_out((uint32_t)(q * 16), j, x[j]);
}

// ---------------------------------------------------------------------------
// Encode extension for k = 2^KL with T not a power of two: coefficients
// d = IFFT_k(message, omega_{n-k}); parity block b = FFT_k(d, omega_{b*k}).
template <int KL, class IO>
__device__ __forceinline__ void body_encode_kpow2_small(IO &io, uint32_t blocks, uint32_t msg_base,
                                                        const MulTab *tw) {
    constexpr int K = 1 << KL;
    uint32_t d[K], x[K];
#pragma unroll
    for (int j = 0; j < K; ++j) d[j] = io.ld_in(0u, j);
    run_levels<KL, 0, 0, KL, true, false>(d, msg_base, tw);
    for (uint32_t b = 0; b < blocks; ++b) {
#pragma unroll
        for (int j = 0; j < K; ++j) x[j] = d[j];
        run_levels<KL, 0, 0, KL, false, false>(x, b * K, tw);
#pragma unroll
        for (int j = 0; j < K; ++j) io.st_out(b * K, j, x[j]);
    }
}

// UT: scalar table words, as body_encode_tile.
template <int KL, int PQ, int NBUF, class IO, bool UT = false>
__device__ __forceinline__ void body_encode_kpow2_tile(IO &io, int q, int p, uint32_t *tile,
                                                       uint32_t blocks, uint32_t msg_base,
                                                       const MulTab *tw, const MulTab *gtw = nullptr) {
    static_assert(!UT || PQ % 64 == 0, "scalar table words need one shard group per wave");
    using TT = Tile<KL, PQ>;
    constexpr int K = TT::N, Q = TT::Q, GB = TT::GB, RB = TT::RB;
    uint32_t x[16], d[GB][Q], y[GB][Q];
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = io.ld_in((uint32_t)(q * 16), j);
    uni_levels<4, 0, 0, 4, true, false, UT>(x, msg_base + q * 16, tw, gtw);
    TT::put_a(tile, x, q, p);
    __syncthreads();
    TT::get_b(tile, d, q, p);
#pragma unroll
    for (int m = 0; m < GB; ++m) uni_levels<RB, 4, 4, KL, true, false, UT>(d[m], msg_base, tw, gtw);
    for (uint32_t b = 0; b < blocks; ++b) {
#pragma unroll
        for (int m = 0; m < GB; ++m) {
#pragma unroll
            for (int h = 0; h < Q; ++h) y[m][h] = d[m][h];
            uni_levels<RB, 4, 4, KL, false, false, UT>(y[m], b * K, tw, gtw);
        }
        uint32_t *tb = tile + (NBUF == 2 ? ((b + 1) & 1) * TT::WORDS : 0);
        if (NBUF == 1) __syncthreads();
        TT::put_b(tb, y, q, p);
        __syncthreads();
        TT::get_a(tb, x, q, p);
        uni_levels<4, 0, 0, 4, false, false, UT>(x, b * K + q * 16, tw, gtw);
#pragma unroll
        for (int j = 0; j < 16; ++j) io.st_out(b * K + q * 16, j, x[j]);
    }
}

// ---------------------------------------------------------------------------
// Recover by LCH formal-derivative erasure decoding (all n points):
//   u_i = r_i * Pi(omega_i)            (0 at erased i)
//   D   = IFFT_n(u, 0)                 (coefficients of f*Pi, deg < n)
//   D'[i] = xor over b with bit b of i clear of D[i | 2^b]
//   f(omega_p) = FFT_n(D', 0)[p] / Pi'(omega_p)  for erased p
// and, as the reference does (it writes the error value, rs/src/codec.rs:560-566),
// the erased shard's previous content g_p is XORed in: out_p = c_p ^ g_p.
// Input and output rows are the same codeword.
// UT: the twiddles' high words by scalar loads (every lane runs the same
// transform; the row tables stay in LDS, they may differ per lane in a batch)
template <int NL, class IO, class TABS, bool UT = false>
__device__ __forceinline__ void body_recover_small(IO &io, const TABS &tabs, const MulTab *tw,
                                                   const MulTab *gtw = nullptr) {
    constexpr int N = 1 << NL;
    uint32_t x[N], orig[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        orig[i] = io.ld_in(0u, i);
        const uint32_t u = gmul4(orig[i], tabs.tab(i));
        x[i] = tabs.erased(i) ? 0u : u;
    }
    uni_levels<NL, 0, 0, NL, true, true, UT>(x, 0u, tw, gtw);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        uint32_t a = 0;
#pragma unroll
        for (int b = 0; b < NL; ++b)
            if (!((i >> b) & 1)) a ^= x[i | (1 << b)];
        x[i] = a;
    }
    uni_levels<NL, 0, 0, NL, false, true, UT>(x, 0u, tw, gtw);
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (tabs.erased(i)) io.st_out(0u, i, gmul4(x[i], tabs.tab(i)) ^ orig[i]);
}

// Low-bit part of the novel-basis formal derivative on 16 consecutive
// points: d[j] = xor over bits b < 4 clear in j of x[j | 2^b].
__device__ __forceinline__ void low_derivative(const uint32_t (&x)[16], uint32_t (&d)[16]) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        uint32_t a = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (!((j >> b) & 1)) a ^= x[j | (1 << b)];
        d[j] = a;
    }
}

// NBUF = 2 (two exchange buffers): the derivative's low-bit terms are taken
// from the phase-A output (see below), which saves the LDS copy of D and its
// cross-lane reads: two exchange barriers instead of four.  NBUF = 1 keeps the
// direct form: D is copied to LDS and the low-bit terms are read across lanes.
template <int NL, int PQ, int NBUF, class IO, class TABS>
__device__ __forceinline__ void body_recover_tile(IO &io, const TABS &tabs, int q, int p, uint32_t *tile,
                                                  const MulTab *tw) {
    using TT = Tile<NL, PQ>;
    constexpr int Q = TT::Q, GB = TT::GB, RB = TT::RB;
    uint32_t x[16], orig[16], y[GB][Q];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        orig[j] = io.ld_in((uint32_t)(q * 16), j);
        const int s = q * 16 + j;
        const uint32_t u = gmul4(orig[j], tabs.tab(s));
        x[j] = tabs.erased(s) ? 0u : u;
    }
    run_levels<4, 0, 0, 4, true, false>(x, q * 16, tw);
    uint32_t *t0 = tile;
    uint32_t *t1 = tile + (NBUF - 1) * TT::WORDS;
    TT::put_a(t0, x, q, p);
    __syncthreads();
    TT::get_b(t0, y, q, p);
#pragma unroll
    for (int m = 0; m < GB; ++m) run_levels<RB, 4, 4, NL, true, true>(y[m], 0u, tw);
    if constexpr (NBUF == 1) {
        // put_b rewrites exactly the words this lane's get_b read: no barrier
        // needed before it
        TT::put_b(t1, y, q, p);
        __syncthreads();
        // formal derivative: bits 4.. from registers, bits 0..3 from LDS
#pragma unroll
        for (int m = 0; m < GB; ++m) {
            const int c = q * GB + m;
#pragma unroll
            for (int h = 0; h < Q; ++h) {
                uint32_t a = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const uint32_t v = t1[(h * 16 + (c | (1 << b))) * PQ + p];
                    a ^= ((c >> b) & 1) ? 0u : v;
                }
#pragma unroll
                for (int b = 0; b < RB; ++b)
                    if (!((h >> b) & 1)) a ^= y[m][h | (1 << b)];
                y[m][h] = a;
            }
            run_levels<RB, 4, 4, NL, false, true>(y[m], 0u, tw);
        }
        __syncthreads();  // the derivative read other lanes' words of t0
        TT::put_b(t0, y, q, p);
        __syncthreads();
        TT::get_a(t0, x, q, p);
    } else {
        // The formal derivative splits by index bits: D' = Dhi(D) + Dlo(D),
        // Dlo the terms of bits 0..3 and Dhi those of bits 4...  Dlo acts on
        // the low 4 index bits alike for every value of the high ones, and the
        // phase-B levels (4..NL-1, beta = 0) act on the high bits with twiddles
        // that do not depend on the low ones, so Dlo commutes with both phase-B
        // transforms and, x being the phase-A output,
        //   FFT_B(D') = FFT_B(Dhi(D)) + Dlo(FFT_B(IFFT_B(x))) = FFT_B(Dhi(D)) + Dlo(x).
        // In the phase-A layout a lane holds the 16 shards q*16 + j, so Dlo(x)
        // needs only the lane's own words.
#pragma unroll
        for (int m = 0; m < GB; ++m) {
            // Dhi in place: ascending h only reads higher, still untouched, indices
#pragma unroll
            for (int h = 0; h < Q; ++h) {
                uint32_t a = 0;
#pragma unroll
                for (int b = 0; b < RB; ++b)
                    if (!((h >> b) & 1)) a ^= y[m][h | (1 << b)];
                y[m][h] = a;
            }
            run_levels<RB, 4, 4, NL, false, true>(y[m], 0u, tw);
        }
        TT::put_b(t1, y, q, p);
        __syncthreads();
        // t0 still holds x (after put_a it was only read), so the lane's own
        // words give Dlo(x) without keeping 16 more registers across phase B
        uint32_t dlo[16];
        TT::get_a(t0, x, q, p);
        low_derivative(x, dlo);
        TT::get_a(t1, x, q, p);
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] ^= dlo[j];
    }
    run_levels<4, 0, 0, 4, false, false>(x, q * 16, tw);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int s = q * 16 + j;
        if (tabs.erased(s)) io.st_out((uint32_t)(q * 16), j, gmul4(x[j], tabs.tab(s)) ^ orig[j]);
    }
}

// ---------------------------------------------------------------------------
// Recover, one exchange buffer (the default for dword-addressable codewords).
// The same algorithm as body_recover_tile, arranged so that every LDS word is
// written and read once per exchange and nothing is held for the output:
//   * the 16 erased flags of the lane's shard group are read in one LDS read
//     (Flags16), and erased rows are neither loaded nor multiplied (u_i = 0);
//   * Dlo(x), the derivative's low-bit terms of the phase-A output x (see
//     body_recover_tile), is computed in registers before the exchange and
//     added after it, so phase B's result goes back through the same buffer;
//   * the recovered value goes into the erased shard by XOR (io.xor_out), the
//     reference's c_p ^ g_p (rs/src/codec.rs:560-566), so the erased content
//     g_p is never read into registers.
// rows holds the lane's 16 input rows on entry (erased rows 0, see
// recover_rows_in) and is read only at the start.  after_exchange() runs once
// the tile is back in the phase-A layout, before the phase-A FFT; the
// one-codeword kernels load the next tile's rows into `rows` there.  The caller synchronises before reusing
// `tile` (other lanes' get_a reads of it).
// Bit j: row q*16 + j is erased (one VGPR; the 16 flag bytes would take four).
// all / none: every row of every shard group this wave holds is erased / none
// is (wave-uniform, from a ballot).
struct Flags16 {
    uint32_t m;
    bool all, none;
    __device__ __forceinline__ bool at(int j) const { return (m >> j) & 1u; }
};

// UNI: every lane of the wave holds the same shard group (64+ quads per
// workgroup), so the mask is moved to an SGPR and the per-row tests below
// become scalar branches that skip erased rows' loads and multiplies, and
// surviving rows' outputs, for the whole wave.
template <bool UNI, class TABS>
__device__ __forceinline__ Flags16 group_flags(const TABS &tabs, int q) {
    const uint4 v = *reinterpret_cast<const uint4 *>(tabs.e + q * 16);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) m |= ((w[j >> 2] >> (8 * (j & 3))) & 0xffu) ? (1u << j) : 0u;
    if constexpr (UNI) m = __builtin_amdgcn_readfirstlane(m);
    const bool all = __builtin_amdgcn_ballot_w64(m != 0xffffu) == 0;
    const bool none = __builtin_amdgcn_ballot_w64(m != 0u) == 0;
    return Flags16{m, all, none};
}

// OOB: erased rows are read at io.oob, past the end of the buffer, where a
// buffer load returns 0 -- 16 loads without a branch between them.  The
// offset must stay out of range after the row's scalar offset is added
// (the launchers choose it when (n + 15) * L < 2^32); otherwise each load
// of a surviving row sits in a branch.  With wave-uniform flags
// (IO::kUniformGroups) the branch is scalar and skips the load outright.
template <bool OOB, class IO>
__device__ __forceinline__ void recover_rows_in(const IO &io, const Flags16 &f, int q, uint32_t (&x)[16]) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if constexpr (OOB && !IO::kUniformGroups) {
            x[j] = io.ld_in_or0((uint32_t)(q * 16), j, f.at(j));
        } else {
            uint32_t v = 0;
            if (!f.at(j)) v = io.ld_in((uint32_t)(q * 16), j);
            x[j] = v;
        }
    }
}

// GRP: multiplies per lookup burst (see level_pass).
template <int NL, int PQ, int GRP, class IO, class TABS, class AFTER>
__device__ __forceinline__ void body_recover_tile1(const IO &io, const TABS &tabs, const Flags16 &f, int q, int p,
                                                   uint32_t *tile, const MulTab *tw, const uint32_t (&rows)[16],
                                                   AFTER &&after_exchange) {
    using TT = Tile<NL, PQ>;
    constexpr int Q = TT::Q, GB = TT::GB, RB = TT::RB;
    uint32_t x[16], y[GB][Q], dlo[16];
    if (f.all) {
        // every input row of the wave is erased, i.e. 0: so is its phase-A
        // IFFT and that output's derivative terms
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = dlo[j] = 0u;
    } else {
        // Erased rows are 0 and stay 0.  With a wave-uniform mask (PQ >= 64)
        // their multiply is skipped; otherwise it is done (a divergent branch
        // here costs registers).
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if (PQ % 64 == 0 && f.at(j))
                x[j] = 0u;
            else
                x[j] = gmul4(rows[j], tabs.tab(q * 16 + j));
        }
        run_levels<4, 0, 0, 4, true, false, GRP>(x, q * 16, tw);
        low_derivative(x, dlo);
    }
    TT::put_a(tile, x, q, p);
    __syncthreads();
    TT::get_b(tile, y, q, p);
#pragma unroll
    for (int m = 0; m < GB; ++m) {
        run_levels<RB, 4, 4, NL, true, true, GRP>(y[m], 0u, tw);
        // Dhi in place: ascending h only reads higher, still untouched, indices
#pragma unroll
        for (int h = 0; h < Q; ++h) {
            uint32_t a = 0;
#pragma unroll
            for (int b = 0; b < RB; ++b)
                if (!((h >> b) & 1)) a ^= y[m][h | (1 << b)];
            y[m][h] = a;
        }
        run_levels<RB, 4, 4, NL, false, true, GRP>(y[m], 0u, tw);
    }
    // put_b rewrites exactly the words this lane's get_b read
    TT::put_b(tile, y, q, p);
    __syncthreads();
    TT::get_a(tile, x, q, p);
    after_exchange();
    // no output row in the wave: its phase-A FFT is not needed
    if (f.none) return;
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] ^= dlo[j];
    run_levels<4, 0, 0, 4, false, false, GRP>(x, q * 16, tw);
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (f.at(j)) io.xor_out((uint32_t)(q * 16), j, gmul4(x[j], tabs.tab(q * 16 + j)));
}

// ---------------------------------------------------------------------------
// Codec::recompute_data_from_parity_sharded (rs/src/codec.rs:321-340):
// message = FFT_T(IFFT_T(parity, 0), omega_T), n = 2T.  Input rows: parity
// shards 0..T-1; output rows: the same buffer, shards T..2T-1.
template <int TL, class IO>
__device__ __forceinline__ void body_recompute_small(IO &io, const MulTab *tw) {
    constexpr int T = 1 << TL;
    uint32_t x[T];
#pragma unroll
    for (int j = 0; j < T; ++j) x[j] = io.ld_in(0u, j);
    run_levels<TL, 0, 0, TL, true, true>(x, 0u, tw);
    run_levels<TL, 0, 0, TL, false, false>(x, (uint32_t)T, tw);
#pragma unroll
    for (int j = 0; j < T; ++j) io.st_out((uint32_t)T, j, x[j]);
}

// UT: scalar table words, as body_encode_tile.
template <int TL, int PQ, class IO, bool UT = false>
__device__ __forceinline__ void body_recompute_tile(IO &io, int q, int p, uint32_t *tile, const MulTab *tw,
                                                    const MulTab *gtw = nullptr) {
    static_assert(!UT || PQ % 64 == 0, "scalar table words need one shard group per wave");
    using TT = Tile<TL, PQ>;
    constexpr int T = TT::N, Q = TT::Q, GB = TT::GB, RB = TT::RB;
    uint32_t x[16], y[GB][Q];
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = io.ld_in((uint32_t)(q * 16), j);
    uni_levels<4, 0, 0, 4, true, false, UT>(x, q * 16, tw, gtw);
    TT::put_a(tile, x, q, p);
    __syncthreads();
    TT::get_b(tile, y, q, p);
#pragma unroll
    for (int m = 0; m < GB; ++m) {
        uni_levels<RB, 4, 4, TL, true, true, UT>(y[m], 0u, tw, gtw);
        uni_levels<RB, 4, 4, TL, false, false, UT>(y[m], (uint32_t)T, tw, gtw);
    }
    uint32_t *t1 = tile + TT::WORDS;
    TT::put_b(t1, y, q, p);
    __syncthreads();
    TT::get_a(t1, x, q, p);
    uni_levels<4, 0, 0, 4, false, false, UT>(x, T + q * 16, tw, gtw);
#pragma unroll
    for (int j = 0; j < 16; ++j) io.st_out((uint32_t)(T + q * 16), j, x[j]);
}

// ---------------------------------------------------------------------------
// log_2 a in GF(2^8) / 0x11d (a != 0), at compile time.
constexpr uint32_t gf_clog(uint32_t a) {
    uint32_t x = 1, i = 0;
    while (x != a) {
        x <<= 1;
        if (x & 0x100u) x ^= 0x11du;
        ++i;
    }
    return i;
}

// log(v << 3k) for the entries v = 1..7 of the three parts k of a MulTab
// (part 2 has entries 1..3 only).
struct PartLogs {
    uint32_t l[3][8];
};
constexpr PartLogs make_part_logs() {
    PartLogs c{};
    for (uint32_t v = 1; v < 8; ++v) {
        c.l[0][v] = gf_clog(v);
        c.l[1][v] = gf_clog(v << 3);
        c.l[2][v] = v < 4 ? gf_clog(v << 6) : 0u;
    }
    return c;
}
constexpr PartLogs kPartLog = make_part_logs();

// ---------------------------------------------------------------------------
// Recover: per-job multiplier tables.  For shard i, Pi(omega_i) =
// prod_{p in E} (omega_i + omega_p); erased shards get 1 / Pi'(omega_i) with
// Pi'(omega_i) = prod_{p in E, p != i} (omega_i + omega_p) -- the
// reference's `denoms` (rs/src/codec.rs:457-462).

// LDS scratch of the plan computation: exp[512], log[256], the points
// omega_i (sub[256]), erased flags er[256] and the e erased points, 4 per
// word (pts4[64]); 1.5 KiB in all.
struct PlanScratch {
    uint8_t *exp, *log, *sub, *er;
    uint32_t *pts4;
};

// Where the plan goes: a RecPrep in global memory (the plan kernel) ...
struct PlanToGlobal {
    RecPrep *p;
    __device__ __forceinline__ void set(uint32_t i, const MulTab &m, uint8_t er) const {
        p->tab[i] = m;
        p->erased[i] = er;
    }
};
// ... or a recover workgroup's own LDS tables.
struct PlanToLds {
    MulTab *tab;
    uint8_t *er;
    __device__ __forceinline__ void set(uint32_t i, const MulTab &m, uint8_t e) const {
        tab[i] = m;
        er[i] = e;
    }
};

// The plan of one erasure pattern (n points, e erased at pos[0..e)), by a
// workgroup of THREADS threads; ends with the outputs written by each thread
// (the caller synchronises before reading them).
template <int THREADS, class OUT>
__device__ __forceinline__ void recover_plan(const OUT &out, uint32_t n, uint32_t e, const uint8_t *pos,
                                             const uint8_t *__restrict__ gtab, const PlanScratch &sc) {
    uint8_t *s_pts = reinterpret_cast<uint8_t *>(sc.pts4);
    for (uint32_t t = threadIdx.x; t < 256u; t += THREADS) {
        sc.exp[t] = gtab[t];
        sc.exp[t + 256] = gtab[t + 256];
        sc.log[t] = gtab[512 + t];
        sc.sub[t] = gtab[768 + t];
        sc.er[t] = 0;
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < e; t += THREADS) {
        sc.er[pos[t]] = 1;
        s_pts[t] = sc.sub[pos[t]];
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < n; t += THREADS) {
        const uint8_t x = sc.sub[t];
        // Sum of log(x + omega_p) over the erased p.  The terms depend on each
        // other only through the add, so the loop is branch-free and unrolled:
        // a thread keeps several LDS lookups in flight.  A zero difference (p = t)
        // adds log[0] = 0 (the uploaded table's entry 0).  Points are read 4 per
        // word (the word read is the same for every lane).
        const uint32_t x4 = (uint32_t)x * 0x01010101u;
        uint32_t lsum = 0;
#pragma unroll 4
        for (uint32_t i = 0; i < (e >> 2); ++i) {
            const uint32_t d = x4 ^ sc.pts4[i];
            lsum += (uint32_t)sc.log[d & 0xffu] + sc.log[(d >> 8) & 0xffu] + sc.log[(d >> 16) & 0xffu] +
                    sc.log[d >> 24];
        }
        for (uint32_t i = e & ~3u; i < e; ++i) lsum += sc.log[x ^ s_pts[i]];
        lsum %= 255u;
        const uint32_t er = sc.er[t];
        const uint32_t lv = er ? (255u - lsum) % 255u : lsum;
        // table entry v of part k is c * (v << 3k) = exp[log c + log(v << 3k)];
        // the second log is a compile-time constant
        uint8_t tb[3][8] = {};
#pragma unroll
        for (int v = 1; v < 8; ++v) {
            tb[0][v] = sc.exp[lv + kPartLog.l[0][v]];
            tb[1][v] = sc.exp[lv + kPartLog.l[1][v]];
            if (v < 4) tb[2][v] = sc.exp[lv + kPartLog.l[2][v]];
        }
        MulTab m;
        for (int k = 0; k < 3; ++k) {
            m.w[2 * k] = tb[k][0] | (tb[k][1] << 8) | (tb[k][2] << 16) | ((uint32_t)tb[k][3] << 24);
            m.w[2 * k + 1] = tb[k][4] | (tb[k][5] << 8) | (tb[k][6] << 16) | ((uint32_t)tb[k][7] << 24);
        }
        out.set(t, m, (uint8_t)er);
    }
}

// One workgroup of 256 threads per job, writing the plan to global memory.
__device__ __forceinline__ void recover_prep_body(RecPrep *__restrict__ out, uint32_t n, uint32_t e,
                                                  const uint8_t *pos, const uint8_t *__restrict__ gtab) {
    __shared__ uint8_t s_exp[512], s_log[256], s_sub[256], s_er[256];
    __shared__ uint32_t s_pts4[64];
    recover_plan<256>(PlanToGlobal{out}, n, e, pos, gtab, PlanScratch{s_exp, s_log, s_sub, s_er, s_pts4});
}

// ---------------------------------------------------------------------------
// I/O policies and multiplier sources.

// One codeword per launch: buffer resources, scalar per-shard offsets.
// Access modes (AM):
//   0  bytes: any base and shard_len; a quad is read and written byte by byte;
//   1  dwords: shard_len % 4 == 0 and 4-byte aligned buffers;
//   2  shifted dwords: shard_len % 4 == 0 and input and output both r = 1..3
//      bytes past a dword boundary, so every row is too.  The rows are
//      addressed from the dword boundary below the base, and a lane's quad is
//      the dword that holds columns 4*quad - r .. 4*quad - r + 3 (columns are
//      independent, so where a quad starts does not matter).  The first and
//      last quad of a row hold bytes of the neighbouring rows or past the
//      buffer; the transform computes them but the lane writes only the bytes
//      of its own row (`bm`).
//   3  wide dwords: mode 1's conditions for codewords of 4 GiB and more (the
//      specialisation below).
// A bool AL converts to 0 or 1.
template <int AM, bool UNI = false>
struct BufIO {
    static_assert(AM >= 0 && AM <= 2, "access mode");
    static constexpr bool AL = AM != 0;           // dword loads
    static constexpr bool kUniformGroups = UNI;  // every lane of a wave holds one shard group
    Rows in, out;
    uint32_t L, off, nb, oob;  // oob: the input buffer's size (see ld_in_or0)
    uint32_t bm;               // AM 2: the quad's bytes that belong to the row (bit per byte)
    __device__ __forceinline__ uint32_t ld_in(uint32_t g, int j) const {
        return ldq<AL>(in, off + g * L, (uint32_t)j * L, nb);
    }
    __device__ __forceinline__ uint32_t ld_out(uint32_t g, int j) const {
        return ldq<AL>(out, off + g * L, (uint32_t)j * L, nb);
    }
    __device__ __forceinline__ void st_out(uint32_t g, int j, uint32_t v) const {
        if constexpr (AM == 2) {
            const uint32_t vo = off + g * L, so = (uint32_t)j * L;
            if (bm == 0xfu) {
                __builtin_amdgcn_raw_buffer_store_b32(v, out.r, (int)vo, (int)so, 0);
            } else if (bm) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if ((bm >> i) & 1u)
                        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)(v >> (8 * i)), out.r, (int)(vo + i), (int)so, 0);
            }
        } else {
            stq<AL>(out, off + g * L, (uint32_t)j * L, v, nb);
        }
    }
    // input row g + j, or 0 if `zero` (read at `oob`, past the buffer's end)
    __device__ __forceinline__ uint32_t ld_in_or0(uint32_t g, int j, bool zero) const {
        static_assert(AL, "ld_in_or0 needs dword-addressable rows");
        return __builtin_amdgcn_raw_buffer_load_b32(in.r, (int)(zero ? oob : off + g * L), (int)((uint32_t)j * L), 0);
    }
    // out ^= v (dword rows only; AM 2: the bytes outside the row XOR in 0)
    __device__ __forceinline__ void xor_out(uint32_t g, int j, uint32_t v) const {
        static_assert(AL, "xor_out needs dword-addressable rows");
        if constexpr (AM == 2) {
            const uint32_t keep = (bm & 1u ? 0xffu : 0u) | (bm & 2u ? 0xff00u : 0u) | (bm & 4u ? 0xff0000u : 0u) |
                                  (bm & 8u ? 0xff000000u : 0u);
            if (bm) ars_buffer_atomic_xor((int)(v & keep), out.r, (int)(off + g * L), (int)((uint32_t)j * L), 0);
        } else {
            if (nb) ars_buffer_atomic_xor((int)v, out.r, (int)(off + g * L), (int)((uint32_t)j * L), 0);
        }
    }
};

// Access mode 3: rows whose byte offsets pass 2^32 (n * shard_len >= 4 GiB).
// Every access goes through a resource based at its own row, (g + j) * L
// bytes past the buffer (a 64-bit address), sized to the row, so the lane's
// column is the only offset and any shard_len below 4 GiB works.  The kernels
// use this mode only where g is wave-uniform (one shard group per wave, chunk
// and block indices), so the row address and the resource are formed on the
// scalar unit.  Every row the bodies address exists in these launches; a lane
// past the row's end reads column 0 and stores nothing (as in mode 1).
template <bool UNI>
struct BufIO<3, UNI> {
    static constexpr bool AL = true;
    static constexpr bool kUniformGroups = UNI;
    const uint8_t *in;
    uint8_t *out;
    uint32_t L, off, nb;
    __device__ __forceinline__ Rows row(const void *base, uint32_t g, int j) const {
        return make_rows(static_cast<const uint8_t *>(base) + (uint64_t)(g + (uint32_t)j) * L, L);
    }
    __device__ __forceinline__ uint32_t ld_in(uint32_t g, int j) const {
        return __builtin_amdgcn_raw_buffer_load_b32(row(in, g, j).r, (int)off, 0, 0);
    }
    __device__ __forceinline__ uint32_t ld_out(uint32_t g, int j) const {
        return __builtin_amdgcn_raw_buffer_load_b32(row(out, g, j).r, (int)off, 0, 0);
    }
    __device__ __forceinline__ void st_out(uint32_t g, int j, uint32_t v) const {
        if (nb) __builtin_amdgcn_raw_buffer_store_b32(v, row(out, g, j).r, (int)off, 0, 0);
    }
    // input row g + j, or 0 if `zero` (read at offset L, the row resource's end)
    __device__ __forceinline__ uint32_t ld_in_or0(uint32_t g, int j, bool zero) const {
        return __builtin_amdgcn_raw_buffer_load_b32(row(in, g, j).r, (int)(zero ? L : off), 0, 0);
    }
    __device__ __forceinline__ void xor_out(uint32_t g, int j, uint32_t v) const {
        if (nb) ars_buffer_atomic_xor((int)v, row(out, g, j).r, (int)off, 0, 0);
    }
};

// Byte offset of p past the dword boundary below it (access mode 2's r).
__device__ __forceinline__ uint32_t dword_shift(const void *p) { return (uint32_t)(reinterpret_cast<uintptr_t>(p) & 3u); }

template <int AM, bool UNI = false>
__device__ __forceinline__ BufIO<AM, UNI> make_bufio(const void *in, uint32_t in_bytes, void *out,
                                                     uint32_t out_bytes, uint32_t quad, uint32_t L) {
    BufIO<AM, UNI> io;
    io.L = L;
    const uint32_t col = quad * 4u;
    if constexpr (AM == 3) {
        // the byte counts are not used: every row is sized L (see BufIO<3>)
        (void)in_bytes;
        (void)out_bytes;
        io.in = static_cast<const uint8_t *>(in);
        io.out = static_cast<uint8_t *>(out);
        io.nb = col < L ? 4u : 0u;
        io.off = io.nb ? col : 0u;
    } else if constexpr (AM == 2) {
        // rows addressed from the dword boundary below each base; both bases
        // have the same r (the launchers check it)
        const uint32_t r = dword_shift(in);
        const uint32_t inb = (in_bytes + r + 3u) & ~3u, outb = (out_bytes + r + 3u) & ~3u;
        io.in = make_rows(static_cast<const uint8_t *>(in) - r, inb);
        io.out = make_rows(static_cast<uint8_t *>(out) - r, outb);
        io.oob = inb;
        // columns c0 .. c0 + 3 of the row, c0 = col - r; keep those in [0, L)
        const int c0 = (int)col - (int)r;
        const int lo = c0 < 0 ? -c0 : 0;
        const int hi = (int)L - c0 < 4 ? (int)L - c0 : 4;
        io.bm = hi > lo ? ((1u << hi) - 1u) & ~((1u << lo) - 1u) : 0u;
        io.nb = io.bm ? 4u : 0u;
        io.off = io.bm ? col : 0u;
    } else {
        io.in = make_rows(in, in_bytes);
        io.out = make_rows(out, out_bytes);
        io.oob = in_bytes;
        io.bm = 0;
        io.nb = col < L ? ((L - col) < 4u ? (L - col) : 4u) : 0u;
        io.off = (AM == 1 && io.nb == 0u) ? 0u : col;
    }
    return io;
}

// Per-shard multipliers staged in LDS (one codeword per workgroup).
struct LdsTabs {
    const MulTab *t;
    const uint8_t *e;
    __device__ __forceinline__ MulTab tab(int s) const { return lds_tab(t, (uint32_t)s); }
    __device__ __forceinline__ bool erased(int s) const { return e[s] != 0; }
};

}  // namespace ars
