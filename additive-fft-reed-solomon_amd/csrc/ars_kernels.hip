// ars_kernels.hip -- fused gfx950 kernels for the sharded encode / erasure
// recover path (one HBM read of the inputs, one HBM write of the outputs).
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
// Twiddle multipliers (255 of them, 6 KiB) are staged into LDS once per
// workgroup.  Point indices are absolute (they include the transform's shift
// beta = omega_B): the twiddle of the block whose first point is o is
// s_l(omega_o) (see level_pass).
//
// Global memory is addressed through buffer resources: a per-lane byte
// offset (the lane's column, plus its shard group when lanes of one wave own
// different shard groups) and a scalar per-shard offset.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_device.hpp"
#include "ars_kernels.hpp"

namespace ars {

__device__ __forceinline__ void stage_words(uint32_t *dst, const uint32_t *__restrict__ src, uint32_t words) {
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ void stage_twiddles(MulTab *s_tw, const MulTab *__restrict__ g_tw) {
    stage_words(reinterpret_cast<uint32_t *>(s_tw), reinterpret_cast<const uint32_t *>(g_tw),
                (uint32_t)kTwiddleEntries * 6u);
}

// Column quad of this lane: byte column `col`, valid bytes `nb`, and the
// per-lane offset to use for loads (clamped in-bounds on the dword path).
struct Quad {
    uint32_t col, nb, off;
};

template <bool AL>
__device__ __forceinline__ Quad lane_quad(uint32_t quad, uint32_t L) {
    Quad r;
    r.col = quad * 4u;
    r.nb = r.col < L ? ((L - r.col) < 4u ? (L - r.col) : 4u) : 0u;
    r.off = (AL && r.nb == 0u) ? 0u : r.col;
    return r;
}

// ---------------------------------------------------------------------------
// Encode, reference algorithm (rs/src/codec.rs:241-265), T = 2^TL <= 16.
template <int TL, bool AL, bool ACC>
__global__ void __launch_bounds__(256) k_encode_small(const uint8_t *__restrict__ msg,
                                                      uint8_t *__restrict__ par, uint32_t L,
                                                      uint32_t chunks,
                                                      const MulTab *__restrict__ g_tw) {
    constexpr int T = 1 << TL;
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles(tw, g_tw);
    __syncthreads();
    const Quad qd = lane_quad<AL>(blockIdx.x * blockDim.x + threadIdx.x, L);
    const Rows rm = make_rows(msg, chunks * T * L), rp = make_rows(par, T * L);
    uint32_t acc[T], x[T];
#pragma unroll
    for (int j = 0; j < T; ++j) acc[j] = ACC ? ldq<AL>(rp, qd.off, j * L, qd.nb) : 0u;
    for (uint32_t c = 0; c < chunks; ++c) {
#pragma unroll
        for (int j = 0; j < T; ++j) x[j] = ldq<AL>(rm, qd.off, (c * T + j) * L, qd.nb);
        run_levels<TL, 0, 0, TL, true, false>(x, (c + 1) * T, tw);
#pragma unroll
        for (int j = 0; j < T; ++j) acc[j] ^= x[j];
    }
    run_levels<TL, 0, 0, TL, false, true>(acc, 0u, tw);
#pragma unroll
    for (int j = 0; j < T; ++j) stq<AL>(rp, qd.off, j * L, acc[j], qd.nb);
}

// Two-phase helpers for TL >= 5 (see the file comment).
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

// Lane coordinates inside a tile: quad p, shard group q (wave-uniform when a
// wave holds one group).
template <int PQ>
__device__ __forceinline__ int tile_q() {
    const int q = threadIdx.x / PQ;
    if constexpr (PQ % 64 == 0) return __builtin_amdgcn_readfirstlane(q);
    return q;
}

// Encode, reference algorithm, T = 2^TL >= 32.
template <int TL, int PQ, bool AL, bool ACC>
__global__ void __launch_bounds__((Tile<TL, PQ>::THREADS))
    k_encode_tile(const uint8_t *__restrict__ msg, uint8_t *__restrict__ par, uint32_t L,
                  uint32_t chunks, const MulTab *__restrict__ g_tw) {
    using TT = Tile<TL, PQ>;
    constexpr int T = TT::N, Q = TT::Q, GB = TT::GB, RB = TT::RB;
    __shared__ uint32_t tile[2 * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles(tw, g_tw);
    __syncthreads();
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    const Quad qd = lane_quad<AL>(blockIdx.x * PQ + p, L);
    const Rows rm = make_rows(msg, chunks * T * L), rp = make_rows(par, T * L);
    const uint32_t va = qd.off + (uint32_t)(q * 16) * L;  // phase-A rows q*16 + j
    uint32_t acc[GB][Q], y[GB][Q], x[16];
#pragma unroll
    for (int m = 0; m < GB; ++m)
#pragma unroll
        for (int h = 0; h < Q; ++h)
            acc[m][h] = ACC ? ldq<AL>(rp, qd.off + (uint32_t)(q * GB + m) * L, h * 16 * L, qd.nb) : 0u;
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t B = (c + 1) * T;
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = ldq<AL>(rm, va, (c * T + j) * L, qd.nb);
        run_levels<4, 0, 0, 4, true, false>(x, B + q * 16, tw);
        uint32_t *tb = tile + (c & 1) * TT::WORDS;
        TT::put_a(tb, x, q, p);
        __syncthreads();
        TT::get_b(tb, y, q, p);
#pragma unroll
        for (int m = 0; m < GB; ++m) {
            run_levels<RB, 4, 4, TL, true, false>(y[m], B, tw);
#pragma unroll
            for (int h = 0; h < Q; ++h) acc[m][h] ^= y[m][h];
        }
    }
#pragma unroll
    for (int m = 0; m < GB; ++m) run_levels<RB, 4, 4, TL, false, true>(acc[m], 0u, tw);
    uint32_t *tb = tile + (chunks & 1) * TT::WORDS;
    TT::put_b(tb, acc, q, p);
    __syncthreads();
    TT::get_a(tb, x, q, p);
    run_levels<4, 0, 0, 4, false, false>(x, q * 16, tw);
#pragma unroll
    for (int j = 0; j < 16; ++j) stq<AL>(rp, va, j * L, x[j], qd.nb);
}

// ---------------------------------------------------------------------------
// Encode extension for k = 2^KL with T not a power of two: coefficients
// d = IFFT_k(message, omega_{n-k}); parity block b = FFT_k(d, omega_{b*k}).
template <int KL, bool AL>
__global__ void __launch_bounds__(256) k_encode_kpow2_small(const uint8_t *__restrict__ msg,
                                                            uint8_t *__restrict__ par, uint32_t L,
                                                            uint32_t blocks, uint32_t msg_base,
                                                            const MulTab *__restrict__ g_tw) {
    constexpr int K = 1 << KL;
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles(tw, g_tw);
    __syncthreads();
    const Quad qd = lane_quad<AL>(blockIdx.x * blockDim.x + threadIdx.x, L);
    const Rows rm = make_rows(msg, K * L), rp = make_rows(par, blocks * K * L);
    uint32_t d[K], x[K];
#pragma unroll
    for (int j = 0; j < K; ++j) d[j] = ldq<AL>(rm, qd.off, j * L, qd.nb);
    run_levels<KL, 0, 0, KL, true, false>(d, msg_base, tw);
    for (uint32_t b = 0; b < blocks; ++b) {
#pragma unroll
        for (int j = 0; j < K; ++j) x[j] = d[j];
        run_levels<KL, 0, 0, KL, false, false>(x, b * K, tw);
#pragma unroll
        for (int j = 0; j < K; ++j) stq<AL>(rp, qd.off, (b * K + j) * L, x[j], qd.nb);
    }
}

template <int KL, int PQ, bool AL>
__global__ void __launch_bounds__((Tile<KL, PQ>::THREADS))
    k_encode_kpow2_tile(const uint8_t *__restrict__ msg, uint8_t *__restrict__ par, uint32_t L,
                        uint32_t blocks, uint32_t msg_base, const MulTab *__restrict__ g_tw) {
    using TT = Tile<KL, PQ>;
    constexpr int K = TT::N, Q = TT::Q, GB = TT::GB, RB = TT::RB;
    __shared__ uint32_t tile[2 * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles(tw, g_tw);
    __syncthreads();
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    const Quad qd = lane_quad<AL>(blockIdx.x * PQ + p, L);
    const Rows rm = make_rows(msg, K * L), rp = make_rows(par, blocks * K * L);
    const uint32_t va = qd.off + (uint32_t)(q * 16) * L;
    uint32_t x[16], d[GB][Q], y[GB][Q];
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = ldq<AL>(rm, va, j * L, qd.nb);
    run_levels<4, 0, 0, 4, true, false>(x, msg_base + q * 16, tw);
    TT::put_a(tile, x, q, p);
    __syncthreads();
    TT::get_b(tile, d, q, p);
#pragma unroll
    for (int m = 0; m < GB; ++m) run_levels<RB, 4, 4, KL, true, false>(d[m], msg_base, tw);
    for (uint32_t b = 0; b < blocks; ++b) {
#pragma unroll
        for (int m = 0; m < GB; ++m) {
#pragma unroll
            for (int h = 0; h < Q; ++h) y[m][h] = d[m][h];
            run_levels<RB, 4, 4, KL, false, false>(y[m], b * K, tw);
        }
        uint32_t *tb = tile + ((b + 1) & 1) * TT::WORDS;
        TT::put_b(tb, y, q, p);
        __syncthreads();
        TT::get_a(tb, x, q, p);
        run_levels<4, 0, 0, 4, false, false>(x, b * K + q * 16, tw);
#pragma unroll
        for (int j = 0; j < 16; ++j) stq<AL>(rp, va, (b * K + j) * L, x[j], qd.nb);
    }
}

// ---------------------------------------------------------------------------
// Recover: per-job multiplier tables.  For shard i, Pi(omega_i) =
// prod_{p in E} (omega_i + omega_p); erased shards get 1 / Pi'(omega_i) with
// Pi'(omega_i) = prod_{p in E, p != i} (omega_i + omega_p) -- the
// reference's `denoms` (rs/src/codec.rs:457-462).
__global__ void __launch_bounds__(256) k_recover_prep(RecPrep *__restrict__ out, uint32_t n,
                                                      uint32_t e, PosArg pos,
                                                      const uint8_t *__restrict__ gtab) {
    __shared__ uint8_t s_exp[512], s_log[256], s_sub[256], s_er[256], s_pts[256];
    const uint32_t t = threadIdx.x;
    s_exp[t] = gtab[t];
    s_exp[t + 256] = gtab[t + 256];
    s_log[t] = gtab[512 + t];
    s_sub[t] = gtab[768 + t];
    s_er[t] = 0;
    __syncthreads();
    if (t < e) {
        s_er[pos.p[t]] = 1;
        s_pts[t] = s_sub[pos.p[t]];
    }
    __syncthreads();
    if (t < n) {
        const uint8_t x = s_sub[t];
        uint32_t lsum = 0;
        for (uint32_t i = 0; i < e; ++i) {
            const uint8_t dlt = x ^ s_pts[i];
            if (dlt) lsum += s_log[dlt];
        }
        lsum %= 255u;
        const uint32_t er = s_er[t];
        const uint32_t lv = er ? (255u - lsum) % 255u : lsum;
        uint8_t tb[3][8];
        for (int v = 0; v < 8; ++v) {
            const uint32_t m0 = (uint32_t)v, m1 = (uint32_t)v << 3, m2 = (uint32_t)v << 6;
            tb[0][v] = m0 ? s_exp[lv + s_log[m0]] : 0;
            tb[1][v] = m1 ? s_exp[lv + s_log[m1]] : 0;
            tb[2][v] = (v < 4 && m2) ? s_exp[lv + s_log[m2 & 0xff]] : 0;
        }
        MulTab m;
        for (int k = 0; k < 3; ++k) {
            m.w[2 * k] = tb[k][0] | (tb[k][1] << 8) | (tb[k][2] << 16) | ((uint32_t)tb[k][3] << 24);
            m.w[2 * k + 1] = tb[k][4] | (tb[k][5] << 8) | (tb[k][6] << 16) | ((uint32_t)tb[k][7] << 24);
        }
        out->tab[t] = m;
        out->erased[t] = (uint8_t)er;
    }
}

template <int N>
__device__ __forceinline__ void stage_prep(MulTab *s_tab, uint8_t *s_er, const RecPrep *__restrict__ prep) {
    stage_words(reinterpret_cast<uint32_t *>(s_tab), reinterpret_cast<const uint32_t *>(prep->tab),
                (uint32_t)N * 6u);
    for (uint32_t i = threadIdx.x; i < (uint32_t)N; i += blockDim.x) s_er[i] = prep->erased[i];
}

// LCH formal-derivative erasure decoding (all n points):
//   u_i = r_i * Pi(omega_i)            (0 at erased i)
//   D   = IFFT_n(u, 0)                 (coefficients of f*Pi, deg < n)
//   D'[i] = xor over b with bit b of i clear of D[i | 2^b]
//   f(omega_p) = FFT_n(D', 0)[p] / Pi'(omega_p)  for erased p
// and, as the reference does (it writes the error value), the erased shard's
// previous content g_p is XORed in: out_p = c_p ^ g_p.
template <int NL, bool AL>
__global__ void __launch_bounds__(256) k_recover_small(uint8_t *__restrict__ data, uint32_t L,
                                                       const RecPrep *__restrict__ prep,
                                                       const MulTab *__restrict__ g_tw) {
    constexpr int N = 1 << NL;
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ MulTab s_tab[N];
    __shared__ uint8_t s_er[N];
    stage_twiddles(tw, g_tw);
    stage_prep<N>(s_tab, s_er, prep);
    __syncthreads();
    const Quad qd = lane_quad<AL>(blockIdx.x * blockDim.x + threadIdx.x, L);
    const Rows rd = make_rows(data, N * L);
    uint32_t x[N], orig[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        orig[i] = ldq<AL>(rd, qd.off, i * L, qd.nb);
        x[i] = s_er[i] ? 0u : gmul4(orig[i], lds_tab(s_tab, i));
    }
    run_levels<NL, 0, 0, NL, true, true>(x, 0u, tw);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        uint32_t a = 0;
#pragma unroll
        for (int b = 0; b < NL; ++b)
            if (!((i >> b) & 1)) a ^= x[i | (1 << b)];
        x[i] = a;
    }
    run_levels<NL, 0, 0, NL, false, true>(x, 0u, tw);
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (s_er[i]) stq<AL>(rd, qd.off, i * L, gmul4(x[i], lds_tab(s_tab, i)) ^ orig[i], qd.nb);
}

// Two-phase recover.  NBUF = 1 reuses one LDS tile for the three exchanges
// (two extra barriers, half the LDS); NBUF = 2 ping-pongs.
template <int NL, int PQ, bool AL, int NBUF>
__global__ void __launch_bounds__((Tile<NL, PQ>::THREADS))
    k_recover_tile(uint8_t *__restrict__ data, uint32_t L, const RecPrep *__restrict__ prep,
                   const MulTab *__restrict__ g_tw) {
    using TT = Tile<NL, PQ>;
    constexpr int N = TT::N, Q = TT::Q, GB = TT::GB, RB = TT::RB;
    __shared__ uint32_t tile[NBUF * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ MulTab s_tab[N];
    __shared__ uint8_t s_er[N];
    stage_twiddles(tw, g_tw);
    stage_prep<N>(s_tab, s_er, prep);
    __syncthreads();
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    const Quad qd = lane_quad<AL>(blockIdx.x * PQ + p, L);
    const Rows rd = make_rows(data, N * L);
    const uint32_t va = qd.off + (uint32_t)(q * 16) * L;
    uint32_t x[16], orig[16], y[GB][Q];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        orig[j] = ldq<AL>(rd, va, j * L, qd.nb);
        const int s = q * 16 + j;
        const uint32_t u = gmul4(orig[j], lds_tab(s_tab, s));
        x[j] = s_er[s] ? 0u : u;
    }
    run_levels<4, 0, 0, 4, true, false>(x, q * 16, tw);
    uint32_t *t0 = tile;
    uint32_t *t1 = tile + (NBUF - 1) * TT::WORDS;
    TT::put_a(t0, x, q, p);
    __syncthreads();
    TT::get_b(t0, y, q, p);
#pragma unroll
    for (int m = 0; m < GB; ++m) run_levels<RB, 4, 4, NL, true, true>(y[m], 0u, tw);
    if constexpr (NBUF == 1) __syncthreads();
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
    if constexpr (NBUF == 1) __syncthreads();
    TT::put_b(t0, y, q, p);
    __syncthreads();
    TT::get_a(t0, x, q, p);
    run_levels<4, 0, 0, 4, false, false>(x, q * 16, tw);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int s = q * 16 + j;
        if (s_er[s]) stq<AL>(rd, va, j * L, gmul4(x[j], lds_tab(s_tab, s)) ^ orig[j], qd.nb);
    }
}

// ---------------------------------------------------------------------------
// Codec::recompute_data_from_parity_sharded (rs/src/codec.rs:321-340):
// message = FFT_T(IFFT_T(parity, 0), omega_T), n = 2T.
template <int TL, bool AL>
__global__ void __launch_bounds__(256) k_recompute_small(uint8_t *__restrict__ data, uint32_t L,
                                                         const MulTab *__restrict__ g_tw) {
    constexpr int T = 1 << TL;
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles(tw, g_tw);
    __syncthreads();
    const Quad qd = lane_quad<AL>(blockIdx.x * blockDim.x + threadIdx.x, L);
    const Rows rd = make_rows(data, 2 * T * L);
    uint32_t x[T];
#pragma unroll
    for (int j = 0; j < T; ++j) x[j] = ldq<AL>(rd, qd.off, j * L, qd.nb);
    run_levels<TL, 0, 0, TL, true, true>(x, 0u, tw);
    run_levels<TL, 0, 0, TL, false, false>(x, (uint32_t)T, tw);
#pragma unroll
    for (int j = 0; j < T; ++j) stq<AL>(rd, qd.off, (T + j) * L, x[j], qd.nb);
}

template <int TL, int PQ, bool AL>
__global__ void __launch_bounds__((Tile<TL, PQ>::THREADS))
    k_recompute_tile(uint8_t *__restrict__ data, uint32_t L, const MulTab *__restrict__ g_tw) {
    using TT = Tile<TL, PQ>;
    constexpr int T = TT::N, Q = TT::Q, GB = TT::GB, RB = TT::RB;
    __shared__ uint32_t tile[2 * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles(tw, g_tw);
    __syncthreads();
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    const Quad qd = lane_quad<AL>(blockIdx.x * PQ + p, L);
    const Rows rd = make_rows(data, 2 * T * L);
    const uint32_t va = qd.off + (uint32_t)(q * 16) * L;
    uint32_t x[16], y[GB][Q];
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = ldq<AL>(rd, va, j * L, qd.nb);
    run_levels<4, 0, 0, 4, true, false>(x, q * 16, tw);
    TT::put_a(tile, x, q, p);
    __syncthreads();
    TT::get_b(tile, y, q, p);
#pragma unroll
    for (int m = 0; m < GB; ++m) {
        run_levels<RB, 4, 4, TL, true, true>(y[m], 0u, tw);
        run_levels<RB, 4, 4, TL, false, false>(y[m], (uint32_t)T, tw);
    }
    uint32_t *t1 = tile + TT::WORDS;
    TT::put_b(t1, y, q, p);
    __syncthreads();
    TT::get_a(t1, x, q, p);
    run_levels<4, 0, 0, 4, false, false>(x, T + q * 16, tw);
#pragma unroll
    for (int j = 0; j < 16; ++j) stq<AL>(rd, va, (T + j) * L, x[j], qd.nb);
}

// ---------------------------------------------------------------------------
// Launchers.
static inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

template <int TL>
constexpr int tile_pq() { return TL >= 8 ? 32 : 64; }

template <int TL, bool AL, bool ACC>
static hipError_t enc_one(const uint8_t *msg, uint8_t *par, uint32_t L, uint32_t chunks,
                          const MulTab *tw, hipStream_t s) {
    const uint64_t quads = ((uint64_t)L + 3) / 4;
    if constexpr (TL <= 4) {
        hipLaunchKernelGGL((k_encode_small<TL, AL, ACC>), dim3(ceil_div(quads, 256)), dim3(256), 0, s,
                           msg, par, L, chunks, tw);
    } else {
        constexpr int PQ = tile_pq<TL>();
        hipLaunchKernelGGL((k_encode_tile<TL, PQ, AL, ACC>), dim3(ceil_div(quads, PQ)),
                           dim3(Tile<TL, PQ>::THREADS), 0, s, msg, par, L, chunks, tw);
    }
    return hipGetLastError();
}

template <int TL>
static hipError_t enc_acc(bool al, bool acc, const uint8_t *msg, uint8_t *par, uint32_t L,
                          uint32_t chunks, const MulTab *tw, hipStream_t s) {
    if (al) return acc ? enc_one<TL, true, true>(msg, par, L, chunks, tw, s)
                       : enc_one<TL, true, false>(msg, par, L, chunks, tw, s);
    return acc ? enc_one<TL, false, true>(msg, par, L, chunks, tw, s)
               : enc_one<TL, false, false>(msg, par, L, chunks, tw, s);
}

hipError_t launch_encode_ref(int tl, bool al, bool acc, const uint8_t *msg, uint8_t *par,
                             uint32_t L, uint32_t chunks, const MulTab *tw, hipStream_t s) {
    switch (tl) {
        case 0: return enc_acc<0>(al, acc, msg, par, L, chunks, tw, s);
        case 1: return enc_acc<1>(al, acc, msg, par, L, chunks, tw, s);
        case 2: return enc_acc<2>(al, acc, msg, par, L, chunks, tw, s);
        case 3: return enc_acc<3>(al, acc, msg, par, L, chunks, tw, s);
        case 4: return enc_acc<4>(al, acc, msg, par, L, chunks, tw, s);
        case 5: return enc_acc<5>(al, acc, msg, par, L, chunks, tw, s);
        case 6: return enc_acc<6>(al, acc, msg, par, L, chunks, tw, s);
        case 7: return enc_acc<7>(al, acc, msg, par, L, chunks, tw, s);
        default: return hipErrorInvalidValue;
    }
}

template <int KL, bool AL>
static hipError_t kp_one(const uint8_t *msg, uint8_t *par, uint32_t L, uint32_t blocks,
                         uint32_t msg_base, const MulTab *tw, hipStream_t s) {
    const uint64_t quads = ((uint64_t)L + 3) / 4;
    if constexpr (KL <= 4) {
        hipLaunchKernelGGL((k_encode_kpow2_small<KL, AL>), dim3(ceil_div(quads, 256)), dim3(256), 0, s,
                           msg, par, L, blocks, msg_base, tw);
    } else {
        constexpr int PQ = tile_pq<KL>();
        hipLaunchKernelGGL((k_encode_kpow2_tile<KL, PQ, AL>), dim3(ceil_div(quads, PQ)),
                           dim3(Tile<KL, PQ>::THREADS), 0, s, msg, par, L, blocks, msg_base, tw);
    }
    return hipGetLastError();
}

hipError_t launch_encode_kpow2(int kl, bool al, const uint8_t *msg, uint8_t *par, uint32_t L,
                               uint32_t blocks, uint32_t msg_base, const MulTab *tw, hipStream_t s) {
#define KP_CASE(K) \
    case K: return al ? kp_one<K, true>(msg, par, L, blocks, msg_base, tw, s) \
                      : kp_one<K, false>(msg, par, L, blocks, msg_base, tw, s);
    switch (kl) {
        KP_CASE(0) KP_CASE(1) KP_CASE(2) KP_CASE(3) KP_CASE(4) KP_CASE(5) KP_CASE(6) KP_CASE(7)
        default: return hipErrorInvalidValue;
    }
#undef KP_CASE
}

hipError_t launch_recover_prep(RecPrep *out, uint32_t n, uint32_t e, const PosArg &pos,
                               const uint8_t *gtab, hipStream_t s) {
    hipLaunchKernelGGL(k_recover_prep, dim3(1), dim3(256), 0, s, out, n, e, pos, gtab);
    return hipGetLastError();
}

template <int NL, bool AL>
static hipError_t rec_one(uint8_t *data, uint32_t L, const RecPrep *prep, const MulTab *tw,
                          hipStream_t s) {
    const uint64_t quads = ((uint64_t)L + 3) / 4;
    if constexpr (NL <= 4) {
        hipLaunchKernelGGL((k_recover_small<NL, AL>), dim3(ceil_div(quads, 256)), dim3(256), 0, s,
                           data, L, prep, tw);
    } else {
        constexpr int PQ = tile_pq<NL>();
        constexpr int NBUF = NL >= 8 ? 1 : 2;
        hipLaunchKernelGGL((k_recover_tile<NL, PQ, AL, NBUF>), dim3(ceil_div(quads, PQ)),
                           dim3(Tile<NL, PQ>::THREADS), 0, s, data, L, prep, tw);
    }
    return hipGetLastError();
}

hipError_t launch_recover(int nl, bool al, uint8_t *data, uint32_t L, const RecPrep *prep,
                          const MulTab *tw, hipStream_t s) {
#define RC_CASE(K) \
    case K: return al ? rec_one<K, true>(data, L, prep, tw, s) : rec_one<K, false>(data, L, prep, tw, s);
    switch (nl) {
        RC_CASE(1) RC_CASE(2) RC_CASE(3) RC_CASE(4) RC_CASE(5) RC_CASE(6) RC_CASE(7) RC_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef RC_CASE
}

template <int TL, bool AL>
static hipError_t rcp_one(uint8_t *data, uint32_t L, const MulTab *tw, hipStream_t s) {
    const uint64_t quads = ((uint64_t)L + 3) / 4;
    if constexpr (TL <= 4) {
        hipLaunchKernelGGL((k_recompute_small<TL, AL>), dim3(ceil_div(quads, 256)), dim3(256), 0, s,
                           data, L, tw);
    } else {
        constexpr int PQ = tile_pq<TL>();
        hipLaunchKernelGGL((k_recompute_tile<TL, PQ, AL>), dim3(ceil_div(quads, PQ)),
                           dim3(Tile<TL, PQ>::THREADS), 0, s, data, L, tw);
    }
    return hipGetLastError();
}

hipError_t launch_recompute(int tl, bool al, uint8_t *data, uint32_t L, const MulTab *tw,
                            hipStream_t s) {
#define RP_CASE(K) \
    case K: return al ? rcp_one<K, true>(data, L, tw, s) : rcp_one<K, false>(data, L, tw, s);
    switch (tl) {
        RP_CASE(0) RP_CASE(1) RP_CASE(2) RP_CASE(3) RP_CASE(4) RP_CASE(5) RP_CASE(6) RP_CASE(7)
        default: return hipErrorInvalidValue;
    }
#undef RP_CASE
}

}  // namespace ars
