// ars_decode.hip -- device side of sharded error decoding.
//
// Every byte column of an n x L codeword is one received word, decoded as
// Codec::decode_systematic_scalar decodes it (rs/src/codec.rs:341-420).  The
// device does the FFT-heavy steps; the host solves the key equation of the
// few columns that need it (ars_keyeq.hpp):
//   k_syndrome_*   s = sum_c IFFT_T(chunk c, omega_{cT}) per column
//                  (rs/src/codec.rs:266-301), in the encode tile layout; a
//                  column whose syndrome is not zero is appended to a list
//                  with its T syndrome bytes.  Clean columns write nothing,
//                  so a scrub of clean data only reads the codeword.
//   k_decode_tail  one wave per listed column: the root search (one T-point
//                  FFT of lambda per chunk, rs/src/codec.rs:383-396, the first
//                  deg(lambda) roots in index order) and the error values
//                  v(w_p) / lambda'(w_p) (rs/src/codec.rs:405-431), applied in
//                  place when deg(lambda) roots were found.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_decode.hpp"
#include "ars_device.hpp"
#include "ars_tiles.hpp"

namespace ars {

namespace {

__device__ __forceinline__ void stage_tw(MulTab *s_tw, const MulTab *__restrict__ g_tw) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(g_tw);
    uint32_t *dst = reinterpret_cast<uint32_t *>(s_tw);
    for (uint32_t i = threadIdx.x; i < (uint32_t)kTwiddleEntries * 6u; i += blockDim.x) dst[i] = src[i];
}

// Byte mask of the quad's listed columns: valid (b < nb) and non-zero.
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t orw, uint32_t nb) {
    uint32_t m = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if ((uint32_t)b < nb && ((orw >> (8 * b)) & 0xffu)) m |= 1u << b;
    return m;
}

// Appends the columns of `mask` (quad starting at col0); returns their first
// list index.
__device__ __forceinline__ uint32_t list_columns(const SynOut &so, uint32_t mask, uint32_t col0) {
    const uint32_t base = atomicAdd(so.count, (uint32_t)__popc(mask));
    uint32_t k = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if (mask & (1u << b)) so.cols[base + k++] = col0 + (uint32_t)b;
    return base;
}

template <int TL, bool AL>
__global__ void __launch_bounds__(256) k_syndrome_small(const uint8_t *__restrict__ rx, uint32_t L,
                                                        uint32_t chunks, const MulTab *__restrict__ g_tw,
                                                        SynOut so) {
    constexpr int T = 1 << TL;
    __shared__ MulTab tw[kTwiddleEntries];
    stage_tw(tw, g_tw);
    __syncthreads();
    const uint32_t quad = blockIdx.x * blockDim.x + threadIdx.x;
    BufIO<AL> io = make_bufio<AL>(rx, chunks * T * L, const_cast<uint8_t *>(rx), 0u, quad, L);
    uint32_t acc[T], x[T];
#pragma unroll
    for (int j = 0; j < T; ++j) acc[j] = 0u;
    for (uint32_t c = 0; c < chunks; ++c) {
#pragma unroll
        for (int j = 0; j < T; ++j) x[j] = io.ld_in(c * T, j);
        run_levels<TL, 0, 0, TL, true, false>(x, c * T, tw);
#pragma unroll
        for (int j = 0; j < T; ++j) acc[j] ^= x[j];
    }
    uint32_t orw = 0;
#pragma unroll
    for (int j = 0; j < T; ++j) orw |= acc[j];
    const uint32_t mask = nonzero_bytes(orw, io.nb);
    if (mask) {
        uint32_t k = list_columns(so, mask, quad * 4u);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (!(mask & (1u << b))) continue;
            uint8_t *dst = so.syn + (size_t)k++ * T;
#pragma unroll
            for (int j = 0; j < T; ++j) dst[j] = (uint8_t)(acc[j] >> (8 * b));
        }
    }
}

// Tile variant for T >= 32 (see ars_tiles.hpp): each chunk's IFFT runs levels
// 0..3 in registers, one LDS exchange, levels 4..TL-1 in registers; the
// accumulated syndrome stays in the phase-B layout (lane q holds rows
// h*16 + q*GB + m).  The lanes of a quad OR their rows into LDS, lane q = 0
// lists the quad's non-zero columns, and every lane then writes its rows of
// those columns.
template <int TL, int PQ, int NBUF, bool AL>
__global__ void __launch_bounds__((Tile<TL, PQ>::THREADS))
    k_syndrome_tile(const uint8_t *__restrict__ rx, uint32_t L, uint32_t chunks,
                    const MulTab *__restrict__ g_tw, SynOut so) {
    using TT = Tile<TL, PQ>;
    constexpr int T = TT::N, Q = TT::Q, GB = TT::GB, RB = TT::RB;
    __shared__ uint32_t tile[NBUF * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ uint32_t s_or[PQ], s_base[PQ];
    stage_tw(tw, g_tw);
    if (threadIdx.x < PQ) s_or[threadIdx.x] = 0u;
    __syncthreads();
    const int p = threadIdx.x % PQ, q = threadIdx.x / PQ;
    const uint32_t quad = blockIdx.x * PQ + p;
    BufIO<AL> io = make_bufio<AL>(rx, chunks * T * L, const_cast<uint8_t *>(rx), 0u, quad, L);
    uint32_t acc[GB][Q], y[GB][Q], x[16];
#pragma unroll
    for (int m = 0; m < GB; ++m)
#pragma unroll
        for (int h = 0; h < Q; ++h) acc[m][h] = 0u;
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t B = c * T;
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = io.ld_in(B + q * 16, j);
        run_levels<4, 0, 0, 4, true, false>(x, B + q * 16, tw);
        uint32_t *tb = tile + (NBUF == 2 ? (c & 1) * TT::WORDS : 0);
        if (NBUF == 1 && c > 0) __syncthreads();
        TT::put_a(tb, x, q, p);
        __syncthreads();
        TT::get_b(tb, y, q, p);
        // chunk 0 sits at beta = 0: its phase-B blocks at offset 0 have the
        // zero twiddle, XOR-only at compile time
        if (c == 0) {
#pragma unroll
            for (int m = 0; m < GB; ++m) run_levels<RB, 4, 4, TL, true, true>(y[m], 0u, tw);
        } else {
#pragma unroll
            for (int m = 0; m < GB; ++m) run_levels<RB, 4, 4, TL, true, false>(y[m], B, tw);
        }
#pragma unroll
        for (int m = 0; m < GB; ++m)
#pragma unroll
            for (int h = 0; h < Q; ++h) acc[m][h] ^= y[m][h];
    }
    uint32_t orw = 0;
#pragma unroll
    for (int m = 0; m < GB; ++m)
#pragma unroll
        for (int h = 0; h < Q; ++h) orw |= acc[m][h];
    if (orw) atomicOr(&s_or[p], orw);
    __syncthreads();
    const uint32_t mask = nonzero_bytes(s_or[p], io.nb);
    if (q == 0 && mask) s_base[p] = list_columns(so, mask, quad * 4u);
    __syncthreads();
    if (mask) {
        uint32_t k = s_base[p];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (!(mask & (1u << b))) continue;
            uint8_t *dst = so.syn + (size_t)k++ * T;
#pragma unroll
            for (int m = 0; m < GB; ++m)
#pragma unroll
                for (int h = 0; h < Q; ++h) dst[h * 16 + q * GB + m] = (uint8_t)(acc[m][h] >> (8 * b));
        }
    }
}

__device__ __forceinline__ uint8_t lmul(const uint8_t *ex, const uint8_t *lg, uint8_t a, uint8_t b) {
    return (a && b) ? ex[lg[a] + lg[b]] : (uint8_t)0;
}

// One wave per listed column (blockIdx.x = list index).
__global__ void __launch_bounds__(64) k_decode_tail(uint8_t *__restrict__ rx, uint32_t L, uint32_t n, uint32_t tl,
                                                    const uint32_t *__restrict__ cols,
                                                    const DecPlan *__restrict__ plans,
                                                    uint8_t *__restrict__ status, uint32_t *__restrict__ failed,
                                                    const uint8_t *__restrict__ gtab,
                                                    const MulTab *__restrict__ g_tw) {
    __shared__ uint8_t s_exp[512], s_log[256], s_tw[256];
    __shared__ uint8_t ev[3][128];
    __shared__ uint8_t rpos[256], rval[256];
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < 512; i += 64) s_exp[i] = gtab[i];
    for (uint32_t i = lane; i < 256; i += 64) s_log[i] = gtab[512 + i];
    // byte 1 of a MulTab's first table is c * 1 = c
    for (uint32_t i = lane; i < (uint32_t)kTwiddleEntries; i += 64) s_tw[i] = (uint8_t)(g_tw[i].w[0] >> 8);
    const DecPlan *pl = plans + blockIdx.x;
    const uint32_t col = cols[blockIdx.x];
    const int dl = pl->deg;
    __syncthreads();
    if (dl < 0) {  // zero locator: true, nothing changes
        if (lane == 0 && status) status[col] = 1;
        return;
    }
    const uint32_t T = 1u << tl, need = (uint32_t)dl;
    uint32_t cnt = 0;  // roots found so far (uniform)
    for (uint32_t c = 0; c < n / T && cnt < need; ++c) {
        for (uint32_t i = lane; i < T; i += 64) {
            ev[0][i] = pl->lam[i];
            ev[1][i] = pl->v[i];
            ev[2][i] = pl->lp[i];
        }
        __syncthreads();
        // forward FFT of the three polynomials at beta = omega_{cT}
        // (rs/src/kernel/lut_kernel.rs:11-18, 28-54): levels tl-1 .. 0
        for (int l = (int)tl - 1; l >= 0; --l) {
            for (uint32_t i = lane; i < T / 2; i += 64) {
                const uint32_t blk = i >> l, pos = i & ((1u << l) - 1u);
                const uint32_t a = (blk << (l + 1)) + pos, b = a + (1u << l);
                const uint32_t o = c * T + (blk << (l + 1));
                const uint8_t t = s_tw[lvl_off(l) + (o >> (l + 1))];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const uint8_t xa = ev[k][a] ^ lmul(s_exp, s_log, t, ev[k][b]);
                    ev[k][a] = xa;
                    ev[k][b] ^= xa;
                }
            }
            __syncthreads();
        }
        for (uint32_t h = 0; h < T; h += 64) {
            const uint32_t off = h + lane;
            const bool root = off < T && ev[0][off] == 0;
            const uint64_t m = __ballot(root);
            const uint32_t rank = cnt + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (root && rank < need) {
                const uint8_t d = ev[2][off];
                rpos[rank] = (uint8_t)(c * T + off);
                rval[rank] = d ? lmul(s_exp, s_log, ev[1][off], s_exp[255 - s_log[d]]) : (uint8_t)0;
            }
            cnt += (uint32_t)__popcll(m);
        }
        __syncthreads();
    }
    if (cnt < need) {  // too few roots: false, the column is left as it was
        if (lane == 0) {
            if (status) status[col] = 2;
            atomicAdd(failed, 1u);
        }
        return;
    }
    for (uint32_t r = lane; r < need; r += 64) rx[(size_t)rpos[r] * L + col] ^= rval[r];
    if (lane == 0 && status) status[col] = 1;
}

static inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

template <int TL, bool AL>
hipError_t syn_one(const uint8_t *rx, uint32_t L, uint32_t chunks, const MulTab *tw, const SynOut &so,
                   hipStream_t s) {
    const uint64_t quads = ((uint64_t)L + 3) / 4;
    if constexpr (TL <= 4) {
        hipLaunchKernelGGL((k_syndrome_small<TL, AL>), dim3(ceil_div(quads, 256)), dim3(256), 0, s, rx, L, chunks,
                           tw, so);
    } else {
        // the encode tile's measured best shapes: 32 quads and one LDS buffer
        // at 128 points, 64 quads and two buffers below
        constexpr int PQ = TL >= 7 ? 32 : 64, NB = TL >= 7 ? 1 : 2;
        hipLaunchKernelGGL((k_syndrome_tile<TL, PQ, NB, AL>), dim3(ceil_div(quads, PQ)),
                           dim3(Tile<TL, PQ>::THREADS), 0, s, rx, L, chunks, tw, so);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_syndrome(int tl, bool al, const uint8_t *rx, uint32_t L, uint32_t chunks, const MulTab *tw,
                           const SynOut &so, hipStream_t s) {
#define SY_CASE(K) \
    case K: return al ? syn_one<K, true>(rx, L, chunks, tw, so, s) : syn_one<K, false>(rx, L, chunks, tw, so, s);
    switch (tl) {
        SY_CASE(1) SY_CASE(2) SY_CASE(3) SY_CASE(4) SY_CASE(5) SY_CASE(6) SY_CASE(7)
        default: return hipErrorInvalidValue;
    }
#undef SY_CASE
}

hipError_t launch_decode_tail(uint8_t *rx, uint32_t L, uint32_t n, int tl, const uint32_t *cols,
                              const DecPlan *plans, uint32_t count, uint8_t *status, uint32_t *failed,
                              const uint8_t *gtab, const MulTab *tw, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_decode_tail, dim3(count), dim3(64), 0, s, rx, L, n, (uint32_t)tl, cols, plans, status,
                       failed, gtab, tw);
    return hipGetLastError();
}

}  // namespace ars
