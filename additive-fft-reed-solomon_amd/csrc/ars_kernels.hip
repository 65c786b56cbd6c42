// ars_kernels.hip -- one-codeword-per-launch kernels: thin wrappers that stage
// the tables into LDS and run the fused tile bodies of ars_tiles.hpp over
// buffer-resource I/O (a per-lane byte offset plus scalar per-shard offsets).
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_device.hpp"
#include "ars_kernels.hpp"
#include "ars_tiles.hpp"

namespace ars {

// Copies WORDS dwords from global memory to LDS with a block of THREADS
// threads (the block size the kernel is launched with).  The loop is unrolled
// and every load is issued before the first LDS write, so a thread waits for
// one memory round trip rather than one per word.
template <int THREADS, int WORDS>
__device__ __forceinline__ void stage_words(uint32_t *dst, const uint32_t *__restrict__ src) {
    constexpr int K = (WORDS + THREADS - 1) / THREADS;
    uint32_t v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t i = threadIdx.x + (uint32_t)(k * THREADS);
        v[k] = i < (uint32_t)WORDS ? src[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t i = threadIdx.x + (uint32_t)(k * THREADS);
        if (i < (uint32_t)WORDS) dst[i] = v[k];
    }
}

template <int THREADS>
__device__ __forceinline__ void stage_twiddles(MulTab *s_tw, const MulTab *__restrict__ g_tw) {
    stage_words<THREADS, kTwiddleEntries * 6>(reinterpret_cast<uint32_t *>(s_tw),
                                              reinterpret_cast<const uint32_t *>(g_tw));
}

// The erased flags are staged as whole words: s_er must be 4-byte aligned
// and hold N rounded up to a multiple of 4 bytes (RecPrep::erased has 256).
template <int THREADS, int N>
__device__ __forceinline__ void stage_prep(MulTab *s_tab, uint8_t *s_er, const RecPrep *__restrict__ prep) {
    stage_words<THREADS, N * 6>(reinterpret_cast<uint32_t *>(s_tab), reinterpret_cast<const uint32_t *>(prep->tab));
    stage_words<THREADS, (N + 3) / 4>(reinterpret_cast<uint32_t *>(s_er),
                                      reinterpret_cast<const uint32_t *>(prep->erased));
}

// Shard group of this lane (wave-uniform when a wave holds one group).
template <int PQ>
__device__ __forceinline__ int tile_q() {
    const int q = threadIdx.x / PQ;
    if constexpr (PQ % 64 == 0) return __builtin_amdgcn_readfirstlane(q);
    return q;
}

template <int TL, int AM, bool ACC>
__global__ void __launch_bounds__(256) k_encode_small(const uint8_t *__restrict__ msg,
                                                      uint8_t *__restrict__ par, uint32_t L,
                                                      uint32_t chunks,
                                                      const MulTab *__restrict__ g_tw) {
    constexpr int T = 1 << TL;
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles<256>(tw, g_tw);
    __syncthreads();
    BufIO<AM> io = make_bufio<AM>(msg, chunks * T * L, par, T * L, blockIdx.x * blockDim.x + threadIdx.x, L);
    body_encode_small<TL, ACC>(io, chunks, tw);
}

// ONE: the message is a single chunk (k == T), known at compile time, so the
// chunk loop and its accumulator disappear (fewer VGPRs, higher occupancy).
template <int TL, int PQ, int AM, bool ACC, int NBUF, bool ONE>
__global__ void __launch_bounds__((Tile<TL, PQ>::THREADS))
    k_encode_tile(const uint8_t *__restrict__ msg, uint8_t *__restrict__ par, uint32_t L,
                  uint32_t chunks, const MulTab *__restrict__ g_tw) {
    using TT = Tile<TL, PQ>;
    __shared__ uint32_t tile[NBUF * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    BufIO<AM> io = make_bufio<AM>(msg, chunks * TT::N * L, par, TT::N * L, blockIdx.x * PQ + p, L);
    // the first chunk's rows are requested before the twiddles are staged, so
    // the two latencies overlap (measured faster for these 4-wave workgroups)
    uint32_t x0[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) x0[j] = io.ld_in((uint32_t)(q * 16), j);
    stage_twiddles<TT::THREADS>(tw, g_tw);
    __syncthreads();
    // with one shard group per wave the twiddle indices are wave-uniform:
    // high table words by scalar loads (see uni_levels)
    body_encode_tile<TL, PQ, ACC, NBUF, BufIO<AM>, (PQ % 64 == 0)>(io, q, p, tile, ONE ? 1u : chunks, tw, x0, g_tw);
}

template <int KL, int AM>
__global__ void __launch_bounds__(256) k_encode_kpow2_small(const uint8_t *__restrict__ msg,
                                                            uint8_t *__restrict__ par, uint32_t L,
                                                            uint32_t blocks, uint32_t msg_base,
                                                            const MulTab *__restrict__ g_tw) {
    constexpr int K = 1 << KL;
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles<256>(tw, g_tw);
    __syncthreads();
    BufIO<AM> io = make_bufio<AM>(msg, K * L, par, blocks * K * L, blockIdx.x * blockDim.x + threadIdx.x, L);
    body_encode_kpow2_small<KL>(io, blocks, msg_base, tw);
}

template <int KL, int PQ, int AM>
__global__ void __launch_bounds__((Tile<KL, PQ>::THREADS))
    k_encode_kpow2_tile(const uint8_t *__restrict__ msg, uint8_t *__restrict__ par, uint32_t L,
                        uint32_t blocks, uint32_t msg_base, const MulTab *__restrict__ g_tw) {
    using TT = Tile<KL, PQ>;
    __shared__ uint32_t tile[2 * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles<TT::THREADS>(tw, g_tw);
    __syncthreads();
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    BufIO<AM> io = make_bufio<AM>(msg, TT::N * L, par, blocks * TT::N * L, blockIdx.x * PQ + p, L);
    body_encode_kpow2_tile<KL, PQ, 2, BufIO<AM>, (PQ % 64 == 0)>(io, q, p, tile, blocks, msg_base, tw, g_tw);
}

__global__ void __launch_bounds__(256) k_recover_prep(RecPrep *__restrict__ out, uint32_t n,
                                                      uint32_t e, PosArg pos,
                                                      const uint8_t *__restrict__ gtab) {
    recover_prep_body(out, n, e, pos.p, gtab);
}

template <int NL, int AM>
__global__ void __launch_bounds__(256) k_recover_small(uint8_t *__restrict__ data, uint32_t L,
                                                       const RecPrep *__restrict__ prep,
                                                       const MulTab *__restrict__ g_tw) {
    constexpr int N = 1 << NL;
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ MulTab s_tab[N];
    __shared__ __align__(4) uint8_t s_er[(N + 3) / 4 * 4];
    stage_twiddles<256>(tw, g_tw);
    stage_prep<256, N>(s_tab, s_er, prep);
    __syncthreads();
    BufIO<AM> io = make_bufio<AM>(data, N * L, data, N * L, blockIdx.x * blockDim.x + threadIdx.x, L);
    LdsTabs tabs{s_tab, s_er};
    body_recover_small<NL, BufIO<AM>, LdsTabs, true>(io, tabs, tw, g_tw);
}

template <int NL, int PQ, bool AL, int NBUF>
__global__ void __launch_bounds__((Tile<NL, PQ>::THREADS))
    k_recover_tile(uint8_t *__restrict__ data, uint32_t L, const RecPrep *__restrict__ prep,
                   const MulTab *__restrict__ g_tw) {
    using TT = Tile<NL, PQ>;
    constexpr int N = TT::N;
    __shared__ uint32_t tile[NBUF * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ MulTab s_tab[N];
    __shared__ __align__(4) uint8_t s_er[N];
    // The rows are loaded after the barrier: with 8-wave workgroups, loading
    // them first would hold every wave at the barrier until the slowest
    // wave's rows arrived (measured slower than staging first).
    stage_twiddles<TT::THREADS>(tw, g_tw);
    stage_prep<TT::THREADS, N>(s_tab, s_er, prep);
    __syncthreads();
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    BufIO<AL> io = make_bufio<AL>(data, N * L, data, N * L, blockIdx.x * PQ + p, L);
    LdsTabs tabs{s_tab, s_er};
    body_recover_tile<NL, PQ, NBUF>(io, tabs, q, p, tile, tw);
}

// Recover with the erasure plan computed by each workgroup itself, for
// launches of at most one workgroup per compute unit: every workgroup then
// runs in the launch's first wave, so the plan costs one workgroup's time
// instead of a second launch (the plan kernel) ahead of this one.  The tile
// form keeps the plan's scratch in its tile buffer, idle until the first
// exchange.
template <int NL, int AM>
__global__ void __launch_bounds__(256) k_recover_small_fused(uint8_t *__restrict__ data, uint32_t L, PosArg pos,
                                                             uint32_t e, const uint8_t *__restrict__ gtab,
                                                             const MulTab *__restrict__ g_tw) {
    constexpr int N = 1 << NL;
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ MulTab s_tab[N];
    __shared__ __align__(4) uint8_t s_er[(N + 3) / 4 * 4];
    __shared__ __align__(4) uint8_t s_scratch[1536];
    stage_twiddles<256>(tw, g_tw);
    recover_plan<256>(PlanToLds{s_tab, s_er}, N, e, pos.p, gtab,
                      PlanScratch{s_scratch, s_scratch + 512, s_scratch + 768, s_scratch + 1024,
                                  reinterpret_cast<uint32_t *>(s_scratch + 1280)});
    __syncthreads();
    BufIO<AM> io = make_bufio<AM>(data, N * L, data, N * L, blockIdx.x * blockDim.x + threadIdx.x, L);
    LdsTabs tabs{s_tab, s_er};
    body_recover_small<NL, BufIO<AM>, LdsTabs, true>(io, tabs, tw, g_tw);
}

template <int NL, int PQ, bool AL, int NBUF>
__global__ void __launch_bounds__((Tile<NL, PQ>::THREADS))
    k_recover_tile_fused(uint8_t *__restrict__ data, uint32_t L, PosArg pos, uint32_t e,
                         const uint8_t *__restrict__ gtab, const MulTab *__restrict__ g_tw) {
    using TT = Tile<NL, PQ>;
    constexpr int N = TT::N;
    __shared__ uint32_t tile[NBUF * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ MulTab s_tab[N];
    __shared__ __align__(4) uint8_t s_er[N];
    static_assert(NBUF * TT::WORDS * 4 >= 1536, "the plan's scratch must fit the tile buffer");
    stage_twiddles<TT::THREADS>(tw, g_tw);
    uint8_t *sc = reinterpret_cast<uint8_t *>(tile);
    recover_plan<TT::THREADS>(PlanToLds{s_tab, s_er}, N, e, pos.p, gtab,
                              PlanScratch{sc, sc + 512, sc + 768, sc + 1024, reinterpret_cast<uint32_t *>(sc + 1280)});
    __syncthreads();  // the plan is in s_tab / s_er, and the tile buffer is free again
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    BufIO<AL> io = make_bufio<AL>(data, N * L, data, N * L, blockIdx.x * PQ + p, L);
    LdsTabs tabs{s_tab, s_er};
    body_recover_tile<NL, PQ, NBUF>(io, tabs, q, p, tile, tw);
}

// Recover of dword-addressable codewords (one exchange buffer, see
// body_recover_tile1), capped at 128 VGPRs (4 waves per SIMD; uncapped, the
// compiler takes 157 for 256 points): each workgroup stages the tables once and walks the
// tiles blockIdx.x, blockIdx.x + gridDim.x, ...  The next tile's input rows
// are requested once the current tile is back in the phase-A layout, so their
// latency overlaps the phase-A FFT and the output of the current one.
template <int NL, int PQ, bool OOB, int AM>
__device__ __forceinline__ void recover_stream(uint8_t *data, uint32_t L, uint32_t tiles, const MulTab *s_tab,
                                               const uint8_t *s_er, uint32_t *tile, const MulTab *tw) {
    using TT = Tile<NL, PQ>;
    constexpr uint32_t N = TT::N;
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    const LdsTabs tabs{s_tab, s_er};
    const Flags16 f = group_flags<PQ % 64 == 0>(tabs, q);  // the same as UNI below
    uint32_t t = blockIdx.x;
    constexpr bool UNI = PQ % 64 == 0;
    BufIO<AM, UNI> io = make_bufio<AM, UNI>(data, N * L, data, N * L, t * PQ + p, L);
    uint32_t rows[16];
    recover_rows_in<OOB>(io, f, q, rows);
    for (;;) {
        const uint32_t tn = t + gridDim.x;
        // wide rows keep their row addresses on the scalar unit but need two more
        // VGPRs; with four multiplies per lookup burst the 256-point body then
        // spills 8 bytes, with two it fits
        constexpr int GRP = (AM == 3 && NL >= 8) ? 2 : kBurstMuls;
        // the next tile's addressing is formed only where its rows are
        // requested, so it is not live across phase B
        body_recover_tile1<NL, PQ, GRP>(io, tabs, f, q, p, tile, tw, rows, [&] {
            if (tn < tiles)
                recover_rows_in<OOB>(make_bufio<AM, UNI>(data, N * L, data, N * L, tn * PQ + p, L), f, q, rows);
        });
        if (tn >= tiles) break;
        io = make_bufio<AM, UNI>(data, N * L, data, N * L, tn * PQ + p, L);
        t = tn;
        __syncthreads();  // every lane's get_a of this tile is done before the next put_a
    }
}

template <int NL, int PQ, bool OOB, int AM>
__global__ void __launch_bounds__((Tile<NL, PQ>::THREADS)) __attribute__((amdgpu_waves_per_eu(4)))
    k_recover_stream(uint8_t *__restrict__ data, uint32_t L, uint32_t tiles, const RecPrep *__restrict__ prep,
                     const MulTab *__restrict__ g_tw) {
    using TT = Tile<NL, PQ>;
    constexpr int N = TT::N;
    __shared__ uint32_t tile[TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ MulTab s_tab[N];
    __shared__ __align__(16) uint8_t s_er[N];
    stage_twiddles<TT::THREADS>(tw, g_tw);
    stage_prep<TT::THREADS, N>(s_tab, s_er, prep);
    __syncthreads();
    recover_stream<NL, PQ, OOB, AM>(data, L, tiles, s_tab, s_er, tile, tw);
}

// The same with the erasure plan computed by each workgroup (launches of at
// most one workgroup per compute unit; see k_recover_tile_fused).
template <int NL, int PQ, bool OOB, int AM>
__global__ void __launch_bounds__((Tile<NL, PQ>::THREADS)) __attribute__((amdgpu_waves_per_eu(4)))
    k_recover_stream_fused(uint8_t *__restrict__ data, uint32_t L, uint32_t tiles, PosArg pos, uint32_t e,
                           const uint8_t *__restrict__ gtab, const MulTab *__restrict__ g_tw) {
    using TT = Tile<NL, PQ>;
    constexpr int N = TT::N;
    __shared__ uint32_t tile[TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ MulTab s_tab[N];
    __shared__ __align__(16) uint8_t s_er[N];
    static_assert(TT::WORDS * 4 >= 1536, "the plan's scratch must fit the tile buffer");
    stage_twiddles<TT::THREADS>(tw, g_tw);
    uint8_t *sc = reinterpret_cast<uint8_t *>(tile);
    recover_plan<TT::THREADS>(PlanToLds{s_tab, s_er}, N, e, pos.p, gtab,
                              PlanScratch{sc, sc + 512, sc + 768, sc + 1024, reinterpret_cast<uint32_t *>(sc + 1280)});
    __syncthreads();  // the plan is in s_tab / s_er, and the tile buffer is free again
    recover_stream<NL, PQ, OOB, AM>(data, L, tiles, s_tab, s_er, tile, tw);
}

template <int TL, int AM>
__global__ void __launch_bounds__(256) k_recompute_small(uint8_t *__restrict__ data, uint32_t L,
                                                         const MulTab *__restrict__ g_tw) {
    constexpr int T = 1 << TL;
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles<256>(tw, g_tw);
    __syncthreads();
    BufIO<AM> io = make_bufio<AM>(data, 2 * T * L, data, 2 * T * L, blockIdx.x * blockDim.x + threadIdx.x, L);
    body_recompute_small<TL>(io, tw);
}

template <int TL, int PQ, int AM>
__global__ void __launch_bounds__((Tile<TL, PQ>::THREADS))
    k_recompute_tile(uint8_t *__restrict__ data, uint32_t L, const MulTab *__restrict__ g_tw) {
    using TT = Tile<TL, PQ>;
    __shared__ uint32_t tile[2 * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    stage_twiddles<TT::THREADS>(tw, g_tw);
    __syncthreads();
    const int p = threadIdx.x % PQ, q = tile_q<PQ>();
    BufIO<AM> io = make_bufio<AM>(data, 2 * TT::N * L, data, 2 * TT::N * L, blockIdx.x * PQ + p, L);
    body_recompute_tile<TL, PQ, BufIO<AM>, (PQ % 64 == 0)>(io, q, p, tile, tw, g_tw);
}

// ---------------------------------------------------------------------------
// Launchers.
static inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

template <int TL>
constexpr int tile_pq() { return TL >= 8 ? 32 : 64; }

// Tile variants the autotuner chooses among: column quads per workgroup (PQ: 16, 32, 64)
// and LDS buffers (NBUF: 2 = one barrier per exchange, 1 = half the LDS).
static bool variant_ok(int tl, int pq, int nbuf) {
    if (tl <= 4) return true;
    if (pq != 16 && pq != 32 && pq != 64) return false;
    if (nbuf != 1 && nbuf != 2) return false;
    return ((1 << (tl - 4)) * pq) <= 1024;
}

// Column quads a launch covers: access mode 2 starts r bytes before the
// first column (see BufIO), so the quads span r + L bytes.
static inline uint64_t launch_quads(int am, const void *base, uint32_t L) {
    const uint64_t r = am == 2 ? (reinterpret_cast<uintptr_t>(base) & 3u) : 0u;
    return ((uint64_t)L + r + 3) / 4;
}

template <int TL, int AM, bool ACC>
static hipError_t enc_one(const uint8_t *msg, uint8_t *par, uint32_t L, uint32_t chunks,
                          const MulTab *tw, hipStream_t s, int pq, int nbuf) {
    const uint64_t quads = launch_quads(AM, msg, L);
    if constexpr (TL <= 4) {
        hipLaunchKernelGGL((k_encode_small<TL, AM, ACC>), dim3(ceil_div(quads, 256)), dim3(256), 0, s,
                           msg, par, L, chunks, tw);
    } else {
#define ENC_V(PQ_, NB_)                                                                              \
    if (pq == PQ_ && nbuf == NB_) {                                                                  \
        if (chunks == 1)                                                                             \
            hipLaunchKernelGGL((k_encode_tile<TL, PQ_, AM, ACC, NB_, true>), dim3(ceil_div(quads, PQ_)), \
                               dim3(Tile<TL, PQ_>::THREADS), 0, s, msg, par, L, chunks, tw);         \
        else                                                                                         \
            hipLaunchKernelGGL((k_encode_tile<TL, PQ_, AM, ACC, NB_, false>), dim3(ceil_div(quads, PQ_)), \
                               dim3(Tile<TL, PQ_>::THREADS), 0, s, msg, par, L, chunks, tw);         \
        return hipGetLastError();                                                                    \
    }
        if constexpr (AM == 3) {
            ENC_V(64, 1) ENC_V(64, 2)
        } else {
            ENC_V(16, 1) ENC_V(16, 2) ENC_V(32, 1) ENC_V(32, 2) ENC_V(64, 1) ENC_V(64, 2)
        }
#undef ENC_V
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int TL>
static hipError_t enc_acc(int am, bool acc, const uint8_t *msg, uint8_t *par, uint32_t L,
                          uint32_t chunks, const MulTab *tw, hipStream_t s, int pq, int nbuf) {
    switch (am) {
        case 1: return acc ? enc_one<TL, 1, true>(msg, par, L, chunks, tw, s, pq, nbuf)
                           : enc_one<TL, 1, false>(msg, par, L, chunks, tw, s, pq, nbuf);
        case 2: return acc ? enc_one<TL, 2, true>(msg, par, L, chunks, tw, s, pq, nbuf)
                           : enc_one<TL, 2, false>(msg, par, L, chunks, tw, s, pq, nbuf);
        case 3: return acc ? enc_one<TL, 3, true>(msg, par, L, chunks, tw, s, pq, nbuf)
                           : enc_one<TL, 3, false>(msg, par, L, chunks, tw, s, pq, nbuf);
        default: return acc ? enc_one<TL, 0, true>(msg, par, L, chunks, tw, s, pq, nbuf)
                            : enc_one<TL, 0, false>(msg, par, L, chunks, tw, s, pq, nbuf);
    }
}

hipError_t launch_encode_ref(int tl, int am, bool acc, const uint8_t *msg, uint8_t *par,
                             uint32_t L, uint32_t chunks, const MulTab *tw, hipStream_t s,
                             int pq, int nbuf) {
    if (pq == 0 || (am == 3 && tl > 4)) pq = am == 3 ? 64 : default_pq(0, tl);
    if (nbuf == 0) nbuf = default_nbuf(0, tl);
    if (!variant_ok(tl, pq, nbuf)) return hipErrorInvalidValue;
    switch (tl) {
        case 0: return enc_acc<0>(am, acc, msg, par, L, chunks, tw, s, pq, nbuf);
        case 1: return enc_acc<1>(am, acc, msg, par, L, chunks, tw, s, pq, nbuf);
        case 2: return enc_acc<2>(am, acc, msg, par, L, chunks, tw, s, pq, nbuf);
        case 3: return enc_acc<3>(am, acc, msg, par, L, chunks, tw, s, pq, nbuf);
        case 4: return enc_acc<4>(am, acc, msg, par, L, chunks, tw, s, pq, nbuf);
        case 5: return enc_acc<5>(am, acc, msg, par, L, chunks, tw, s, pq, nbuf);
        case 6: return enc_acc<6>(am, acc, msg, par, L, chunks, tw, s, pq, nbuf);
        case 7: return enc_acc<7>(am, acc, msg, par, L, chunks, tw, s, pq, nbuf);
        default: return hipErrorInvalidValue;
    }
}

template <int KL, int AM>
static hipError_t kp_one(const uint8_t *msg, uint8_t *par, uint32_t L, uint32_t blocks,
                         uint32_t msg_base, const MulTab *tw, hipStream_t s) {
    const uint64_t quads = ((uint64_t)L + 3) / 4;
    if constexpr (KL <= 4) {
        hipLaunchKernelGGL((k_encode_kpow2_small<KL, AM>), dim3(ceil_div(quads, 256)), dim3(256), 0, s,
                           msg, par, L, blocks, msg_base, tw);
    } else {
        constexpr int PQ = tile_pq<KL>();
        static_assert(AM != 3 || PQ % 64 == 0, "wide rows need one shard group per wave");
        hipLaunchKernelGGL((k_encode_kpow2_tile<KL, PQ, AM>), dim3(ceil_div(quads, PQ)),
                           dim3(Tile<KL, PQ>::THREADS), 0, s, msg, par, L, blocks, msg_base, tw);
    }
    return hipGetLastError();
}

hipError_t launch_encode_kpow2(int kl, int am, const uint8_t *msg, uint8_t *par, uint32_t L,
                               uint32_t blocks, uint32_t msg_base, const MulTab *tw, hipStream_t s) {
#define KP_CASE(K)                                                                \
    case K: return am == 3 ? kp_one<K, 3>(msg, par, L, blocks, msg_base, tw, s)   \
                 : am == 1 ? kp_one<K, 1>(msg, par, L, blocks, msg_base, tw, s)   \
                           : kp_one<K, 0>(msg, par, L, blocks, msg_base, tw, s);
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

// Erased rows may be read at the buffer's end (see recover_rows_in) when no
// row offset added to it wraps past 2^32 (8 bytes of slack for access mode 2's
// rounding).
// Mode 3 reads them at the end of the row's own resource, which always works.
static bool oob_ok(int nl, int am, uint32_t L) {
    return am == 3 || ((uint64_t)(1u << nl) + 15u) * L + 8u < (1ull << 32);
}

// Workgroups of `kernel` resident per compute unit (at least 1).
template <class K>
static uint32_t resident_groups(K kernel, int threads) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess || nb < 1) nb = 1;
    return (uint32_t)nb;
}

// The streaming recover's grid: twice the workgroups the device holds at
// once (measured faster than exactly one round of resident workgroups), at
// most one per tile.
template <int NL, int PQ, int AM>
static uint32_t stream_grid(uint32_t tiles, uint32_t cus) {
    static const uint32_t occ = resident_groups(k_recover_stream<NL, PQ, true, AM>, Tile<NL, PQ>::THREADS);
    const uint64_t g = 2ull * occ * (cus ? cus : 256u);
    return (uint32_t)(g < tiles ? g : tiles);
}

template <int NL, int AM>
static hipError_t rec_one(uint8_t *data, uint32_t L, const RecPrep *prep, const MulTab *tw,
                          hipStream_t s, int pq, int nbuf, uint32_t cus) {
    const uint64_t quads = launch_quads(AM, data, L);
    if constexpr (NL <= 4) {
        hipLaunchKernelGGL((k_recover_small<NL, AM>), dim3(ceil_div(quads, 256)), dim3(256), 0, s,
                           data, L, prep, tw);
        return hipGetLastError();
    } else {
        if constexpr (AM != 0) {
            if (nbuf == 1 || AM >= 2) {
#define RS_V(PQ_)                                                                                    \
    if (pq == PQ_) {                                                                                 \
        const uint32_t tiles = ceil_div(quads, PQ_), grid = stream_grid<NL, PQ_, AM>(tiles, cus);    \
        if (oob_ok(NL, AM, L))                                                                       \
            hipLaunchKernelGGL((k_recover_stream<NL, PQ_, true, AM>), dim3(grid),                    \
                               dim3(Tile<NL, PQ_>::THREADS), 0, s, data, L, tiles, prep, tw);        \
        else                                                                                         \
            hipLaunchKernelGGL((k_recover_stream<NL, PQ_, false, AM>), dim3(grid),                   \
                               dim3(Tile<NL, PQ_>::THREADS), 0, s, data, L, tiles, prep, tw);        \
        return hipGetLastError();                                                                    \
    }
                if constexpr (AM == 3) {
                    RS_V(64)  // wide rows need one shard group per wave
                } else {
                    RS_V(16) RS_V(32) RS_V(64)
                }
#undef RS_V
                return hipErrorInvalidValue;
            }
        }
        if constexpr (AM <= 1) {
            constexpr bool AL = AM == 1;
#define REC_V(PQ_, NB_)                                                                              \
    if (pq == PQ_ && nbuf == NB_) {                                                                  \
        hipLaunchKernelGGL((k_recover_tile<NL, PQ_, AL, NB_>), dim3(ceil_div(quads, PQ_)),           \
                           dim3(Tile<NL, PQ_>::THREADS), 0, s, data, L, prep, tw);                   \
        return hipGetLastError();                                                                    \
    }
            REC_V(16, 1) REC_V(16, 2) REC_V(32, 1) REC_V(32, 2) REC_V(64, 1) REC_V(64, 2)
#undef REC_V
        }
        return hipErrorInvalidValue;
    }
}

hipError_t launch_recover(int nl, int am, uint8_t *data, uint32_t L, const RecPrep *prep,
                          const MulTab *tw, hipStream_t s, int pq, int nbuf, uint32_t cus) {
    if (pq == 0) pq = (am != 0 || nl < 8) ? default_pq(1, nl) : 32;  // byte path: 32 quads at 256 points
    if (nbuf == 0) nbuf = am != 0 ? default_nbuf(1, nl) : 2;
    if (am == 3 && nl > 4) pq = 64, nbuf = 1;  // the only wide variant
    if (!variant_ok(nl, pq, nbuf)) return hipErrorInvalidValue;
#define RC_CASE(K)                                                                                  \
    case K:                                                                                         \
        return am == 1 ? rec_one<K, 1>(data, L, prep, tw, s, pq, nbuf, cus)                         \
             : am == 2 ? rec_one<K, 2>(data, L, prep, tw, s, pq, nbuf, cus)                         \
             : am == 3 ? rec_one<K, 3>(data, L, prep, tw, s, pq, nbuf, cus)                         \
                       : rec_one<K, 0>(data, L, prep, tw, s, pq, nbuf, cus);
    switch (nl) {
        RC_CASE(1) RC_CASE(2) RC_CASE(3) RC_CASE(4) RC_CASE(5) RC_CASE(6) RC_CASE(7) RC_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef RC_CASE
}

template <int NL, int AM>
static hipError_t rec_fused_one(uint8_t *data, uint32_t L, const PosArg &pos, uint32_t e, const uint8_t *gtab,
                                const MulTab *tw, hipStream_t s, uint32_t max_groups) {
    const uint64_t quads = launch_quads(AM, data, L);
    if constexpr (NL <= 4) {
        const uint32_t grid = ceil_div(quads, 256);
        if (grid > max_groups) return hipErrorNotSupported;
        hipLaunchKernelGGL((k_recover_small_fused<NL, AM>), dim3(grid), dim3(256), 0, s, data, L, pos, e, gtab, tw);
    } else if constexpr (AM == 3) {
        // a codeword of 4 GiB or more has at least 2^22 column quads, far more
        // than one workgroup per compute unit covers
        return hipErrorNotSupported;
    } else if constexpr (AM != 0) {
        constexpr int PQ = default_pq(1, NL);
        const uint32_t grid = ceil_div(quads, PQ);
        if (grid > max_groups) return hipErrorNotSupported;
        if (oob_ok(NL, AM, L))
            hipLaunchKernelGGL((k_recover_stream_fused<NL, PQ, true, AM>), dim3(grid), dim3(Tile<NL, PQ>::THREADS), 0,
                               s, data, L, grid, pos, e, gtab, tw);
        else
            hipLaunchKernelGGL((k_recover_stream_fused<NL, PQ, false, AM>), dim3(grid), dim3(Tile<NL, PQ>::THREADS),
                               0, s, data, L, grid, pos, e, gtab, tw);
    } else {
        constexpr int PQ = NL >= 8 ? 32 : 64, NB = 2;  // the byte path's default variant
        const uint32_t grid = ceil_div(quads, PQ);
        if (grid > max_groups) return hipErrorNotSupported;
        hipLaunchKernelGGL((k_recover_tile_fused<NL, PQ, false, NB>), dim3(grid), dim3(Tile<NL, PQ>::THREADS), 0, s,
                           data, L, pos, e, gtab, tw);
    }
    return hipGetLastError();
}

hipError_t launch_recover_fused(int nl, int am, uint8_t *data, uint32_t L, const PosArg &pos, uint32_t e,
                                const uint8_t *gtab, const MulTab *tw, hipStream_t s, uint32_t max_groups) {
#define RF_CASE(K)                                                                                  \
    case K:                                                                                         \
        return am == 1 ? rec_fused_one<K, 1>(data, L, pos, e, gtab, tw, s, max_groups)              \
             : am == 2 ? rec_fused_one<K, 2>(data, L, pos, e, gtab, tw, s, max_groups)              \
             : am == 3 ? rec_fused_one<K, 3>(data, L, pos, e, gtab, tw, s, max_groups)              \
                       : rec_fused_one<K, 0>(data, L, pos, e, gtab, tw, s, max_groups);
    switch (nl) {
        RF_CASE(1) RF_CASE(2) RF_CASE(3) RF_CASE(4) RF_CASE(5) RF_CASE(6) RF_CASE(7) RF_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef RF_CASE
}

template <int TL, int AM>
static hipError_t rcp_one(uint8_t *data, uint32_t L, const MulTab *tw, hipStream_t s) {
    const uint64_t quads = ((uint64_t)L + 3) / 4;
    if constexpr (TL <= 4) {
        hipLaunchKernelGGL((k_recompute_small<TL, AM>), dim3(ceil_div(quads, 256)), dim3(256), 0, s,
                           data, L, tw);
    } else {
        constexpr int PQ = tile_pq<TL>();
        static_assert(AM != 3 || PQ % 64 == 0, "wide rows need one shard group per wave");
        hipLaunchKernelGGL((k_recompute_tile<TL, PQ, AM>), dim3(ceil_div(quads, PQ)),
                           dim3(Tile<TL, PQ>::THREADS), 0, s, data, L, tw);
    }
    return hipGetLastError();
}

hipError_t launch_recompute(int tl, int am, uint8_t *data, uint32_t L, const MulTab *tw,
                            hipStream_t s) {
#define RP_CASE(K)                                                 \
    case K: return am == 3 ? rcp_one<K, 3>(data, L, tw, s)         \
                 : am == 1 ? rcp_one<K, 1>(data, L, tw, s)         \
                           : rcp_one<K, 0>(data, L, tw, s);
    switch (tl) {
        RP_CASE(0) RP_CASE(1) RP_CASE(2) RP_CASE(3) RP_CASE(4) RP_CASE(5) RP_CASE(6) RP_CASE(7)
        default: return hipErrorInvalidValue;
    }
#undef RP_CASE
}

}  // namespace ars
