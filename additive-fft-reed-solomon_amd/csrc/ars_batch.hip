// ars_batch.hip -- many codewords of uneven sizes in one persistent launch.
//
// Device side.  The host flattens a batch into work items (see BItem): each
// item is one tile of one code shape whose lanes carry column quads of one or
// more codewords (BSeg), so a large codeword spans many items and many small
// codewords share one.  A fixed grid of workgroups (sized to the occupancy)
// walks the regions -- one per transform size, body variant and addressing,
// largest transforms first -- pulling each region's items from its own
// atomic counter until that region is drained.  Each region is its own
// inlined loop around one tile body, so no body's register allocation
// constrains another's and no body sits behind a function call (a call
// would pass the buffer descriptors in VGPRs and save registers to scratch).
// Twiddles are indexed by absolute point index, so one LDS copy serves every
// code shape.
//
// The three launches of a batch -- counter reset, erasure plans, persistent
// kernel -- read their sizes from a device header, so a hipGraph captured
// once replays any batch whose descriptors fit the arena.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_batch.hpp"
#include "ars_device.hpp"
#include "ars_tiles.hpp"

namespace ars {

// Packed items: each lane addresses its own codeword as a 32-bit offset
// from the item's base pointers, through one scalar buffer resource.
struct PackIO {
    Rows in, out;
    uint32_t vin, vout, L, nb;
    __device__ __forceinline__ uint32_t ld_in(uint32_t g, int j) const {
        return __builtin_amdgcn_raw_buffer_load_b32(in.r, (int)(vin + (g + (uint32_t)j) * L), 0, 0);
    }
    __device__ __forceinline__ uint32_t ld_out(uint32_t g, int j) const {
        return __builtin_amdgcn_raw_buffer_load_b32(out.r, (int)(vout + (g + (uint32_t)j) * L), 0, 0);
    }
    __device__ __forceinline__ void st_out(uint32_t g, int j, uint32_t v) const {
        if (nb) __builtin_amdgcn_raw_buffer_store_b32(v, out.r, (int)(vout + (g + (uint32_t)j) * L), 0, 0);
    }
};

__global__ void k_batch_reset(uint8_t *arena) {
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBatchRegions; i += blockDim.x)
        reinterpret_cast<BHeader *>(arena)->counters[i] = 0;
}

// Erasure plans for every recover job of the batch (fixed grid, loops).
__global__ void __launch_bounds__(256) k_batch_prep(uint8_t *__restrict__ arena,
                                                    const uint8_t *__restrict__ gtab) {
    const BHeader *h = reinterpret_cast<const BHeader *>(arena);
    const uint32_t np = h->nprep;
    const BPrep *preps = reinterpret_cast<const BPrep *>(arena + h->off_preps);
    const uint8_t *pos = arena + h->off_pos;
    for (uint32_t i = blockIdx.x; i < np; i += gridDim.x) {
        const BPrep d = preps[i];
        recover_prep_body(d.plan, d.n, d.e, pos + d.pos_off, gtab);
        __syncthreads();
    }
}

// Resolve this lane's codeword and column quad for repetition r of the
// item.  Lanes no segment covers, or past the end of their codeword's shard,
// read quad 0 of the item's first codeword (always in bounds) and store
// nothing.
__device__ __forceinline__ void lane_job(const BItem &it, const BSeg *__restrict__ segs,
                                         const BJob *__restrict__ jobs, uint32_t lane, uint32_t lanes,
                                         uint32_t r, BJob &job, uint32_t &col, uint32_t &nb,
                                         uint32_t &si) {
    const uint32_t nseg = it.nseg_rep & 0xffffu;
    uint32_t jid = segs[it.seg0].job, quad = 0;
    bool valid = false;
    si = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
        const BSeg sg = segs[it.seg0 + s];
        if (lane - sg.lane0 < sg.nlanes) {
            jid = sg.job;
            quad = sg.quad0 + (lane - sg.lane0) + r * lanes;
            valid = true;
            si = s;
        }
    }
    job = jobs[jid];
    valid = valid && quad < job.quads;
    col = valid ? quad * 4u : 0u;
    nb = valid ? 4u : 0u;
}

// Recover items: copy the plans of the item's codewords into LDS
// (segment s at slots [s*N, (s+1)*N)).
template <int TL>
__device__ __forceinline__ void stage_item_plans(const BItem &it, const BSeg *__restrict__ segs,
                                                 const BJob *__restrict__ jobs, MulTab *s_tab,
                                                 uint8_t *s_er) {
    constexpr uint32_t N = 1u << TL;
    const uint32_t nseg = it.nseg_rep & 0xffffu;
    for (uint32_t i = threadIdx.x; i < nseg * N; i += blockDim.x) {
        const RecPrep *pl = jobs[segs[it.seg0 + i / N].job].plan;
        s_tab[i] = pl->tab[i % N];
        s_er[i] = pl->erased[i % N];
    }
}

// Descriptor fields are wave-uniform, but values read with vector loads from
// the (atomically updated) arena sit in VGPRs; a buffer resource built from
// them would force a readfirstlane loop around every load and store.  These
// helpers move them into SGPRs.
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
template <class T>
__device__ __forceinline__ T *uni(T *ptr) {
    const uint64_t v = reinterpret_cast<uint64_t>(ptr);
    const uint64_t lo = uni((uint32_t)v), hi = uni((uint32_t)(v >> 32));
    return reinterpret_cast<T *>(lo | (hi << 32));
}
__device__ __forceinline__ BItem uni(const BItem &x) {
    BItem y;
    y.meta = uni(x.meta);
    y.arg = uni(x.arg);
    y.seg0 = uni(x.seg0);
    y.nseg_rep = uni(x.nseg_rep);
    y.in_base = uni(x.in_base);
    y.out_base = uni(x.out_base);
    return y;
}
__device__ __forceinline__ BJob uni(const BJob &x) {
    BJob y;
    y.in = uni(x.in);
    y.out = uni(x.out);
    y.plan = uni(x.plan);
    y.L = uni(x.L);
    y.quads = uni(x.quads);
    y.in_bytes = uni(x.in_bytes);
    y.out_bytes = uni(x.out_bytes);
    return y;
}

// Tile body of variant V (see batch_region) for transform size 2^TL.
template <int TL, int V, class IO>
__device__ __forceinline__ void run_body(const BItem &it, IO &io, int q, int p, uint32_t *tile,
                                         const MulTab *tw, const MulTab *gtw, const MulTab *s_tab,
                                         const uint8_t *s_er, uint32_t si) {
    constexpr uint32_t PQ = batch_lanes(TL);
    if constexpr (TL <= 4) {
        if constexpr (V == 0 || V == 1) {
            body_encode_small<TL, V == 1>(io, it.arg, tw);
        } else if constexpr (V == 2) {
            body_encode_kpow2_small<TL>(io, it.arg & 0xffffu, it.arg >> 16, tw);
        } else if constexpr (TL >= 1) {
            // every lane runs the same transform: the twiddles' high words by
            // scalar loads (uni_levels)
            LdsTabs tabs{s_tab + si * (1u << TL), s_er + si * (1u << TL)};
            body_recover_small<TL, IO, LdsTabs, true>(io, tabs, tw, gtw);
        }
    } else if constexpr (V == 0 || V == 1) {
        // 64+ quads: a wave holds one shard group, so the tables' high words
        // come by scalar loads (see uni_levels)
        if constexpr (TL <= 7)
            body_encode_tile<TL, PQ, V == 1, 1, IO, (PQ % 64 == 0)>(io, q, p, tile, it.arg, tw, nullptr, gtw);
    } else if constexpr (V == 2) {
        if constexpr (TL <= 7)
            body_encode_kpow2_tile<TL, PQ, 1, IO, (PQ % 64 == 0)>(io, q, p, tile, it.arg & 0xffffu, it.arg >> 16, tw,
                                                                  gtw);
    } else {
        // The one-buffer streaming body (body_recover_tile1) measured slower
        // here: 32-quad items leave the erased flags lane-varying, and this
        // kernel's shared 128-VGPR budget spills, so its scratch reloads wait
        // for the atomic outputs (mixed batch 1.94 vs 2.24 TB/s, one-shape
        // (256,128) 64 KiB recover 1.53 vs 1.69 TB/s).
        LdsTabs tabs{s_tab + si * (1u << TL), s_er + si * (1u << TL)};
        body_recover_tile<TL, PQ, 1>(io, tabs, q, p, tile, tw);
    }
}

// Repetition r of an item: one codeword (PACK = 0: scalar job descriptor,
// buffer-resource addressing) or several (PACK = 1: per-lane offsets from the
// item's base pointers).
template <int TL, int V, int PACK>
__device__ __forceinline__ void run_tile(const BItem &it, const BSeg *segs, const BJob *jobs, uint32_t r,
                                         uint32_t *tile, const MulTab *tw, const MulTab *gtw,
                                         const MulTab *s_tab, const uint8_t *s_er) {
    constexpr uint32_t PQ = batch_lanes(TL);
    const int p = (int)(threadIdx.x % PQ);
    const int q = PQ % 64 == 0 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x / PQ)) : (int)(threadIdx.x / PQ);
    if constexpr (PACK == 0) {
        const uint32_t jid = uni(segs[it.seg0].job), quad0 = uni(segs[it.seg0].quad0);
        const BJob job = uni(jobs[jid]);
        const uint32_t quad = quad0 + (uint32_t)p + r * PQ;
        BufIO<true> io = make_bufio<true>(job.in, job.in_bytes, job.out, job.out_bytes,
                                          quad < job.quads ? quad : job.quads, job.L);
        run_body<TL, V>(it, io, q, p, tile, tw, gtw, s_tab, s_er, 0u);
    } else if constexpr (TL <= kBatchPackMaxTl) {
        BJob job;
        uint32_t col, nb, si;
        lane_job(it, segs, jobs, (uint32_t)p, PQ, r, job, col, nb, si);
        PackIO io;
        io.in = make_rows(it.in_base, 0xffffffffu);
        io.out = make_rows(it.out_base, 0xffffffffu);
        io.vin = (uint32_t)(job.in - it.in_base) + col;
        io.vout = (uint32_t)(job.out - it.out_base) + col;
        io.L = job.L;
        io.nb = nb;
        run_body<TL, V>(it, io, q, p, tile, tw, gtw, s_tab, s_er, si);
    }
}

// Drain region (TL, V, PACK): pull its items until its counter passes its
// range.  The two barriers around the item index also order the previous
// item's LDS reads (tile, plans) before this item's writes.
template <int TL, int V, int PACK>
__device__ __forceinline__ void run_region(BHeader *h, const BItem *items, const BSeg *segs, const BJob *jobs,
                                           uint32_t *tile, const MulTab *tw, const MulTab *gtw, MulTab *s_tab,
                                           uint8_t *s_er, uint32_t *s_item) {
    constexpr uint32_t g = batch_region(TL, V, PACK);
    const uint32_t begin = uni(h->begin[g]), end = uni(h->end[g]);
    if (begin >= end) return;
    for (;;) {
        if (threadIdx.x == 0) *s_item = begin + atomicAdd(&h->counters[g], 1u);
        __syncthreads();
        const uint32_t i = uni(*s_item);
        __syncthreads();
        if (i >= end) break;
        const BItem it = uni(items[i]);
        if constexpr (V == 3) {
            stage_item_plans<TL>(it, segs, jobs, s_tab, s_er);
            __syncthreads();
        }
        const uint32_t rep = it.nseg_rep >> 16;
        for (uint32_t r = 0; r < rep; ++r) {
            if (r) __syncthreads();  // the previous tile's LDS reads are done
            run_tile<TL, V, PACK>(it, segs, jobs, r, tile, tw, gtw, s_tab, s_er);
        }
    }
}

// Every region of transform size 2^TL, then the next smaller size.
template <int TL>
__device__ __forceinline__ void run_sizes(BHeader *h, const BItem *items, const BSeg *segs, const BJob *jobs,
                                          uint32_t *tile, const MulTab *tw, const MulTab *gtw, MulTab *s_tab,
                                          uint8_t *s_er, uint32_t *s_item) {
    if constexpr (TL >= 0) {
        if constexpr (TL <= 7) {  // encode transforms are at most 128 points
            run_region<TL, 0, 0>(h, items, segs, jobs, tile, tw, gtw, s_tab, s_er, s_item);
            run_region<TL, 1, 0>(h, items, segs, jobs, tile, tw, gtw, s_tab, s_er, s_item);
            run_region<TL, 2, 0>(h, items, segs, jobs, tile, tw, gtw, s_tab, s_er, s_item);
        }
        if constexpr (TL >= 1) run_region<TL, 3, 0>(h, items, segs, jobs, tile, tw, gtw, s_tab, s_er, s_item);
        if constexpr (TL <= kBatchPackMaxTl) {
            run_region<TL, 0, 1>(h, items, segs, jobs, tile, tw, gtw, s_tab, s_er, s_item);
            run_region<TL, 1, 1>(h, items, segs, jobs, tile, tw, gtw, s_tab, s_er, s_item);
            run_region<TL, 2, 1>(h, items, segs, jobs, tile, tw, gtw, s_tab, s_er, s_item);
            if constexpr (TL >= 1) run_region<TL, 3, 1>(h, items, segs, jobs, tile, tw, gtw, s_tab, s_er, s_item);
        }
        run_sizes<TL - 1>(h, items, segs, jobs, tile, tw, gtw, s_tab, s_er, s_item);
    }
}

__global__ void __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(4, 8)))
    k_batch(uint8_t *__restrict__ arena, const MulTab *__restrict__ g_tw) {
    __shared__ uint32_t tile[8192];  // 2^TL shards x batch_lanes(TL) quads, every TL >= 5
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ MulTab s_tab[kBatchPlanSlots];
    __shared__ uint8_t s_er[kBatchPlanSlots];
    __shared__ uint32_t s_item;
    BHeader *h = reinterpret_cast<BHeader *>(arena);
    const BItem *items = reinterpret_cast<const BItem *>(arena + h->off_items);
    const BSeg *segs = reinterpret_cast<const BSeg *>(arena + h->off_segs);
    const BJob *jobs = reinterpret_cast<const BJob *>(arena + h->off_jobs);
    {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(g_tw);
        uint32_t *dst = reinterpret_cast<uint32_t *>(tw);
        for (uint32_t i = threadIdx.x; i < (uint32_t)kTwiddleEntries * 6u; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    run_sizes<8>(h, items, segs, jobs, tile, tw, g_tw, s_tab, s_er, &s_item);
}

hipError_t launch_batch_reset(uint8_t *arena, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, arena);
    return hipGetLastError();
}

hipError_t launch_batch_prep(uint8_t *arena, const uint8_t *gtab, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_prep, dim3(grid), dim3(256), 0, s, arena, gtab);
    return hipGetLastError();
}

hipError_t launch_batch(uint8_t *arena, const MulTab *tw, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(k_batch, dim3(grid), dim3(kBatchThreads), 0, s, arena, tw);
    return hipGetLastError();
}

int batch_blocks_per_cu() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch, kBatchThreads, 0) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace ars
