// ars_batch.hip -- many codewords of uneven sizes in one persistent launch.
//
// Device side.  The host flattens a batch into work items (see BItem): each
// item is one tile of one code shape whose lanes carry column quads of one or
// more codewords (BSeg), so a large codeword spans many items and many small
// codewords share one.  A fixed grid of workgroups (sized to the occupancy)
// pulls items from an atomic counter until the queue is empty.  Twiddles are
// indexed by absolute point index, so one LDS copy serves every code shape.
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
    if (threadIdx.x == 0) reinterpret_cast<BHeader *>(arena)->counter = 0;
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

template <int TL, class IO>
__device__ __noinline__ void run_body(const BItem &it, IO &io, int q, int p, uint32_t *tile,
                                         const MulTab *tw, const MulTab *s_tab, const uint8_t *s_er,
                                         uint32_t si) {
    const uint32_t kind = it.meta & 0xffu;
    const bool acc = (it.meta >> 16) & 1u;
    constexpr uint32_t PQ = batch_lanes(TL);
    if constexpr (TL <= 4) {
        if (kind == kEncRef) {
            if (acc) body_encode_small<TL, true>(io, it.arg, tw);
            else body_encode_small<TL, false>(io, it.arg, tw);
        } else if (kind == kEncKpow2) {
            body_encode_kpow2_small<TL>(io, it.arg & 0xffffu, it.arg >> 16, tw);
        } else if constexpr (TL >= 1) {
            LdsTabs tabs{s_tab + si * (1u << TL), s_er + si * (1u << TL)};
            body_recover_small<TL>(io, tabs, tw);
        }
    } else {
        if (kind == kEncRef) {
            if constexpr (TL <= 7) {
                if (acc) body_encode_tile<TL, PQ, true, 1>(io, q, p, tile, it.arg, tw);
                else body_encode_tile<TL, PQ, false, 1>(io, q, p, tile, it.arg, tw);
            }
        } else if (kind == kEncKpow2) {
            if constexpr (TL <= 7) body_encode_kpow2_tile<TL, PQ, 1>(io, q, p, tile, it.arg & 0xffffu, it.arg >> 16, tw);
        } else {
            LdsTabs tabs{s_tab + si * (1u << TL), s_er + si * (1u << TL)};
            body_recover_tile<TL, PQ, 1>(io, tabs, q, p, tile, tw);
        }
    }
}

template <int TL>
__device__ __forceinline__ void run_tile(const BItem &it, const BSeg *segs, const BJob *jobs,
                                         uint32_t r, uint32_t *tile, const MulTab *tw,
                                         const MulTab *s_tab, const uint8_t *s_er) {
    constexpr uint32_t PQ = batch_lanes(TL);
    const int p = (int)(threadIdx.x % PQ);
    const int q = (int)(threadIdx.x / PQ);
    if ((it.nseg_rep & 0xffffu) == 1u) {
        // one codeword: scalar job descriptor, buffer-resource addressing
        const BSeg sg = segs[it.seg0];
        const BJob job = jobs[sg.job];
        const uint32_t quad = sg.quad0 + (uint32_t)p + r * PQ;
        BufIO<true> io = make_bufio<true>(job.in, job.in_bytes, job.out, job.out_bytes,
                                          quad < job.quads ? quad : job.quads, job.L);
        run_body<TL>(it, io, q, p, tile, tw, s_tab, s_er, 0u);
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
        run_body<TL>(it, io, q, p, tile, tw, s_tab, s_er, si);
    }
}

template <int TL>
__device__ __forceinline__ void run_item(const BItem &it, const BSeg *segs, const BJob *jobs,
                                         uint32_t *tile, const MulTab *tw, MulTab *s_tab, uint8_t *s_er) {
    const uint32_t rep = it.nseg_rep >> 16;
    if ((it.meta & 0xffu) == kRecover) {
        stage_item_plans<TL>(it, segs, jobs, s_tab, s_er);
        __syncthreads();
    }
    for (uint32_t r = 0; r < rep; ++r) {
        if (r) __syncthreads();  // the previous tile's LDS reads are done
        run_tile<TL>(it, segs, jobs, r, tile, tw, s_tab, s_er);
    }
}

__global__ void __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) k_batch(uint8_t *__restrict__ arena,
                                                         const MulTab *__restrict__ g_tw) {
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
    const uint32_t nitems = h->nitems;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&h->counter, 1u);
        __syncthreads();
        const uint32_t i = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (i >= nitems) break;
        const BItem it = items[i];
        switch ((it.meta >> 8) & 0xffu) {
            case 0: run_item<0>(it, segs, jobs, tile, tw, s_tab, s_er); break;
            case 1: run_item<1>(it, segs, jobs, tile, tw, s_tab, s_er); break;
            case 2: run_item<2>(it, segs, jobs, tile, tw, s_tab, s_er); break;
            case 3: run_item<3>(it, segs, jobs, tile, tw, s_tab, s_er); break;
            case 4: run_item<4>(it, segs, jobs, tile, tw, s_tab, s_er); break;
            case 5: run_item<5>(it, segs, jobs, tile, tw, s_tab, s_er); break;
            case 6: run_item<6>(it, segs, jobs, tile, tw, s_tab, s_er); break;
            case 7: run_item<7>(it, segs, jobs, tile, tw, s_tab, s_er); break;
            case 8: run_item<8>(it, segs, jobs, tile, tw, s_tab, s_er); break;
            default: break;
        }
    }
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
