// ars_host.cpp -- host-slice entry points: a streaming pipeline over column
// chunks.  Byte columns are independent, so a codeword in host memory is
// processed as chunks of w columns: gather the chunk's rows into pinned
// staging (pitch w), copy in, run the fused kernel with shard_len = w, copy
// the output rows back, scatter them.  Three stream slots keep copy-in,
// kernel and copy-out of consecutive chunks overlapped; the path is bound by
// PCIe and host memcpy, not by HBM.  Gather and scatter run on a small
// worker pool: one thread's memcpy between pageable and pinned memory is far
// below the link's rate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/ars.h"
#include "ars_work_pool.hpp"
#include "ars_ctx.hpp"
#include "ars_kernels.hpp"

namespace {

using ars::WorkPool;

constexpr int kSlots = 3;
constexpr size_t kChunkBytes = 16u << 20;  // device bytes per chunk (all n rows)

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint8_t *dev = nullptr;      // n * w bytes, then e * w bytes of packed erased rows
    uint8_t *h_in = nullptr;     // pinned, n * w bytes
    uint8_t *h_out = nullptr;    // pinned, n * w bytes
    // pending scatter of this slot's output to the caller
    bool pending = false;
    size_t c0 = 0, w = 0;
};

struct Pipe {
    std::mutex mu;
    Slot slots[kSlots];
    ars::RecPrep *plan = nullptr;
    WorkPool *pool = nullptr;
    bool ready = false;
};

// Worker threads besides the caller: enough to approach the link's rate
// without taking over the host.
unsigned copy_workers() {
    const unsigned hc = std::thread::hardware_concurrency();
    return std::min(7u, hc > 1 ? hc - 1 : 0u);
}

std::mutex g_pipes_mu;
std::vector<std::pair<ars_ctx *, Pipe *>> g_pipes;

Pipe *pipe_for(ars_ctx *ctx) {
    std::lock_guard<std::mutex> lk(g_pipes_mu);
    for (auto &p : g_pipes)
        if (p.first == ctx) return p.second;
    Pipe *p = new (std::nothrow) Pipe;
    if (p) g_pipes.emplace_back(ctx, p);
    return p;
}

int init_pipe(ars_ctx *ctx, Pipe *p) {
    if (p->ready) return ARS_OK;
    // After a failed set-up the next call lands here again: create only what
    // is still missing, so nothing made by the earlier attempt is leaked.
    hipError_t e = hipSuccess;
    for (Slot &s : p->slots) {
        if ((!s.stream && (e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)) != hipSuccess) ||
            (!s.done && (e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming)) != hipSuccess) ||
            (!s.dev && (e = hipMalloc(&s.dev, 2 * kChunkBytes)) != hipSuccess) ||
            (!s.h_in && (e = hipHostMalloc(&s.h_in, kChunkBytes, hipHostMallocDefault)) != hipSuccess) ||
            (!s.h_out && (e = hipHostMalloc(&s.h_out, kChunkBytes, hipHostMallocDefault)) != hipSuccess))
            return ars::hip_error(e, "host pipeline set-up");
    }
    if (!p->plan && (e = hipMalloc(&p->plan, sizeof(ars::RecPrep))) != hipSuccess)
        return ars::hip_error(e, "host pipeline plan");
    (void)ctx;
    if (!p->pool) p->pool = new (std::nothrow) WorkPool(copy_workers());
    if (!p->pool) return ars::set_error(ARS_E_NOMEM, "host pipeline workers");
    p->ready = true;
    return ARS_OK;
}

// Chunk width: a multiple of 64 bytes (dword path, whole cache lines) with
// all n rows of a chunk within kChunkBytes; the last chunk takes the rest.
size_t chunk_width(uint32_t n, size_t L) {
    size_t w = (kChunkBytes / n) & ~size_t(63);
    if (w == 0) w = 64;
    return std::min(w, L);
}

// Row list for the pack kernel, passed by value.
struct RowList {
    uint8_t r[256];
};

// dst row i (pitch w) = src row rows.r[i] (pitch w), i = blockIdx.y: packs the
// recovered rows so that one copy brings them back.  16-byte accesses when w
// allows, bytes otherwise (a ragged last chunk).
__global__ void __launch_bounds__(256) k_pack_rows(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst,
                                                   RowList rows, uint32_t w) {
    const uint8_t *a = src + (size_t)rows.r[blockIdx.y] * w;
    uint8_t *b = dst + (size_t)blockIdx.y * w;
    const uint32_t off = (blockIdx.x * 256u + threadIdx.x) * 16u;
    if (off >= w) return;
    if ((w & 15u) == 0) {
        *reinterpret_cast<uint4 *>(b + off) = *reinterpret_cast<const uint4 *>(a + off);
    } else {
        const uint32_t end = min(off + 16u, w);
        for (uint32_t i = off; i < end; ++i) b[i] = a[i];
    }
}

// Row copies between a pitch-L caller buffer (columns [c0, c0 + w)) and a
// pitch-w staging buffer, split into pieces of about 1 MiB over the pool.
void copy_rows(WorkPool *pool, uint8_t *pitched, uint8_t *staged, size_t L, size_t c0, size_t w, uint32_t count,
               bool to_staging) {
    const size_t per = std::max<size_t>(1, (size_t(1) << 20) / std::max<size_t>(w, 1));
    const size_t pieces = (count + per - 1) / per;
    pool->run(pieces, [&](size_t pi) {
        const size_t end = std::min<size_t>(count, (pi + 1) * per);
        for (size_t r = pi * per; r < end; ++r) {
            uint8_t *a = pitched + r * L + c0, *b = staged + r * w;
            if (to_staging)
                std::memcpy(b, a, w);
            else
                std::memcpy(a, b, w);
        }
    });
}
// Packed staging row i -> caller row rows[i].
void copy_packed(WorkPool *pool, uint8_t *dst, const uint8_t *packed, size_t L, size_t c0, size_t w,
                 const uint8_t *rows, uint32_t count) {
    const size_t per = std::max<size_t>(1, (size_t(1) << 20) / std::max<size_t>(w, 1));
    const size_t pieces = (count + per - 1) / per;
    pool->run(pieces, [&](size_t pi) {
        const size_t end = std::min<size_t>(count, (pi + 1) * per);
        for (size_t i = pi * per; i < end; ++i) std::memcpy(dst + (size_t)rows[i] * L + c0, packed + i * w, w);
    });
}
void gather(WorkPool *pool, uint8_t *dst, const uint8_t *src, size_t L, size_t c0, size_t w, uint32_t rows) {
    copy_rows(pool, const_cast<uint8_t *>(src), dst, L, c0, w, rows, true);
}
void scatter(WorkPool *pool, uint8_t *dst, const uint8_t *src, size_t L, size_t c0, size_t w, uint32_t rows) {
    copy_rows(pool, dst, const_cast<uint8_t *>(src), L, c0, w, rows, false);
}

// Waits for chunks a previous call left in flight (it returned an error
// midway) and forgets them: their output belonged to that call's buffers.
void discard_pending(Pipe *p) {
    for (Slot &s : p->slots) {
        if (!s.pending) continue;
        hipEventSynchronize(s.done);
        s.pending = false;
    }
}

// Runs a pipeline body; whatever it returns, no chunk stays pending.
template <class F>
int run_pipe(Pipe *p, F &&body) {
    discard_pending(p);
    const int rc = body();
    if (rc < 0) discard_pending(p);
    return rc;
}

}  // namespace

void ars::release_host_pipe(ars_ctx *ctx) {
    Pipe *p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_pipes_mu);
        for (auto it = g_pipes.begin(); it != g_pipes.end(); ++it)
            if (it->first == ctx) {
                p = it->second;
                g_pipes.erase(it);
                break;
            }
    }
    if (!p) return;
    for (Slot &s : p->slots) {
        if (s.stream) hipStreamSynchronize(s.stream);
        if (s.done) hipEventDestroy(s.done);
        if (s.stream) hipStreamDestroy(s.stream);
        hipFree(s.dev);
        hipHostFree(s.h_in);
        hipHostFree(s.h_out);
    }
    hipFree(p->plan);
    delete p->pool;
    delete p;
}

extern "C" {

int ars_encode_host(ars_ctx *ctx, uint32_t n, uint32_t k, const uint8_t *message, uint8_t *parity,
                    size_t L, uint32_t flags) {
    if (!ctx || !message || !parity) return ars::set_error(ARS_E_NULL, "null context or buffer");
    // any codeword size: the device sees column chunks of at most kChunkBytes
    int rc = ars::check_code(n, k, L, true);
    if (rc) return rc;
    if (flags & ~ARS_ENC_ACCUMULATE) return ars::set_error(ARS_E_INVAL, "unknown encode flags");
    const uint32_t T = n - k;
    if ((flags & ARS_ENC_ACCUMULATE) && !ars::is_pow2(T))
        return ars::set_error(ARS_E_UNSUPPORTED, "ACCUMULATE is defined only when T is a power of two");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return ars::hip_error(e, "hipSetDevice");
    Pipe *p = pipe_for(ctx);
    if (!p) return ars::set_error(ARS_E_NOMEM, "host pipeline allocation failed");
    std::lock_guard<std::mutex> lk(p->mu);
    if ((rc = init_pipe(ctx, p))) return rc;
    return run_pipe(p, [&]() -> int {
        const bool acc = flags & ARS_ENC_ACCUMULATE;
        const size_t W = chunk_width(n, L);
        int si = 0;
        for (size_t c0 = 0; c0 < L; c0 += W, si = (si + 1) % kSlots) {
            Slot &s = p->slots[si];
            const size_t w = std::min(W, L - c0);
            if (s.pending) {  // retire this slot's previous chunk
                if ((e = hipEventSynchronize(s.done)) != hipSuccess) return ars::hip_error(e, "host pipeline");
                scatter(p->pool, parity, s.h_out, L, s.c0, s.w, T);
                s.pending = false;
            }
            uint8_t *d_par = s.dev, *d_msg = s.dev + (size_t)T * w;
            gather(p->pool, s.h_in + (size_t)T * w, message, L, c0, w, k);
            if (acc) gather(p->pool, s.h_in, parity, L, c0, w, T);
            const size_t in_off = acc ? 0 : (size_t)T * w;
            if ((e = hipMemcpyAsync(s.dev + in_off, s.h_in + in_off, (size_t)n * w - in_off, hipMemcpyHostToDevice,
                                    s.stream)) != hipSuccess)
                return ars::hip_error(e, "host pipeline copy-in");
            if ((rc = ars_encode(ctx, n, k, d_msg, d_par, w, flags, s.stream)) < 0) return rc;
            if ((e = hipMemcpyAsync(s.h_out, d_par, (size_t)T * w, hipMemcpyDeviceToHost, s.stream)) != hipSuccess ||
                (e = hipEventRecord(s.done, s.stream)) != hipSuccess)
                return ars::hip_error(e, "host pipeline copy-out");
            s.pending = true;
            s.c0 = c0;
            s.w = w;
        }
        for (Slot &s : p->slots) {
            if (!s.pending) continue;
            if ((e = hipEventSynchronize(s.done)) != hipSuccess) return ars::hip_error(e, "host pipeline");
            scatter(p->pool, parity, s.h_out, L, s.c0, s.w, T);
            s.pending = false;
        }
        return ARS_OK;
    });
}

int ars_recover_host(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *received, size_t L,
                     const uint8_t *positions, uint32_t e_cnt) {
    if (!ctx || !received) return ars::set_error(ARS_E_NULL, "null context or buffer");
    int rc = ars::check_code(n, k, L, true);  // column chunks, as in ars_encode_host
    if (rc) return rc;
    if (e_cnt > n - k) return ARS_FALSE;
    if (e_cnt == 0) return ARS_TRUE;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return ars::hip_error(e, "hipSetDevice");
    Pipe *p = pipe_for(ctx);
    if (!p) return ars::set_error(ARS_E_NOMEM, "host pipeline allocation failed");
    std::lock_guard<std::mutex> lk(p->mu);
    if ((rc = init_pipe(ctx, p))) return rc;
    return run_pipe(p, [&]() -> int {
        // one erasure plan for every chunk (validates the positions)
        if ((rc = ars_recover_plan(ctx, n, k, positions, e_cnt, p->plan, p->slots[0].stream)) < 0) return rc;
        if ((e = hipStreamSynchronize(p->slots[0].stream)) != hipSuccess) return ars::hip_error(e, "recover plan");
        std::vector<uint8_t> er(positions, positions + e_cnt);
        RowList rows{};
        std::memcpy(rows.r, positions, e_cnt);
        const size_t W = chunk_width(n, L);
        auto retire = [&](Slot &s) -> int {
            hipError_t ee = hipEventSynchronize(s.done);
            if (ee != hipSuccess) return ars::hip_error(ee, "host pipeline");
            copy_packed(p->pool, received, s.h_out, L, s.c0, s.w, er.data(), e_cnt);
            s.pending = false;
            return ARS_OK;
        };
        int si = 0;
        for (size_t c0 = 0; c0 < L; c0 += W, si = (si + 1) % kSlots) {
            Slot &s = p->slots[si];
            const size_t w = std::min(W, L - c0);
            if (s.pending && (rc = retire(s))) return rc;
            gather(p->pool, s.h_in, received, L, c0, w, n);
            if ((e = hipMemcpyAsync(s.dev, s.h_in, (size_t)n * w, hipMemcpyHostToDevice, s.stream)) != hipSuccess)
                return ars::hip_error(e, "host pipeline copy-in");
            if ((rc = ars_recover_with_plan(ctx, n, k, s.dev, w, p->plan, s.stream)) < 0) return rc;
            uint8_t *packed = s.dev + (size_t)n * w;
            k_pack_rows<<<dim3((unsigned)((w + 4095) / 4096), e_cnt), 256, 0, s.stream>>>(s.dev, packed, rows,
                                                                                        (uint32_t)w);
            if ((e = hipGetLastError()) != hipSuccess) return ars::hip_error(e, "pack launch");
            if ((e = hipMemcpyAsync(s.h_out, packed, (size_t)e_cnt * w, hipMemcpyDeviceToHost, s.stream)) != hipSuccess)
                return ars::hip_error(e, "host pipeline copy-out");
            if ((e = hipEventRecord(s.done, s.stream)) != hipSuccess) return ars::hip_error(e, "hipEventRecord");
            s.pending = true;
            s.c0 = c0;
            s.w = w;
        }
        for (Slot &s : p->slots)
            if (s.pending && (rc = retire(s))) return rc;
        return ARS_TRUE;
    });
}

}  // extern "C"
