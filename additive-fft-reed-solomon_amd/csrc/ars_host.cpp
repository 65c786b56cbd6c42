// ars_host.cpp -- host-slice entry points: a streaming pipeline over column
// chunks.  Byte columns are independent, so a codeword in host memory is
// processed as chunks of w columns: gather the chunk's rows into pinned
// staging (pitch w), copy in, run the fused kernel with shard_len = w, copy
// the output rows back, scatter them.  Three stream slots keep copy-in,
// kernel and copy-out of consecutive chunks overlapped; the path is bound by
// PCIe and host memcpy, not by HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/ars.h"
#include "ars_ctx.hpp"
#include "ars_kernels.hpp"

namespace {

constexpr int kSlots = 3;
constexpr size_t kChunkBytes = 16u << 20;  // device bytes per chunk (all n rows)

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint8_t *dev = nullptr;      // n * w bytes
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
    bool ready = false;
};

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
    hipError_t e;
    for (Slot &s : p->slots) {
        if ((e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming)) != hipSuccess ||
            (e = hipMalloc(&s.dev, kChunkBytes)) != hipSuccess ||
            (e = hipHostMalloc(&s.h_in, kChunkBytes, hipHostMallocDefault)) != hipSuccess ||
            (e = hipHostMalloc(&s.h_out, kChunkBytes, hipHostMallocDefault)) != hipSuccess)
            return ars::hip_error(e, "host pipeline set-up");
    }
    if ((e = hipMalloc(&p->plan, sizeof(ars::RecPrep))) != hipSuccess) return ars::hip_error(e, "host pipeline plan");
    (void)ctx;
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

// Copies rows [r0, r0 + rows) of width w at column c0 from a pitch-L buffer
// into a pitch-w buffer, or back.
void gather(uint8_t *dst, const uint8_t *src, size_t L, size_t c0, size_t w, uint32_t rows) {
    for (uint32_t r = 0; r < rows; ++r) std::memcpy(dst + (size_t)r * w, src + (size_t)r * L + c0, w);
}
void scatter(uint8_t *dst, const uint8_t *src, size_t L, size_t c0, size_t w, uint32_t rows) {
    for (uint32_t r = 0; r < rows; ++r) std::memcpy(dst + (size_t)r * L + c0, src + (size_t)r * w, w);
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
    delete p;
}

extern "C" {

int ars_encode_host(ars_ctx *ctx, uint32_t n, uint32_t k, const uint8_t *message, uint8_t *parity,
                    size_t L, uint32_t flags) {
    if (!ctx || !message || !parity) return ars::set_error(ARS_E_NULL, "null context or buffer");
    int rc = ars::check_code(n, k, L);
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
    const bool acc = flags & ARS_ENC_ACCUMULATE;
    const size_t W = chunk_width(n, L);
    int si = 0;
    for (size_t c0 = 0; c0 < L; c0 += W, si = (si + 1) % kSlots) {
        Slot &s = p->slots[si];
        const size_t w = std::min(W, L - c0);
        if (s.pending) {  // retire this slot's previous chunk
            if ((e = hipEventSynchronize(s.done)) != hipSuccess) return ars::hip_error(e, "host pipeline");
            scatter(parity, s.h_out, L, s.c0, s.w, T);
            s.pending = false;
        }
        uint8_t *d_par = s.dev, *d_msg = s.dev + (size_t)T * w;
        gather(s.h_in + (size_t)T * w, message, L, c0, w, k);
        if (acc) gather(s.h_in, parity, L, c0, w, T);
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
        scatter(parity, s.h_out, L, s.c0, s.w, T);
        s.pending = false;
    }
    return ARS_OK;
}

int ars_recover_host(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *received, size_t L,
                     const uint8_t *positions, uint32_t e_cnt) {
    if (!ctx || !received) return ars::set_error(ARS_E_NULL, "null context or buffer");
    int rc = ars::check_code(n, k, L);
    if (rc) return rc;
    if (e_cnt > n - k) return ARS_FALSE;
    if (e_cnt == 0) return ARS_TRUE;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return ars::hip_error(e, "hipSetDevice");
    Pipe *p = pipe_for(ctx);
    if (!p) return ars::set_error(ARS_E_NOMEM, "host pipeline allocation failed");
    std::lock_guard<std::mutex> lk(p->mu);
    if ((rc = init_pipe(ctx, p))) return rc;
    // one erasure plan for every chunk (validates the positions)
    if ((rc = ars_recover_plan(ctx, n, k, positions, e_cnt, p->plan, p->slots[0].stream)) < 0) return rc;
    if ((e = hipStreamSynchronize(p->slots[0].stream)) != hipSuccess) return ars::hip_error(e, "recover plan");
    std::vector<uint8_t> er(positions, positions + e_cnt);
    const size_t W = chunk_width(n, L);
    auto retire = [&](Slot &s) -> int {
        hipError_t ee = hipEventSynchronize(s.done);
        if (ee != hipSuccess) return ars::hip_error(ee, "host pipeline");
        for (uint8_t r : er) std::memcpy(received + (size_t)r * L + s.c0, s.h_out + (size_t)r * s.w, s.w);
        s.pending = false;
        return ARS_OK;
    };
    int si = 0;
    for (size_t c0 = 0; c0 < L; c0 += W, si = (si + 1) % kSlots) {
        Slot &s = p->slots[si];
        const size_t w = std::min(W, L - c0);
        if (s.pending && (rc = retire(s))) return rc;
        gather(s.h_in, received, L, c0, w, n);
        if ((e = hipMemcpyAsync(s.dev, s.h_in, (size_t)n * w, hipMemcpyHostToDevice, s.stream)) != hipSuccess)
            return ars::hip_error(e, "host pipeline copy-in");
        if ((rc = ars_recover_with_plan(ctx, n, k, s.dev, w, p->plan, s.stream)) < 0) return rc;
        for (uint8_t r : er)
            if ((e = hipMemcpyAsync(s.h_out + (size_t)r * w, s.dev + (size_t)r * w, w, hipMemcpyDeviceToHost,
                                    s.stream)) != hipSuccess)
                return ars::hip_error(e, "host pipeline copy-out");
        if ((e = hipEventRecord(s.done, s.stream)) != hipSuccess) return ars::hip_error(e, "hipEventRecord");
        s.pending = true;
        s.c0 = c0;
        s.w = w;
    }
    for (Slot &s : p->slots)
        if (s.pending && (rc = retire(s))) return rc;
    return ARS_TRUE;
}

}  // extern "C"
