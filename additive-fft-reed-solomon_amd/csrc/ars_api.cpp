// ars_api.cpp -- the C ABI (include/ars.h): argument validation, resident
// tables, and kernel dispatch.  No CPU fallback exists: every entry point
// runs the HIP kernels or returns an error.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "../../include/ars.h"
#include "ars_ctx.hpp"
#include "ars_gf.hpp"
#include "ars_kernels.hpp"

namespace ars {

thread_local std::string g_err;

int set_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int hip_error(hipError_t e, const char *where) {
    g_err = std::string(where) + ": " + hipGetErrorString(e);
    return ARS_E_HIP;
}

bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }
int ilog2(uint32_t v) {
    int r = 0;
    while ((1u << r) < v) ++r;
    return r;
}

}  // namespace ars

namespace {

using ars::g_err;
using ars::ilog2;
using ars::is_pow2;

int fail(int code, const char *msg) { return ars::set_error(code, msg); }
int hip_fail(hipError_t e, const char *where) { return ars::hip_error(e, where); }

}  // namespace

int ars::check_code(uint32_t n, uint32_t k, size_t L, bool wide) {
    if (n < 2 || n > 256 || !is_pow2(n)) return fail(ARS_E_INVAL, "n must be a power of two in [2, 256]");
    if (k < 1 || k >= n) return fail(ARS_E_INVAL, "k must satisfy 1 <= k < n");
    if (L == 0) return fail(ARS_E_INVAL, "shard_len must be >= 1");
    if (L > 0xFFFFFFFFull) return fail(ARS_E_INVAL, "shard_len must be below 4 GiB");
    if (!wide && (uint64_t)n * L > 0xFFFFFFFFull)
        return fail(ARS_E_INVAL, "codeword of 4 GiB or more: this call needs n * shard_len below 4 GiB");
    const uint32_t T = n - k;
    if (!is_pow2(T) && !is_pow2(k))
        return fail(ARS_E_UNSUPPORTED, "neither T = n - k nor k is a power of two");
    return ARS_OK;
}

namespace {
using ars::check_code;

// Access mode of a call on buffers a and b (ars::io_mode).  Codewords of 4 GiB
// or more run only as aligned dwords (mode 3).
int call_mode(const void *a, const void *b, uint32_t n, size_t L, int &am) {
    am = ars::io_mode(a, b, L, (uint64_t)n * L);
    if (am < 0)
        return fail(ARS_E_UNSUPPORTED,
                    "a codeword of 4 GiB or more needs 4-byte aligned buffers and shard_len % 4 == 0");
    return ARS_OK;
}

// Scratch of ars_recover_noalloc: room for one erasure plan (n <= 256).  Every
// GPU the library runs on holds its own copy until the process ends.  Host
// code must not use this name as a device address: each context takes its
// device's address with hipGetSymbolAddress in ars_create and passes that.
constexpr size_t kNoAllocScratchBytes = 6656;
static_assert(sizeof(ars::RecPrep) <= kNoAllocScratchBytes, "one plan must fit the scratch");
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// The plan buffer ars_recover keeps for `s`, created on the stream's first
// call.  Work on one stream runs in issue order, so a call's plan kernel
// overwrites the buffer only after the previous call's tile kernel has read
// it; a stream handle is reused only after the stream it named was destroyed,
// which HIP does once that stream's work has completed.  Keeping the buffer
// saves the stream-ordered free of a per-call allocation, whose completion
// marker leaves the GPU idle for about 6 us before the next kernel.  Returns
// null (the caller then allocates on the stream) for hipStreamPerThread, one
// handle that names a different stream on each host thread, for a stream
// under graph capture, whose graph may be replayed on any stream, and once
// the context holds kMaxPlanSlots buffers.
ars::PlanSlot *plan_slot(ars_ctx *ctx, hipStream_t s) {
    if (s == hipStreamPerThread) return nullptr;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
    std::lock_guard<std::mutex> lk(ctx->plan_mu);
    for (auto &p : ctx->plan_slots)
        if (p->stream == s) return p.get();
    if (ctx->plan_slots.size() >= ars::kMaxPlanSlots) return nullptr;
    std::unique_ptr<ars::PlanSlot> slot(new (std::nothrow) ars::PlanSlot);
    if (!slot || hipMalloc(reinterpret_cast<void **>(&slot->buf), sizeof(ars::RecPrep)) != hipSuccess) return nullptr;
    slot->stream = s;
    ctx->plan_slots.push_back(std::move(slot));
    return ctx->plan_slots.back().get();
}
}  // namespace

extern "C" {

const char *ars_last_error(void) { return g_err.c_str(); }
const char *ars_version(void) { return "ars-mi355x 0.1 (gfx950)"; }

int ars_create(int device, ars_ctx **out) {
    if (!out) return fail(ARS_E_NULL, "out is null");
    *out = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    ars_ctx *c = new (std::nothrow) ars_ctx;
    if (!c) return fail(ARS_E_NOMEM, "context allocation failed");
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    const ars::GfTables &g = ars::gf();
    ars::MulTab tw[ars::kTwiddleEntries];
    for (int l = 0; l < 8; ++l)
        for (int j = 0; j < (128 >> l); ++j)
            tw[ars::lvl_off(l) + j] = g.multab(g.spv[l][g.sub[j << (l + 1)]]);
    uint8_t gtab[ars::kGtabBytes];
    std::memcpy(gtab, g.exp, 512);
    std::memcpy(gtab + 512, g.log, 256);
    std::memcpy(gtab + 768, g.sub, 256);
    if ((e = hipMalloc(&c->d_tw, sizeof tw)) != hipSuccess ||
        (e = hipMalloc(&c->d_gtab, sizeof gtab)) != hipSuccess ||
        (e = hipMemcpy(c->d_tw, tw, sizeof tw, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c->d_gtab, gtab, sizeof gtab, hipMemcpyHostToDevice)) != hipSuccess) {
        hipFree(c->d_tw);
        hipFree(c->d_gtab);
        delete c;
        return hip_fail(e, "context tables");
    }
    // this device's copy of the no-allocation scratch (the current device is `device`)
    if ((e = hipGetSymbolAddress(reinterpret_cast<void **>(&c->noalloc_plan), HIP_SYMBOL(scratch))) != hipSuccess) {
        hipFree(c->d_tw);
        hipFree(c->d_gtab);
        delete c;
        return hip_fail(e, "hipGetSymbolAddress(scratch)");
    }
    *out = c;
    return ARS_OK;
}

int ars_destroy(ars_ctx *ctx) {
    if (!ctx) return ARS_OK;
    hipSetDevice(ctx->device);
    ars::release_host_pipe(ctx);
    ars::release_decode_state(ctx);
    hipFree(ctx->d_tw);
    hipFree(ctx->d_gtab);
    for (auto &p : ctx->plan_slots) hipFree(p->buf);
    delete ctx;
    return ARS_OK;
}

int ars_encode(ars_ctx *ctx, uint32_t n, uint32_t k, const uint8_t *d_message, uint8_t *d_parity,
               size_t L, uint32_t flags, void *stream) {
    if (!ctx || !d_message || !d_parity) return fail(ARS_E_NULL, "null context or buffer");
    int rc = check_code(n, k, L, true);
    if (rc) return rc;
    if (flags & ~ARS_ENC_ACCUMULATE) return fail(ARS_E_INVAL, "unknown encode flags");
    int am;
    if ((rc = call_mode(d_message, d_parity, n, L, am))) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t T = n - k;
    if (is_pow2(T)) {
        const int tl = ilog2(T);
        e = ars::launch_encode_ref(tl, am, (flags & ARS_ENC_ACCUMULATE) != 0, d_message, d_parity,
                                   L, k / T, ctx->d_tw, s, ctx->tune_pq[0][tl], ctx->tune_nbuf[0][tl]);
    } else {
        if (flags & ARS_ENC_ACCUMULATE)
            return fail(ARS_E_UNSUPPORTED, "ACCUMULATE is defined only when T is a power of two");
        e = ars::launch_encode_kpow2(ilog2(k), am, d_message, d_parity, L, n / k - 1, n - k,
                                     ctx->d_tw, s);
    }
    if (e != hipSuccess) return hip_fail(e, "encode launch");
    return ARS_OK;
}

namespace {
int validate_positions(uint32_t n, const uint8_t *positions, uint32_t e_cnt, ars::PosArg &pa) {
    if (!positions) return fail(ARS_E_NULL, "positions is null");
    bool seen[256] = {false};
    for (uint32_t i = 0; i < e_cnt; ++i) {
        const uint8_t p = positions[i];
        if (p >= n) return fail(ARS_E_POSITION, "erasure position >= n");
        if (seen[p]) return fail(ARS_E_DUPLICATE, "duplicate erasure position");
        seen[p] = true;
        pa.p[i] = p;
    }
    return ARS_OK;
}
}  // namespace

size_t ars_recover_plan_size(void) { return sizeof(ars::RecPrep); }

int ars_recover_plan(ars_ctx *ctx, uint32_t n, uint32_t k, const uint8_t *positions, uint32_t e_cnt,
                     void *d_plan, void *stream) {
    if (!ctx || !d_plan) return fail(ARS_E_NULL, "null context or plan");
    int rc = check_code(n, k, 1);
    if (rc) return rc;
    if (e_cnt > n - k) return ARS_FALSE;
    ars::PosArg pa;
    if (e_cnt && (rc = validate_positions(n, positions, e_cnt, pa))) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    e = ars::launch_recover_prep(static_cast<ars::RecPrep *>(d_plan), n, e_cnt, pa, ctx->d_gtab,
                                 static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "recover plan launch");
    return ARS_TRUE;
}

int ars_recover_with_plan(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received, size_t L,
                          const void *d_plan, void *stream) {
    if (!ctx || !d_received || !d_plan) return fail(ARS_E_NULL, "null context, buffer or plan");
    int rc = check_code(n, k, L, true);
    if (rc) return rc;
    int am;
    if ((rc = call_mode(d_received, d_received, n, L, am))) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int nl = ilog2(n);
    e = ars::launch_recover(nl, am, d_received, L, static_cast<const ars::RecPrep *>(d_plan),
                            ctx->d_tw, static_cast<hipStream_t>(stream), ctx->tune_pq[1][nl], ctx->tune_nbuf[1][nl],
                            (uint32_t)ctx->num_cus);
    if (e != hipSuccess) return hip_fail(e, "recover launch");
    return ARS_TRUE;
}

int ars_recover(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received, size_t L,
                const uint8_t *positions, uint32_t e_cnt, void *stream) {
    if (!ctx || !d_received) return fail(ARS_E_NULL, "null context or buffer");
    int rc = check_code(n, k, L, true);
    if (rc) return rc;
    int am;
    if ((rc = call_mode(d_received, d_received, n, L, am))) return rc;
    const uint32_t T = n - k;
    // rs/src/codec.rs:491-496: too many erasures -> false, none -> true.
    if (e_cnt > T) return ARS_FALSE;
    if (e_cnt == 0) return ARS_TRUE;
    ars::PosArg pa;
    if ((rc = validate_positions(n, positions, e_cnt, pa))) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nl = ilog2(n);
    // Launches of at most one workgroup per compute unit compute the plan in
    // the recover kernel itself: one launch instead of two.
    e = ars::launch_recover_fused(nl, am, d_received, L, pa, e_cnt, ctx->d_gtab, ctx->d_tw, s,
                                  (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256));
    if (e == hipSuccess) return ARS_TRUE;
    if (e != hipErrorNotSupported) return hip_fail(e, "recover launch");
    auto enqueue = [&](ars::RecPrep *prep) {
        hipError_t le = ars::launch_recover_prep(prep, n, e_cnt, pa, ctx->d_gtab, s);
        if (le == hipSuccess)
            le = ars::launch_recover(nl, am, d_received, L, prep, ctx->d_tw, s, ctx->tune_pq[1][nl],
                                     ctx->tune_nbuf[1][nl], (uint32_t)ctx->num_cus);
        return le;
    };
    if (ars::PlanSlot *slot = plan_slot(ctx, s)) {
        std::lock_guard<std::mutex> lk(slot->mu);
        if ((e = enqueue(slot->buf)) != hipSuccess) return hip_fail(e, "recover launch");
        return ARS_TRUE;
    }
    ars::RecPrep *prep = nullptr;
    if ((e = hipMallocAsync(reinterpret_cast<void **>(&prep), sizeof(ars::RecPrep), s)) != hipSuccess)
        return hip_fail(e, "hipMallocAsync");
    e = enqueue(prep);
    hipError_t ef = hipFreeAsync(prep, s);
    if (e != hipSuccess) return hip_fail(e, "recover launch");
    if (ef != hipSuccess) return hip_fail(ef, "hipFreeAsync");
    return ARS_TRUE;
}

int ars_recover_noalloc(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received, size_t L,
                        const uint8_t *positions, uint32_t e_cnt, void *stream) {
    if (!ctx || !d_received) return fail(ARS_E_NULL, "null context or buffer");
    int rc = check_code(n, k, L, true);  // n <= 256: the plan fits the scratch (static_assert above)
    if (rc) return rc;
    int am;
    if ((rc = call_mode(d_received, d_received, n, L, am))) return rc;
    const uint32_t T = n - k;
    if (e_cnt > T) return ARS_FALSE;
    if (e_cnt == 0) return ARS_TRUE;
    ars::PosArg pa;
    if ((rc = validate_positions(n, positions, e_cnt, pa))) return rc;
    if (!ctx->noalloc_plan) return fail(ARS_E_STATE, "context has no scratch address");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nl = ilog2(n);
    e = ars::launch_recover_fused(nl, am, d_received, L, pa, e_cnt, ctx->d_gtab, ctx->d_tw, s,
                                  (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256));
    if (e == hipSuccess) return ARS_TRUE;
    if (e != hipErrorNotSupported) return hip_fail(e, "recover launch");
    if ((e = ars::launch_recover_prep(ctx->noalloc_plan, n, e_cnt, pa, ctx->d_gtab, s)) != hipSuccess ||
        (e = ars::launch_recover(nl, am, d_received, L, ctx->noalloc_plan, ctx->d_tw, s, ctx->tune_pq[1][nl],
                                 ctx->tune_nbuf[1][nl], (uint32_t)ctx->num_cus)) != hipSuccess)
        return hip_fail(e, "recover launch");
    return ARS_TRUE;
}

int ars_recompute_data(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received, size_t L,
                       void *stream) {
    if (!ctx || !d_received) return fail(ARS_E_NULL, "null context or buffer");
    int rc = check_code(n, k, L, true);
    if (rc) return rc;
    if (n != 2 * (n - k)) return fail(ARS_E_INVAL, "recompute_data requires n = 2T");
    int am;
    if ((rc = call_mode(d_received, d_received, n, L, am))) return rc;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // mode 2 runs as bytes here (launch_recompute)
    e = ars::launch_recompute(ilog2(n - k), am, d_received, L, ctx->d_tw, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "recompute launch");
    return ARS_OK;
}

}  // extern "C"
