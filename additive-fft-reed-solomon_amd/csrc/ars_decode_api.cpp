// ars_decode_api.cpp -- ars_decode: sharded error decoding (see ars.h).
//
// 1. k_syndrome lists the columns whose syndrome is not zero (device).
// 2. k_decode_columns solves each listed column's key equation (the step the
//    reference runs as HGCD) and runs its root search and error evaluation,
//    one wave per column, reading the list length on the device; the host
//    waits once, for the counts.
// With ARS_DEC_HOST_KEY_EQUATION the key equations run on a host worker pool
// instead (ars_keyeq.hpp) and k_decode_tail does the rest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/ars.h"
#include "ars_ctx.hpp"
#include "ars_decode.hpp"
#include "ars_keyeq.hpp"
#include "ars_work_pool.hpp"

namespace {

ars::WorkPool &solver_pool() {
    static ars::WorkPool pool([] {
        const unsigned hc = std::thread::hardware_concurrency();
        return std::min(15u, hc > 1 ? hc - 1 : 0u);
    }());
    return pool;
}
std::mutex g_solver_mu;  // one job at a time on the pool

bool aligned4(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 3u) == 0; }

// Per-context decode scratch, kept between calls and grown on demand:
// device [listed, failed] counters | column list | syndromes, the device
// plans, and a pinned copy of the counters.  Calls on one context are
// serialised by `mu`.
struct DecState {
    std::mutex mu;
    uint8_t *dev = nullptr;
    size_t dev_bytes = 0;
    ars::DecPlan *plans = nullptr;
    size_t plan_count = 0;
    uint32_t *h_counts = nullptr;
    uint8_t *xt = nullptr;  // monomial coefficients of X_j, 256 x 256, log domain
};

std::mutex g_states_mu;
std::vector<std::pair<ars_ctx *, DecState *>> g_states;

DecState *state_for(ars_ctx *ctx) {
    std::lock_guard<std::mutex> lk(g_states_mu);
    for (auto &p : g_states)
        if (p.first == ctx) return p.second;
    DecState *st = new (std::nothrow) DecState;
    if (st) g_states.emplace_back(ctx, st);
    return st;
}

}  // namespace

void ars::release_decode_state(ars_ctx *ctx) {
    DecState *st = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_states_mu);
        for (auto it = g_states.begin(); it != g_states.end(); ++it)
            if (it->first == ctx) {
                st = it->second;
                g_states.erase(it);
                break;
            }
    }
    if (!st) return;
    hipFree(st->dev);
    hipFree(st->plans);
    hipHostFree(st->h_counts);
    hipFree(st->xt);
    delete st;
}

extern "C" int ars_decode(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received, size_t L, uint32_t flags,
                          uint8_t *d_status, uint64_t *counts, void *stream) {
    if (!ctx || !d_received) return ars::set_error(ARS_E_NULL, "null context or buffer");
    int rc = ars::check_code(n, k, L);
    if (rc) return rc;
    if (flags & ~(ARS_DEC_FORMAL_DERIVATIVE | ARS_DEC_HOST_KEY_EQUATION))
        return ars::set_error(ARS_E_INVAL, "unknown decode flags");
    const uint32_t T = n - k;
    if (!ars::is_pow2(T) || T < 2)
        return ars::set_error(ARS_E_UNSUPPORTED, "error decoding needs T = n - k a power of two >= 2");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return ars::hip_error(e, "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int tl = ars::ilog2(T);

    DecState *st = state_for(ctx);
    if (!st) return ars::set_error(ARS_E_NOMEM, "decode state allocation failed");
    std::lock_guard<std::mutex> lk(st->mu);
    // scratch: [listed, failed] | cols[L] | syn[L * T]
    const size_t off_cols = 256, off_syn = off_cols + ((4 * L + 255) & ~size_t(255));
    const size_t need = off_syn + L * T;
    if (st->dev_bytes < need) {
        hipFree(st->dev);
        st->dev = nullptr;
        st->dev_bytes = 0;
        if ((e = hipMalloc(reinterpret_cast<void **>(&st->dev), need)) != hipSuccess)
            return ars::hip_error(e, "decode scratch");
        st->dev_bytes = need;
    }
    if (!st->h_counts &&
        (e = hipHostMalloc(reinterpret_cast<void **>(&st->h_counts), 8, hipHostMallocDefault)) != hipSuccess)
        return ars::hip_error(e, "decode counters");
    uint8_t *base = st->dev;
    uint32_t *d_count = reinterpret_cast<uint32_t *>(base), *d_failed = d_count + 1;
    uint32_t *d_cols = reinterpret_cast<uint32_t *>(base + off_cols);
    uint8_t *d_syn = base + off_syn;
    if ((e = hipMemsetAsync(base, 0, 8, s)) != hipSuccess) return ars::hip_error(e, "decode reset");
    if (d_status && (e = hipMemsetAsync(d_status, 0, L, s)) != hipSuccess) return ars::hip_error(e, "status reset");

    const bool al = (L % 4 == 0) && aligned4(d_received);
    const bool formal = flags & ARS_DEC_FORMAL_DERIVATIVE;
    ars::SynOut so{d_count, d_cols, d_syn};
    if ((e = ars::launch_syndrome(tl, al, d_received, (uint32_t)L, n / T, ctx->d_tw, so, s)) != hipSuccess)
        return ars::hip_error(e, "syndrome launch");
    uint32_t listed = 0, failed = 0;
    if (!(flags & ARS_DEC_HOST_KEY_EQUATION)) {
        if (!st->xt) {
            // monomial coefficients of X_j in the log domain, 255 for zero
            std::vector<uint8_t> xl(256 * 256);
            const ars::KeyTables &kt = ars::key_tables();
            const ars::GfTables &g = ars::gf();
            for (int j = 0; j < 256; ++j)
                for (int i = 0; i < 256; ++i) xl[j * 256 + i] = kt.x[j][i] ? g.log[kt.x[j][i]] : (uint8_t)255;
            if ((e = hipMalloc(reinterpret_cast<void **>(&st->xt), 256 * 256)) != hipSuccess)
                return ars::hip_error(e, "decode basis table");
            if ((e = hipMemcpy(st->xt, xl.data(), 256 * 256, hipMemcpyHostToDevice)) != hipSuccess) {
                hipFree(st->xt);
                st->xt = nullptr;
                return ars::hip_error(e, "decode basis table copy");
            }
        }
        // enough waves to fill the device; each strides over the list
        const uint32_t grid = (uint32_t)std::min<size_t>(L, (size_t)std::max(ctx->num_cus, 1) * 32);
        if ((e = ars::launch_decode_columns(d_received, (uint32_t)L, n, tl, d_count, d_cols, d_syn, formal, grid,
                                            d_status, d_failed, ctx->d_gtab, ctx->d_tw, st->xt, s)) != hipSuccess)
            return ars::hip_error(e, "decode launch");
        if ((e = hipMemcpyAsync(st->h_counts, d_count, 8, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return ars::hip_error(e, "decode result");
        listed = st->h_counts[0];
        failed = st->h_counts[1];
        if (counts) {
            counts[0] = L - listed;
            counts[1] = listed - failed;
            counts[2] = failed;
        }
        return failed ? ARS_FALSE : ARS_TRUE;
    }
    if ((e = hipMemcpyAsync(st->h_counts, d_count, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return ars::hip_error(e, "syndrome count");
    listed = st->h_counts[0];

    if (listed) {
        std::vector<uint8_t> syn((size_t)listed * T);
        if ((e = hipMemcpyAsync(syn.data(), d_syn, syn.size(), hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return ars::hip_error(e, "syndrome copy");
        std::vector<ars::DecPlan> plans(listed);
        {
            std::lock_guard<std::mutex> lk(g_solver_mu);
            const size_t per = 16, pieces = (listed + per - 1) / per;
            solver_pool().run(pieces, [&](size_t pi) {
                const size_t end = std::min<size_t>(listed, (pi + 1) * per);
                for (size_t i = pi * per; i < end; ++i)
                    if (!ars::solve_key_equation(&syn[i * T], n, T, formal, plans[i])) plans[i].deg = -1;
            });
        }
        if (st->plan_count < listed) {
            hipFree(st->plans);
            st->plans = nullptr;
            st->plan_count = 0;
            if ((e = hipMalloc(reinterpret_cast<void **>(&st->plans), listed * sizeof(ars::DecPlan))) != hipSuccess)
                return ars::hip_error(e, "decode plans");
            st->plan_count = listed;
        }
        if ((e = hipMemcpyAsync(st->plans, plans.data(), plans.size() * sizeof(ars::DecPlan),
                                hipMemcpyHostToDevice, s)) != hipSuccess)
            return ars::hip_error(e, "decode plans copy");
        if ((e = ars::launch_decode_tail(d_received, (uint32_t)L, n, tl, d_cols, st->plans, listed, d_status,
                                         d_failed, ctx->d_gtab, ctx->d_tw, s)) != hipSuccess)
            return ars::hip_error(e, "decode launch");
        if ((e = hipMemcpyAsync(st->h_counts + 1, d_failed, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return ars::hip_error(e, "decode result");
        failed = st->h_counts[1];
    }
    if (counts) {
        counts[0] = L - listed;
        counts[1] = listed - failed;
        counts[2] = failed;
    }
    return failed ? ARS_FALSE : ARS_TRUE;
}
