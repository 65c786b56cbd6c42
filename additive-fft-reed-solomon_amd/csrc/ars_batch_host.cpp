// ars_batch_host.cpp -- the batch half of the C ABI: validation, work-item
// construction, descriptor upload, hipGraph capture and replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/ars.h"
#include "ars_batch.hpp"
#include "ars_batch_plan.hpp"
#include "ars_ctx.hpp"

struct ars_batch {
    ars_ctx *ctx = nullptr;
    uint32_t max_jobs = 0;
    uint8_t *d_arena = nullptr;
    size_t arena_cap = 0;
    uint8_t *h_stage = nullptr;  // pinned mirror of the arena's descriptor part
    size_t stage_cap = 0;
    hipEvent_t upload_done = nullptr;
    hipEvent_t run_done = nullptr;  // the last run's kernels (any stream) have finished reading the arena
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t grid = 0, prep_grid = 0, blocks_per_cu = 0;
    ars_batch_info info{};
    // host scratch, reused across runs
    std::vector<ars::BItem> items;
    std::vector<ars::BSeg> segs;
    std::vector<ars::BJob> jobs;
    std::vector<ars::BPrep> preps;
    std::vector<uint8_t> pos;
    // outgrown arenas and staging buffers, freed by ars_batch_destroy
    std::vector<uint8_t *> retired_arenas, retired_stages;
};

namespace {

using ars::ClassKey;
using ars::Planned;

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }
bool aligned4(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 3u) == 0; }

void destroy_graph(ars_batch *b) {
    if (b->exec) {
        hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
}

int capture_graph(ars_batch *b) {
    destroy_graph(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return ars::hip_error(e, "hipStreamBeginCapture");
    hipError_t e1 = ars::launch_batch_reset(b->d_arena, b->cap_stream);
    hipError_t e2 = ars::launch_batch_prep(b->d_arena, b->ctx->d_gtab, b->prep_grid, b->cap_stream);
    hipError_t e3 = ars::launch_batch(b->d_arena, b->ctx->d_tw, b->grid, b->cap_stream);
    e = hipStreamEndCapture(b->cap_stream, &g);
    if (e1 != hipSuccess) e = e1;
    if (e2 != hipSuccess) e = e2;
    if (e3 != hipSuccess) e = e3;
    if (e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return ars::hip_error(e, "batch graph capture");
    }
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->exec = nullptr;
        return ars::hip_error(e, "hipGraphInstantiate");
    }
    return ARS_OK;
}

int ensure_capacity(ars_batch *b, size_t arena_bytes, size_t stage_bytes) {
    hipError_t e;
    // An outgrown buffer may still be read by work in flight: a previous
    // run's kernels (on any stream) or its upload.  It is kept until
    // ars_batch_destroy rather than freed here, since a free synchronises the
    // whole device and would stall every stream's work.  Capacities grow by
    // half, so what is kept stays below twice the final size.
    if (arena_bytes > b->arena_cap) {
        destroy_graph(b);  // captured with the old arena's address
        if (b->d_arena) b->retired_arenas.push_back(b->d_arena);
        b->d_arena = nullptr;
        b->arena_cap = 0;
        const size_t cap = arena_bytes + arena_bytes / 2;
        if ((e = hipMalloc(&b->d_arena, cap)) != hipSuccess) return ars::hip_error(e, "hipMalloc(batch arena)");
        b->arena_cap = cap;
    }
    if (stage_bytes > b->stage_cap) {
        if (b->h_stage) b->retired_stages.push_back(b->h_stage);
        b->h_stage = nullptr;
        b->stage_cap = 0;
        const size_t cap = stage_bytes + stage_bytes / 2;
        if ((e = hipHostMalloc(&b->h_stage, cap, hipHostMallocDefault)) != hipSuccess)
            return ars::hip_error(e, "hipHostMalloc(batch staging)");
        b->stage_cap = cap;
    }
    return ARS_OK;
}

}  // namespace

namespace ars {
ars_ctx *batch_context(const ars_batch *b) { return b->ctx; }
}  // namespace ars

extern "C" {

int ars_batch_create(ars_ctx *ctx, uint32_t max_jobs, ars_batch **out) {
    if (!ctx || !out) return ars::set_error(ARS_E_NULL, "null context or out");
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return ars::hip_error(e, "hipSetDevice");
    ars_batch *b = new (std::nothrow) ars_batch;
    if (!b) return ars::set_error(ARS_E_NOMEM, "batch allocation failed");
    b->ctx = ctx;
    b->max_jobs = max_jobs ? max_jobs : 1;
    if ((e = hipEventCreateWithFlags(&b->upload_done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->run_done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking)) != hipSuccess) {
        ars_batch_destroy(b);
        return ars::hip_error(e, "batch events/streams");
    }
    b->prep_grid = 256;
    const int bpc = ars::batch_blocks_per_cu();
    b->grid = (uint32_t)(bpc * (ctx->num_cus > 0 ? ctx->num_cus : 256));
    // size the arena and its staging mirror up front for up to 4096 jobs of a
    // few work items each, so that a serving loop's first large batches do
    // not grow them (and recapture the graph) while other work is in flight
    const size_t jobs = std::min<size_t>(b->max_jobs, 4096);
    const size_t per_job = sizeof(ars::BJob) + sizeof(ars::BPrep) + 4 * (sizeof(ars::BItem) + sizeof(ars::BSeg)) + 256;
    const size_t stage = align_up(sizeof(ars::BHeader)) + jobs * per_job;
    int rc = ensure_capacity(b, stage + jobs * sizeof(ars::RecPrep), stage);
    // capture (and instantiate) the graph now rather than on the first run
    if (rc == ARS_OK) rc = capture_graph(b);
    if (rc) {
        ars_batch_destroy(b);
        return rc;
    }
    *out = b;
    return ARS_OK;
}

int ars_batch_destroy(ars_batch *b) {
    if (!b) return ARS_OK;
    hipSetDevice(b->ctx->device);
    if (b->upload_done) hipEventSynchronize(b->upload_done);
    hipDeviceSynchronize();
    destroy_graph(b);
    if (b->cap_stream) hipStreamDestroy(b->cap_stream);
    if (b->upload_done) hipEventDestroy(b->upload_done);
    if (b->run_done) hipEventDestroy(b->run_done);
    hipFree(b->d_arena);
    hipHostFree(b->h_stage);
    for (uint8_t *p : b->retired_arenas) hipFree(p);
    for (uint8_t *p : b->retired_stages) hipHostFree(p);
    delete b;
    return ARS_OK;
}

int ars_batch_set_blocks_per_cu(ars_batch *b, uint32_t bpc) {
    if (!b) return ars::set_error(ARS_E_NULL, "null batch");
    const uint32_t occ = (uint32_t)ars::batch_blocks_per_cu();
    const uint32_t use = (bpc == 0 || bpc > occ) ? occ : bpc;
    const uint32_t grid = use * (uint32_t)(b->ctx->num_cus > 0 ? b->ctx->num_cus : 256);
    if (grid != b->grid) {
        b->grid = grid;
        destroy_graph(b);
    }
    b->blocks_per_cu = use;
    return ARS_OK;
}

int ars_run_jobs_plain(ars_ctx *ctx, ars_job *jobs, uint32_t count, void *stream) {
    if (!ctx || (!jobs && count)) return ars::set_error(ARS_E_NULL, "null context or jobs");
    for (uint32_t i = 0; i < count; ++i) {
        ars_job &j = jobs[i];
        if (j.op == ARS_OP_ENCODE)
            j.status = ars_encode(ctx, j.n, j.k, j.d_message, j.d_parity, (size_t)j.shard_len, j.flags, stream);
        else if (j.op == ARS_OP_RECOVER)
            j.status = ars_recover(ctx, j.n, j.k, j.d_codeword, (size_t)j.shard_len, j.positions, j.e, stream);
        else
            j.status = ars::set_error(ARS_E_INVAL, "unknown op");
    }
    return ARS_OK;
}

int ars_batch_info_get(const ars_batch *b, ars_batch_info *info) {
    if (!b || !info) return ars::set_error(ARS_E_NULL, "null batch or info");
    *info = b->info;
    return ARS_OK;
}

int ars_batch_run(ars_batch *b, ars_job *jobs, uint32_t count, uint32_t flags, void *stream) {
    if (!b || (!jobs && count)) return ars::set_error(ARS_E_NULL, "null batch or jobs");
    if (count > b->max_jobs) return ars::set_error(ARS_E_INVAL, "more jobs than the batch was created for");
    if (flags & ~ARS_BATCH_NO_GRAPH) return ars::set_error(ARS_E_INVAL, "unknown batch flags");
    hipError_t e = hipSetDevice(b->ctx->device);
    if (e != hipSuccess) return ars::hip_error(e, "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);

    b->items.clear();
    b->segs.clear();
    b->jobs.clear();
    b->preps.clear();
    b->pos.clear();
    std::vector<Planned> planned;
    std::vector<uint32_t> fallback;
    std::vector<int32_t> prep_of;  // per BJob: index into preps, or -1
    uint64_t alg_bytes = 0;

    // 1. validate every job exactly as the single-call entry points do
    for (uint32_t i = 0; i < count; ++i) {
        ars_job &j = jobs[i];
        const uint64_t L64 = j.shard_len;
        int rc = ars::check_code(j.n, j.k, (size_t)L64);
        if (rc) {
            j.status = rc;
            continue;
        }
        const size_t L = (size_t)L64;
        const uint32_t n = j.n, k = j.k, T = n - k;
        if (j.op == ARS_OP_ENCODE) {
            if (!j.d_message || !j.d_parity) {
                j.status = ars::set_error(ARS_E_NULL, "encode job without buffers");
                continue;
            }
            if (j.flags & ~ARS_ENC_ACCUMULATE) {
                j.status = ars::set_error(ARS_E_INVAL, "unknown encode flags");
                continue;
            }
            const bool acc = (j.flags & ARS_ENC_ACCUMULATE) != 0;
            if (!ars::is_pow2(T) && acc) {
                j.status = ars::set_error(ARS_E_UNSUPPORTED, "ACCUMULATE needs T a power of two");
                continue;
            }
            j.status = ARS_OK;
            alg_bytes += (uint64_t)n * L + (acc ? (uint64_t)T * L : 0);
            if (L % 4 || !aligned4(j.d_message) || !aligned4(j.d_parity)) {
                fallback.push_back(i);
                continue;
            }
            ClassKey key;
            if (ars::is_pow2(T)) key = {ars::kEncRef, (uint32_t)ars::ilog2(T), acc ? 1u : 0u, k / T};
            else key = {ars::kEncKpow2, (uint32_t)ars::ilog2(k), 0u, (n / k - 1) | ((n - k) << 16)};
            planned.push_back({key, (uint32_t)b->jobs.size()});
            b->jobs.push_back({j.d_message, j.d_parity, nullptr, (uint32_t)L, (uint32_t)(L / 4),
                               (uint32_t)((size_t)k * L), (uint32_t)((size_t)T * L)});
            prep_of.push_back(-1);
        } else if (j.op == ARS_OP_RECOVER) {
            if (!j.d_codeword) {
                j.status = ars::set_error(ARS_E_NULL, "recover job without a codeword");
                continue;
            }
            if (j.e > T) {  // rs/src/codec.rs:491-493
                j.status = ARS_FALSE;
                continue;
            }
            if (j.e == 0) {
                j.status = ARS_TRUE;
                continue;
            }
            if (!j.positions) {
                j.status = ars::set_error(ARS_E_NULL, "positions is null");
                continue;
            }
            bool seen[256] = {false}, bad = false;
            for (uint32_t t = 0; t < j.e && !bad; ++t) {
                const uint8_t p = j.positions[t];
                if (p >= n) {
                    j.status = ars::set_error(ARS_E_POSITION, "erasure position >= n");
                    bad = true;
                } else if (seen[p]) {
                    j.status = ars::set_error(ARS_E_DUPLICATE, "duplicate erasure position");
                    bad = true;
                }
                seen[p] = true;
            }
            if (bad) continue;
            j.status = ARS_TRUE;
            alg_bytes += (uint64_t)(n + j.e) * L;
            if (L % 4 || !aligned4(j.d_codeword)) {
                fallback.push_back(i);
                continue;
            }
            prep_of.push_back((int32_t)b->preps.size());
            b->preps.push_back({nullptr, n, j.e, (uint32_t)b->pos.size(), 0});
            b->pos.insert(b->pos.end(), j.positions, j.positions + j.e);
            b->pos.resize((b->pos.size() + 15) & ~size_t(15));
            planned.push_back({{ars::kRecover, (uint32_t)ars::ilog2(n), 0u, 0u}, (uint32_t)b->jobs.size()});
            b->jobs.push_back({j.d_codeword, j.d_codeword, nullptr, (uint32_t)L, (uint32_t)(L / 4),
                               (uint32_t)((size_t)n * L), (uint32_t)((size_t)n * L)});
        } else {
            j.status = ars::set_error(ARS_E_INVAL, "unknown op");
        }
    }

    // 2. work items (ars_batch_plan.hpp)
    ars::plan_items(planned, b->jobs, b->items, b->segs);

    // 3. arena layout: header | items | segs | jobs | preps | positions | plans
    ars::BHeader hdr{};
    hdr.nitems = (uint32_t)b->items.size();
    hdr.nprep = (uint32_t)b->preps.size();
    ars::order_regions(b->items, hdr.begin, hdr.end);
    size_t off = align_up(sizeof(ars::BHeader));
    hdr.off_items = off;
    off = align_up(off + b->items.size() * sizeof(ars::BItem));
    hdr.off_segs = off;
    off = align_up(off + b->segs.size() * sizeof(ars::BSeg));
    hdr.off_jobs = off;
    off = align_up(off + b->jobs.size() * sizeof(ars::BJob));
    hdr.off_preps = off;
    off = align_up(off + b->preps.size() * sizeof(ars::BPrep));
    hdr.off_pos = off;
    off = align_up(off + b->pos.size());
    const size_t stage_bytes = off;
    hdr.off_plans = off;
    const size_t arena_bytes = off + b->preps.size() * sizeof(ars::RecPrep);
    int rc = ensure_capacity(b, arena_bytes, stage_bytes);
    if (rc) return rc;
    for (size_t t = 0; t < b->jobs.size(); ++t) {
        if (prep_of[t] >= 0) {
            ars::RecPrep *plan = reinterpret_cast<ars::RecPrep *>(b->d_arena + hdr.off_plans) + prep_of[t];
            b->jobs[t].plan = plan;
            b->preps[(size_t)prep_of[t]].plan = plan;
        }
    }

    // 4. upload through the pinned staging buffer (reused only after the
    //    previous upload finished)
    if ((e = hipEventSynchronize(b->upload_done)) != hipSuccess) return ars::hip_error(e, "hipEventSynchronize");
    uint8_t *h = b->h_stage;
    std::memcpy(h, &hdr, sizeof hdr);
    if (!b->items.empty()) std::memcpy(h + hdr.off_items, b->items.data(), b->items.size() * sizeof(ars::BItem));
    if (!b->segs.empty()) std::memcpy(h + hdr.off_segs, b->segs.data(), b->segs.size() * sizeof(ars::BSeg));
    if (!b->jobs.empty()) std::memcpy(h + hdr.off_jobs, b->jobs.data(), b->jobs.size() * sizeof(ars::BJob));
    if (!b->preps.empty()) std::memcpy(h + hdr.off_preps, b->preps.data(), b->preps.size() * sizeof(ars::BPrep));
    if (!b->pos.empty()) std::memcpy(h + hdr.off_pos, b->pos.data(), b->pos.size());
    // a previous run on another stream may still read the arena: order the
    // upload after it on the device
    if ((e = hipStreamWaitEvent(s, b->run_done, 0)) != hipSuccess) return ars::hip_error(e, "hipStreamWaitEvent");
    if ((e = hipMemcpyAsync(b->d_arena, h, stage_bytes, hipMemcpyHostToDevice, s)) != hipSuccess)
        return ars::hip_error(e, "batch descriptor upload");
    if ((e = hipEventRecord(b->upload_done, s)) != hipSuccess) return ars::hip_error(e, "hipEventRecord");

    // 5. the batch: graph replay (captured once per arena / grid) or direct
    if (!(flags & ARS_BATCH_NO_GRAPH)) {
        if (!b->exec && (rc = capture_graph(b))) return rc;
        if ((e = hipGraphLaunch(b->exec, s)) != hipSuccess) return ars::hip_error(e, "hipGraphLaunch");
    } else {
        if ((e = ars::launch_batch_reset(b->d_arena, s)) != hipSuccess ||
            (e = ars::launch_batch_prep(b->d_arena, b->ctx->d_gtab, b->prep_grid, s)) != hipSuccess ||
            (e = ars::launch_batch(b->d_arena, b->ctx->d_tw, b->grid, s)) != hipSuccess)
            return ars::hip_error(e, "batch launch");
    }
    if ((e = hipEventRecord(b->run_done, s)) != hipSuccess) return ars::hip_error(e, "hipEventRecord");

    // 6. jobs the batched kernel cannot take run as plain calls
    for (uint32_t i : fallback) {
        ars_job &j = jobs[i];
        if (j.op == ARS_OP_ENCODE)
            j.status = ars_encode(b->ctx, j.n, j.k, j.d_message, j.d_parity, (size_t)j.shard_len, j.flags, stream);
        else
            j.status = ars_recover(b->ctx, j.n, j.k, j.d_codeword, (size_t)j.shard_len, j.positions, j.e, stream);
    }

    b->info.jobs = (uint32_t)b->jobs.size();
    b->info.items = (uint32_t)b->items.size();
    b->info.segments = (uint32_t)b->segs.size();
    b->info.plans = (uint32_t)b->preps.size();
    b->info.fallback_jobs = (uint32_t)fallback.size();
    b->info.grid = b->grid;
    b->info.algorithmic_bytes = alg_bytes;
    return ARS_OK;
}

}  // extern "C"
