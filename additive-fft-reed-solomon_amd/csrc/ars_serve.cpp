// ars_serve.cpp -- the serving half of the C ABI: a request loop with
// continuous batching over the batched persistent kernel.
//
// Requests are submitted from any thread.  One dispatcher thread:
//   1. retires finished requests (event queries): statuses, latencies, pool
//      space -- each request at the completion mark recorded after its own
//      work, so a small request does not wait for a large one launched after
//      it on the same slot;
//   2. admits waiting requests in arrival order while the HBM budget (a
//      device pool holding each admitted request's codeword) has room --
//      a request that does not fit waits, later smaller ones may pass it;
//   3. as soon as one of `pipeline_depth` batch slots (batch object + stream)
//      is free, launches everything admitted (up to max_batch_jobs) as one
//      batch, so requests arriving while others run join the next batch and
//      overlap with it on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ars.h"
#include "ars_ctx.hpp"
#include "ars_serve_poll.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kPoolAlign = 256;
// Default batch thresholds (ars_server_config::min_batch_jobs = 0): slots
// with fewer batchable jobs run them as plain calls, the second bound while
// no other slot is busy.
constexpr size_t kMinBatchJobs = 4;
constexpr size_t kIdleBatchJobs = 16;

// First-fit allocator over [0, cap) with coalescing frees.
class Pool {
  public:
    explicit Pool(size_t cap) { free_[0] = cap; }
    bool alloc(size_t bytes, size_t &off) {
        bytes = (bytes + kPoolAlign - 1) & ~(kPoolAlign - 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + kPoolAlign - 1) & ~(kPoolAlign - 1);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }

  private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0;
};

struct Req {
    uint64_t ticket = 0;
    ars_request r{};
    std::vector<uint8_t> pos;
    size_t bytes = 0, off = 0;
    uint64_t alg_bytes = 0;
    Clock::time_point t_submit, t_done;
    int32_t status = 0;
    bool done = false;
};

struct Slot {
    ars_batch *batch = nullptr;
    hipStream_t stream = nullptr;
    std::vector<Req *> reqs;
    std::vector<ars_job> jobs;    // run in the batched kernel
    std::vector<ars_job> direct;  // run as plain calls (codewords of direct_min_bytes or more)
    // Completion marks, recorded on the stream in order: mark m completes
    // reqs[order[mark_end[m - 1]] .. order[mark_end[m] - 1]] (mark_end[-1] = 0).
    std::vector<hipEvent_t> evs;  // grown on demand, kept until the server is destroyed
    std::vector<size_t> mark_end, order;
    size_t marks = 0, next_mark = 0, next_req = 0;
    bool synced = false;  // a mark could not be recorded: the stream was synchronised instead
    bool busy = false;
    bool ran_batch = false;  // the last launch used the batched kernel
};

// Records a completion mark for sl.order[prev .. end); false if no event
// could be created or recorded.
bool record_mark(Slot &sl) {
    if (sl.marks == sl.evs.size()) {
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return false;
        sl.evs.push_back(ev);
        sl.mark_end.push_back(0);
    }
    if (hipEventRecord(sl.evs[sl.marks], sl.stream) != hipSuccess) return false;
    sl.mark_end[sl.marks++] = sl.order.size();
    return true;
}

}  // namespace

struct ars_server {
    ars_ctx *ctx = nullptr;
    ars_server_config cfg{};
    uint8_t *d_pool = nullptr;
    std::unique_ptr<Pool> pool;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> waiting;
    std::vector<Req *> admitted;
    std::unordered_map<uint64_t, Req *> tickets;
    std::vector<Slot> slots;
    std::thread disp;
    bool stop = false;
    std::atomic<uint64_t> pending{0};  // submissions (and the stop request) the dispatcher has not yet seen
    uint64_t next_ticket = 1, outstanding = 0;
    Clock::time_point t0;
    // statistics since the last reset
    std::vector<double> lat;
    uint64_t submitted = 0, completed = 0, rejected = 0, batches = 0, batch_jobs = 0, direct_jobs = 0, alg = 0;
    double first_arrival = -1, last_completion = 0;
};

namespace {

double secs(const ars_server *s, Clock::time_point t) {
    return std::chrono::duration<double>(t - s->t0).count();
}

void launch_slot(ars_server *s, Slot &sl, bool loaded) {
    // a request whose input copy cannot be queued fails with that error and
    // stays out of the batch; codewords of direct_min_bytes or more run as
    // plain calls after the batch, on the same stream
    std::vector<size_t> in_batch, in_direct, in_failed;
    in_batch.reserve(sl.reqs.size());
    sl.jobs.clear();
    sl.direct.clear();
    sl.order.clear();
    sl.marks = sl.next_mark = sl.next_req = 0;
    sl.synced = false;
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        Req *q = sl.reqs[i];
        const ars_request &r = q->r;
        const size_t L = (size_t)r.shard_len, T = r.n - r.k;
        uint8_t *d = s->d_pool + q->off;
        ars_job j{};
        j.op = r.op;
        j.n = r.n;
        j.k = r.k;
        j.flags = r.flags;
        j.shard_len = r.shard_len;
        hipError_t e = hipSuccess;
        if (r.op == ARS_OP_ENCODE) {
            j.d_parity = d;
            j.d_message = d + T * L;
            if (r.host_in) e = hipMemcpyAsync(d + T * L, r.host_in, r.k * L, hipMemcpyHostToDevice, sl.stream);
        } else {
            j.d_codeword = d;
            j.positions = q->pos.data();
            j.e = r.e;
            if (r.host_in) e = hipMemcpyAsync(d, r.host_in, r.n * L, hipMemcpyHostToDevice, sl.stream);
        }
        if (e != hipSuccess) {
            q->status = ars::hip_error(e, "request input copy");
            in_failed.push_back(i);
            continue;
        }
        if ((uint64_t)r.n * r.shard_len >= s->cfg.direct_min_bytes) {
            in_direct.push_back(i);
            sl.direct.push_back(j);
        } else {
            in_batch.push_back(i);
            sl.jobs.push_back(j);
        }
    }
    // a batch's fixed cost (planning, descriptor upload, its launches) pays
    // off only over several jobs: fewer run as plain calls.  With no other
    // slot busy the GPU would otherwise idle through that fixed cost, so the
    // bar is higher; under load, batching keeps the launches few.
    const size_t min_jobs =
        (size_t)ars::min_batch_jobs(s->cfg.min_batch_jobs, loaded, kMinBatchJobs, kIdleBatchJobs, s->cfg.max_batch_jobs);
    if (!sl.jobs.empty() && sl.jobs.size() < min_jobs) {
        in_direct.insert(in_direct.end(), in_batch.begin(), in_batch.end());
        sl.direct.insert(sl.direct.end(), sl.jobs.begin(), sl.jobs.end());
        in_batch.clear();
        sl.jobs.clear();
    }
    sl.ran_batch = !sl.jobs.empty();
    auto finish_request = [&](Req *q, int32_t status) {
        const ars_request &r = q->r;
        q->status = status;
        if (!r.host_out || q->status < 0) return;
        const size_t L = (size_t)r.shard_len;
        uint8_t *d = s->d_pool + q->off;
        hipError_t e = hipSuccess;
        if (r.op == ARS_OP_ENCODE) e = hipMemcpyAsync(r.host_out, d, (r.n - r.k) * L, hipMemcpyDeviceToHost, sl.stream);
        else if (q->status == ARS_TRUE) e = hipMemcpyAsync(r.host_out, d, r.n * L, hipMemcpyDeviceToHost, sl.stream);
        if (e != hipSuccess) q->status = ars::hip_error(e, "request output copy");
    };
    // requests whose input copy failed complete with the first mark
    sl.order.insert(sl.order.end(), in_failed.begin(), in_failed.end());
    // the batched jobs complete together, then each plain call on its own;
    // every request's output copy precedes its mark, so its pool space may
    // be reused as soon as the mark has completed
    if (!sl.jobs.empty()) {
        const int rc_b = ars_batch_run(sl.batch, sl.jobs.data(), (uint32_t)sl.jobs.size(), 0, sl.stream);
        for (size_t b = 0; b < in_batch.size(); ++b) {
            finish_request(sl.reqs[in_batch[b]], rc_b < 0 ? rc_b : sl.jobs[b].status);
            sl.order.push_back(in_batch[b]);
        }
        (void)record_mark(sl);
    }
    for (size_t b = 0; b < in_direct.size(); ++b) {
        const int rc_d = ars_run_jobs_plain(s->ctx, &sl.direct[b], 1, sl.stream);
        finish_request(sl.reqs[in_direct[b]], rc_d < 0 ? rc_d : sl.direct[b].status);
        sl.order.push_back(in_direct[b]);
        (void)record_mark(sl);
    }
    // requests after the last recorded mark: mark them now, or wait here
    const size_t marked = sl.marks ? sl.mark_end[sl.marks - 1] : 0;
    if (marked < sl.order.size() && !record_mark(sl)) {
        (void)hipStreamSynchronize(sl.stream);
        sl.synced = true;
    }
    sl.busy = true;
}

// Caller holds s->mu.
void finish(ars_server *s, Req *q, Clock::time_point now) {
    q->t_done = now;
    q->done = true;
    s->lat.push_back(std::chrono::duration<double>(now - q->t_submit).count());
    s->completed++;
    if (q->status >= 0) s->alg += q->alg_bytes;
    s->last_completion = secs(s, now);
    s->outstanding--;
}

void dispatcher(ars_server *s) {
    hipSetDevice(s->ctx->device);
    // per slot: the marks found complete by this poll, in order, and how each ended
    std::vector<size_t> ready(s->slots.size(), 0);
    std::vector<std::vector<ars::SlotPoll>> polled(s->slots.size());
    // after work, keep polling this long before sleeping on the condition
    // variable: waking a sleeping thread costs tens of microseconds, which a
    // request arriving shortly after the last one would otherwise pay
    const auto linger = std::chrono::microseconds(500);
    Clock::time_point last_work = Clock::now();
    for (;;) {
        // 1. poll the slots' marks without the lock (the slots belong to
        //    this thread), so submitters are not held up meanwhile.  Marks
        //    complete in stream order: stop at the first one still running.
        bool any_done = false, busy = false;
        for (size_t i = 0; i < s->slots.size(); ++i) {
            Slot &sl = s->slots[i];
            polled[i].clear();
            ready[i] = sl.next_mark;
            if (!sl.busy) continue;
            busy = true;
            while (ready[i] < sl.marks) {
                const ars::SlotPoll p = ars::classify_slot_poll((int)hipEventQuery(sl.evs[ready[i]]), (int)hipSuccess,
                                                                (int)hipErrorNotReady);
                if (p == ars::SlotPoll::kRunning) break;
                polled[i].push_back(p);
                ++ready[i];
            }
            any_done |= ready[i] > sl.next_mark || (sl.synced && ready[i] == sl.marks);
        }
        if (!any_done && s->pending.load(std::memory_order_acquire) == 0) {
            if (busy || Clock::now() - last_work < linger) {
                // work in flight (or just done): poll again at once.  A timed
                // wait of a few microseconds oversleeps by the kernel's timer
                // slack (tens of microseconds), which every request's latency
                // would pay.
                std::this_thread::yield();
                continue;
            }
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->waiting.empty() && s->admitted.empty() && !s->stop)
                s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->pending.exchange(0, std::memory_order_relaxed) != 0 || any_done) last_work = Clock::now();
        // 2. retire finished requests
        const Clock::time_point now = Clock::now();
        for (size_t i = 0; i < s->slots.size(); ++i) {
            Slot &sl = s->slots[i];
            if (!sl.busy) continue;
            auto retire_to = [&](size_t end, ars::SlotPoll p) {
                for (; sl.next_req < end; ++sl.next_req) {
                    Req *q = sl.reqs[sl.order[sl.next_req]];
                    q->status = ars::retired_status(p, q->status, ARS_E_HIP);
                    s->pool->release(q->off, q->bytes);
                    finish(s, q, now);
                }
            };
            bool retired = false;
            for (const ars::SlotPoll p : polled[i]) {
                retire_to(sl.mark_end[sl.next_mark++], p);
                retired = true;
            }
            // past the last mark: the stream was synchronised at launch
            if (sl.next_mark == sl.marks && sl.synced && sl.next_req < sl.order.size()) {
                retire_to(sl.order.size(), ars::SlotPoll::kDone);
                retired = true;
            }
            if (sl.next_req == sl.order.size()) {
                sl.reqs.clear();
                sl.busy = false;
            }
            if (retired) s->cv_done.notify_all();
        }
        // 3. admission against the budget
        for (auto it = s->waiting.begin(); it != s->waiting.end();) {
            Req *q = *it;
            size_t off;
            if (s->pool->alloc(q->bytes, off)) {
                q->off = off;
                s->admitted.push_back(q);
                it = s->waiting.erase(it);
            } else {
                ++it;
            }
        }
        // 4. launch on the free slots
        for (Slot &sl : s->slots) {
            if (s->admitted.empty()) break;
            if (sl.busy) continue;
            const size_t take = std::min<size_t>(s->admitted.size(), s->cfg.max_batch_jobs);
            sl.reqs.assign(s->admitted.begin(), s->admitted.begin() + take);
            s->admitted.erase(s->admitted.begin(), s->admitted.begin() + take);
            bool loaded = false;
            for (const Slot &o : s->slots) loaded |= o.busy;
            lk.unlock();
            launch_slot(s, sl, loaded);
            lk.lock();
            // only slots that ran the batched kernel count as batches
            if (sl.ran_batch) {
                s->batches++;
                s->batch_jobs += sl.jobs.size();
            }
            s->direct_jobs += sl.direct.size();
        }
        bool any_busy = false;
        for (const Slot &sl : s->slots) any_busy |= sl.busy;
        if (s->stop && s->waiting.empty() && s->admitted.empty() && !any_busy) break;
    }
}

}  // namespace

extern "C" {

int ars_server_create(ars_ctx *ctx, const ars_server_config *cfg, ars_server **out) {
    if (!ctx || !out) return ars::set_error(ARS_E_NULL, "null context or out");
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return ars::hip_error(e, "hipSetDevice");
    ars_server *s = new (std::nothrow) ars_server;
    if (!s) return ars::set_error(ARS_E_NOMEM, "server allocation failed");
    s->ctx = ctx;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.hbm_budget) s->cfg.hbm_budget = 64ull << 30;
    if (!s->cfg.max_batch_jobs) s->cfg.max_batch_jobs = 1024;
    if (!s->cfg.pipeline_depth) s->cfg.pipeline_depth = 2;
    if (!s->cfg.direct_min_bytes) s->cfg.direct_min_bytes = 16ull << 20;
    if ((e = hipMalloc(&s->d_pool, s->cfg.hbm_budget)) != hipSuccess) {
        delete s;
        return ars::hip_error(e, "hipMalloc(server pool)");
    }
    s->pool.reset(new Pool(s->cfg.hbm_budget));
    s->slots.resize(s->cfg.pipeline_depth);
    for (Slot &sl : s->slots) {
        int rc = ars_batch_create(ctx, s->cfg.max_batch_jobs, &sl.batch);
        if (rc == ARS_OK && (e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess)
            rc = ars::hip_error(e, "hipStreamCreate");
        if (rc != ARS_OK) {
            ars_server_destroy(s);
            return rc;
        }
    }
    s->t0 = Clock::now();
    s->disp = std::thread(dispatcher, s);
    *out = s;
    return ARS_OK;
}

int ars_server_destroy(ars_server *s) {
    if (!s) return ARS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->pending.fetch_add(1, std::memory_order_release);
    }
    s->cv_work.notify_all();
    if (s->disp.joinable()) s->disp.join();
    hipSetDevice(s->ctx->device);
    for (Slot &sl : s->slots) {
        if (sl.stream) hipStreamSynchronize(sl.stream);
        if (sl.batch) ars_batch_destroy(sl.batch);
        for (hipEvent_t ev : sl.evs) hipEventDestroy(ev);
        if (sl.stream) hipStreamDestroy(sl.stream);
    }
    for (auto &kv : s->tickets) delete kv.second;
    hipFree(s->d_pool);
    delete s;
    return ARS_OK;
}

int ars_server_submit(ars_server *s, const ars_request *req, uint64_t *ticket) {
    if (!s || !req || !ticket) return ars::set_error(ARS_E_NULL, "null server, request or ticket");
    int rc = ars::check_code(req->n, req->k, (size_t)req->shard_len);
    if (rc) return rc;
    if (req->op != ARS_OP_ENCODE && req->op != ARS_OP_RECOVER) return ars::set_error(ARS_E_INVAL, "unknown op");
    if (req->flags) return ars::set_error(ARS_E_INVAL, "server requests take no flags");
    if (req->op == ARS_OP_RECOVER && req->e && !req->positions)
        return ars::set_error(ARS_E_NULL, "positions is null");
    Req *q = new (std::nothrow) Req;
    if (!q) return ars::set_error(ARS_E_NOMEM, "request allocation failed");
    q->r = *req;
    const size_t L = (size_t)req->shard_len;
    q->bytes = (size_t)req->n * L;
    if (q->bytes > s->cfg.hbm_budget) {
        delete q;
        std::lock_guard<std::mutex> lk(s->mu);
        s->rejected++;
        return ars::set_error(ARS_E_BUDGET, "request larger than the server's HBM budget");
    }
    if (req->op == ARS_OP_RECOVER && req->e) {
        q->pos.assign(req->positions, req->positions + std::min<uint32_t>(req->e, 256));
        q->r.positions = nullptr;
        if (req->e > 256) q->r.e = 256;  // e > T anyway: the job reports false
    }
    q->alg_bytes = req->op == ARS_OP_ENCODE ? (uint64_t)req->n * L : (uint64_t)(req->n + req->e) * L;
    q->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        q->ticket = s->next_ticket++;
        s->tickets[q->ticket] = q;
        s->waiting.push_back(q);
        s->submitted++;
        s->outstanding++;
        if (s->first_arrival < 0) s->first_arrival = secs(s, q->t_submit);
        *ticket = q->ticket;
        s->pending.fetch_add(1, std::memory_order_release);
    }
    s->cv_work.notify_one();
    return ARS_OK;
}

int ars_server_wait(ars_server *s, uint64_t ticket, int32_t *status, double *latency_s) {
    if (!s) return ars::set_error(ARS_E_NULL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return ars::set_error(ARS_E_STATE, "unknown or already collected ticket");
    // another thread may collect (and free) the same ticket meanwhile: look
    // it up again after every wake-up
    s->cv_done.wait(lk, [s, ticket] {
        auto f = s->tickets.find(ticket);
        return f == s->tickets.end() || f->second->done;
    });
    it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return ars::set_error(ARS_E_STATE, "ticket collected by another thread");
    Req *q = it->second;
    if (status) *status = q->status;
    if (latency_s) *latency_s = std::chrono::duration<double>(q->t_done - q->t_submit).count();
    s->tickets.erase(it);
    delete q;
    s->cv_done.notify_all();  // a second waiter on this ticket returns
    return ARS_OK;
}

int ars_server_drain(ars_server *s) {
    if (!s) return ars::set_error(ARS_E_NULL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [s] { return s->outstanding == 0; });
    return ARS_OK;
}

int ars_server_stats_get(ars_server *s, ars_server_stats *o) {
    if (!s || !o) return ars::set_error(ARS_E_NULL, "null server or stats");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(o, 0, sizeof *o);
    o->submitted = s->submitted;
    o->completed = s->completed;
    o->rejected = s->rejected;
    o->batches = s->batches;
    o->algorithmic_bytes = s->alg;
    o->mean_batch_jobs = s->batches ? (double)s->batch_jobs / (double)s->batches : 0.0;
    o->batched_jobs = s->batch_jobs;
    o->direct_jobs = s->direct_jobs;
    o->first_arrival_s = s->first_arrival;
    o->last_completion_s = s->last_completion;
    if (!s->lat.empty()) {
        std::vector<double> v = s->lat;
        auto pct = [&v](double p) {
            const size_t i = std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5));
            std::nth_element(v.begin(), v.begin() + i, v.end());
            return v[i];
        };
        o->p50_latency_s = pct(0.50);
        o->p99_latency_s = pct(0.99);
        o->max_latency_s = *std::max_element(v.begin(), v.end());
    }
    return ARS_OK;
}

int ars_server_stats_reset(ars_server *s) {
    if (!s) return ars::set_error(ARS_E_NULL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat.clear();
    s->submitted = s->completed = s->rejected = s->batches = s->batch_jobs = s->direct_jobs = s->alg = 0;
    s->first_arrival = -1;
    s->last_completion = 0;
    return ARS_OK;
}

int ars_server_replay(ars_server *s, const ars_request *reqs, const double *arrival_s, uint32_t count,
                      int32_t *statuses, double *latencies_s) {
    if (!s || (count && (!reqs || !arrival_s))) return ars::set_error(ARS_E_NULL, "null server or trace");
    std::vector<uint64_t> tickets(count, 0);
    std::vector<int> rcs(count, ARS_OK);
    std::thread gen([&] {
        const Clock::time_point start = Clock::now();
        for (uint32_t i = 0; i < count; ++i) {
            const Clock::time_point at = start + std::chrono::duration_cast<Clock::duration>(
                                                     std::chrono::duration<double>(arrival_s[i]));
            while (Clock::now() < at) {
                const auto left = at - Clock::now();
                if (left > std::chrono::microseconds(200)) std::this_thread::sleep_for(left - std::chrono::microseconds(100));
            }
            rcs[i] = ars_server_submit(s, &reqs[i], &tickets[i]);
        }
    });
    gen.join();
    int rc = ARS_OK;
    for (uint32_t i = 0; i < count; ++i) {
        int32_t st = rcs[i];
        double lat = 0;
        if (rcs[i] == ARS_OK) {
            int w = ars_server_wait(s, tickets[i], &st, &lat);
            if (w) rc = w;
        }
        if (statuses) statuses[i] = st;
        if (latencies_s) latencies_s[i] = lat;
    }
    return rc;
}

int ars_plain_replay(ars_ctx *ctx, const ars_request *reqs, const double *arrival_s, uint32_t count,
                     uint8_t *d_buffer, size_t buffer_bytes, int32_t *statuses, double *latencies_s,
                     void *stream) {
    if (!ctx || (count && (!reqs || !arrival_s || !d_buffer)))
        return ars::set_error(ARS_E_NULL, "null context, trace or buffer");
    for (uint32_t i = 0; i < count; ++i) {
        const ars_request &r = reqs[i];
        if (r.op != ARS_OP_ENCODE && r.op != ARS_OP_RECOVER) return ars::set_error(ARS_E_INVAL, "unknown op");
        if ((uint64_t)r.n * r.shard_len > buffer_bytes)
            return ars::set_error(ARS_E_INVAL, "a request's codeword does not fit the buffer");
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return ars::hip_error(e, "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // One event per request in flight, reused round-robin: request i records
    // ring[i % R] only after the completion thread has passed request i - R.
    constexpr uint32_t R = 256;
    std::vector<hipEvent_t> ring(R, nullptr);
    for (hipEvent_t &ev : ring)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
            for (hipEvent_t x : ring)
                if (x) hipEventDestroy(x);
            return ars::hip_error(e, "hipEventCreate");
        }
    std::vector<int32_t> st(count, ARS_OK);
    std::vector<double> lat(count, 0.0);
    std::atomic<uint32_t> issued{0}, retired{0};
    const Clock::time_point start = Clock::now();
    auto due = [&](uint32_t i) {
        return start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
    };
    std::thread completion([&] {
        hipSetDevice(ctx->device);
        for (uint32_t i = 0; i < count; ++i) {
            while (issued.load(std::memory_order_acquire) <= i) std::this_thread::yield();
            // a failed kernel or copy on the stream fails the request (st[i]
            // is not written by the issuing thread after `issued` passes i)
            if (hipEventSynchronize(ring[i % R]) != hipSuccess && st[i] >= 0) st[i] = ARS_E_HIP;
            lat[i] = std::chrono::duration<double>(Clock::now() - due(i)).count();
            retired.store(i + 1, std::memory_order_release);
        }
    });
    // One stream runs its work in order, so every request can use the same
    // device buffer: request i's input copy starts after request i - 1's
    // output copy has finished.
    for (uint32_t i = 0; i < count; ++i) {
        const ars_request &r = reqs[i];
        while (Clock::now() < due(i)) {
            const auto left = due(i) - Clock::now();
            if (left > std::chrono::microseconds(200)) std::this_thread::sleep_for(left - std::chrono::microseconds(100));
        }
        while (i >= R && retired.load(std::memory_order_acquire) < i - R + 1) std::this_thread::yield();
        const size_t L = (size_t)r.shard_len, T = r.n - r.k;
        int rc = ARS_OK;
        if (r.host_in)
            e = hipMemcpyAsync(r.op == ARS_OP_ENCODE ? d_buffer + T * L : d_buffer, r.host_in,
                               (r.op == ARS_OP_ENCODE ? r.k : r.n) * L, hipMemcpyHostToDevice, s);
        if (r.host_in && e != hipSuccess) rc = ars::hip_error(e, "request input copy");
        if (rc == ARS_OK)
            rc = r.op == ARS_OP_ENCODE ? ars_encode(ctx, r.n, r.k, d_buffer + T * L, d_buffer, L, r.flags, s)
                                       : ars_recover(ctx, r.n, r.k, d_buffer, L, r.positions, r.e, s);
        if (rc >= 0 && r.host_out && (r.op == ARS_OP_ENCODE || rc == ARS_TRUE)) {
            e = hipMemcpyAsync(r.host_out, d_buffer, (r.op == ARS_OP_ENCODE ? T : r.n) * L, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) rc = ars::hip_error(e, "request output copy");
        }
        st[i] = rc;
        if ((e = hipEventRecord(ring[i % R], s)) != hipSuccess) (void)hipStreamSynchronize(s);
        issued.store(i + 1, std::memory_order_release);
    }
    completion.join();
    for (hipEvent_t ev : ring) hipEventDestroy(ev);
    for (uint32_t i = 0; i < count; ++i) {
        if (statuses) statuses[i] = st[i];
        if (latencies_s) latencies_s[i] = lat[i];
    }
    return ARS_OK;
}

}  // extern "C"
