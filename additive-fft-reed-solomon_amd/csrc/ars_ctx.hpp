// ars_ctx.hpp -- the per-device context shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ars_gf.hpp"
#include "ars_kernels.hpp"

namespace ars {
// ars_recover's erasure plan buffer for one stream of a context.  `mu` is held
// while a call enqueues its plan kernel and its tile kernel, so two host
// threads on one stream cannot interleave the pair.
struct PlanSlot {
    hipStream_t stream = nullptr;
    RecPrep *buf = nullptr;
    std::mutex mu;
};
constexpr size_t kMaxPlanSlots = 64;  // streams per context with a kept plan buffer
}  // namespace ars

struct ars_ctx {
    int device = 0;
    ars::MulTab *d_tw = nullptr;  // 255 twiddle multipliers, resident
    uint8_t *d_gtab = nullptr;    // exp / log / omega for the recover plan kernels
    int num_cus = 0;
    // autotuned tile variant per (op, log2 transform size): 0 = default
    // op 0 = encode (reference algorithm), 1 = recover
    uint8_t tune_pq[2][9] = {{0}};
    uint8_t tune_nbuf[2][9] = {{0}};
    // ars_recover's plan buffers, one per stream the context has recovered on
    // (at most kMaxPlanSlots; later streams allocate per call)
    std::mutex plan_mu;
    std::vector<std::unique_ptr<ars::PlanSlot>> plan_slots;
    // this device's address of the file-scope plan scratch of
    // ars_recover_noalloc (ars_api.cpp), from hipGetSymbolAddress at creation
    ars::RecPrep *noalloc_plan = nullptr;
};

namespace ars {
int set_error(int code, const std::string &msg);
int hip_error(hipError_t e, const char *where);
// wide: the call runs codewords of 4 GiB or more (n * shard_len >= 2^32);
// otherwise those are rejected.  shard_len itself stays below 4 GiB.
int check_code(uint32_t n, uint32_t k, size_t L, bool wide = false);
bool is_pow2(uint32_t v);
int ilog2(uint32_t v);
void release_host_pipe(ars_ctx *ctx);     // ars_host.cpp
void release_decode_state(ars_ctx *ctx);  // ars_decode_api.cpp
ars_ctx *batch_context(const ars_batch *b);  // ars_batch_host.cpp
}  // namespace ars
