// ars_batch_desc.hpp -- batch descriptors shared by the host planner, the
// batch host code and the persistent kernel.  Plain data: no HIP headers, so
// the planner can be unit-tested on a host without a GPU.
#pragma once
#include <cstdint>

namespace ars {

struct RecPrep;  // per-job recover plan (ars_kernels.hpp)


constexpr int kBatchThreads = 512;   // workgroup size of the persistent kernel
constexpr int kBatchMaxSegs = 64;    // segments (codeword pieces) per work item
constexpr int kBatchPlanSlots = 512; // recover: segments x n multipliers staged in LDS per item

// Most segments an item of this kind / transform size may hold.  Shapes up
// to 64 points pack several codewords into one item (their items are 128+
// column quads wide, so small shards would leave most lanes idle); larger
// shapes run one codeword per item with scalar addressing.
constexpr int kBatchPackMaxTl = 6;
constexpr uint32_t batch_max_segs(uint32_t kind, int tl) {
    return tl > kBatchPackMaxTl ? 1u
                   : kind == 2u ? ((uint32_t)kBatchPlanSlots >> tl < (uint32_t)kBatchMaxSegs
                                       ? (uint32_t)kBatchPlanSlots >> tl : (uint32_t)kBatchMaxSegs)
                                : (uint32_t)kBatchMaxSegs;
}

enum BatchKind : uint32_t { kEncRef = 0, kEncKpow2 = 1, kRecover = 2 };

// One codeword of a batch, as the kernel sees it.
struct alignas(16) BJob {
    const uint8_t *in;     // encode: message shards; recover: codeword
    uint8_t *out;          // encode: parity shards; recover: codeword
    const RecPrep *plan;   // recover: multiplier tables
    uint32_t L;            // shard_len
    uint32_t quads;        // column quads per shard (L / 4)
    uint32_t in_bytes;     // extent of `in`
    uint32_t out_bytes;    // extent of `out`
};

// A contiguous run of column quads of one job mapped onto lanes of an item.
struct alignas(16) BSeg {
    uint32_t job, quad0, lane0, nlanes;
};

// One unit of work for a workgroup: every segment shares the code shape
// (kind, transform size, chunk count), so all lanes run the same barriers.
// rep > 1 (single-segment items of large codewords): the item covers rep
// consecutive tiles, each advancing the segment's quads by the lane count.
// Items packing several codewords (nseg > 1) address all of them relative
// to in_base / out_base through one buffer resource; the host guarantees
// every packed codeword lies within 4 GiB above those bases.
struct alignas(16) BItem {
    uint32_t meta;      // kind | tl << 8 | acc << 16
    uint32_t arg;       // kEncRef: chunks (k/T); kEncKpow2: blocks | msg_base << 16
    uint32_t seg0;
    uint32_t nseg_rep;  // nseg | rep << 16
    const uint8_t *in_base;
    uint8_t *out_base;
};

// Recover plan job for the batched prep kernel.
struct alignas(16) BPrep {
    RecPrep *plan;
    uint32_t n, e, pos_off, pad;
};

// Regions of the persistent kernel: one per (transform size tl, body
// variant, addressing), so each tile body runs in a loop of its own.
// Variant: 0 encode, 1 encode with accumulate, 2 encode k = 2^KL, 3 recover.
// Addressing: 0 one codeword per item (scalar descriptor), 1 packed item.
constexpr int kBatchClasses = 9;
constexpr int kBatchVariants = 4;
constexpr int kBatchRegions = kBatchClasses * kBatchVariants * 2;
constexpr uint32_t batch_region(uint32_t tl, uint32_t variant, uint32_t packed) {
    return (tl * kBatchVariants + variant) * 2u + packed;
}

// Header at the start of the device arena (read by the graph's kernels, so
// the captured launches never change while the work behind them does).
// Offsets are in bytes from the arena base.  Items are sorted by region in
// the order the kernel walks them (largest transforms first); region g owns
// items [begin[g], end[g]) and hands them out through counters[g].
struct alignas(16) BHeader {
    uint32_t nitems, nprep, pad0, pad1;
    uint64_t off_items, off_segs, off_jobs, off_preps, off_pos, off_plans;
    uint32_t begin[kBatchRegions], end[kBatchRegions], counters[kBatchRegions];
};

// Lanes per item: tile kinds use kBatchThreads / Q column quads.
constexpr uint32_t batch_lanes(int tl) { return tl <= 4 ? (uint32_t)kBatchThreads : (uint32_t)(kBatchThreads >> (tl - 4)); }

}  // namespace ars
