// ars_batch_plan.hpp -- turns a batch's validated jobs into work items.
// Host-only, HIP-free (unit-tested on the CPU by tests/test_batch_plan.py).
//
// Every item holds one code shape (kind, transform size, accumulate flag,
// chunk/block argument), so all lanes of a workgroup run the same barriers.
// Codewords of at least one item's width become runs of at most kRepMax whole
// tiles (single-segment items); smaller ones are packed side by side into
// shared items, up to batch_max_segs segments, when their buffers lie within
// 4 GiB above the item's base pointers.  A packed item's segments are
// contiguous in the segment array, so a class's large codewords are emitted
// before its small ones are packed.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

#include "ars_batch_desc.hpp"

namespace ars {

constexpr uint32_t kRepMax = 4;  // tiles per item for large codewords

struct ClassKey {
    uint32_t kind, tl, acc, arg;
    bool operator<(const ClassKey &o) const {
        if (tl != o.tl) return tl > o.tl;  // larger transforms first: they are the long items
        if (kind != o.kind) return kind < o.kind;
        if (acc != o.acc) return acc < o.acc;
        return arg < o.arg;
    }
    bool operator==(const ClassKey &o) const {
        return kind == o.kind && tl == o.tl && acc == o.acc && arg == o.arg;
    }
};

struct Planned {
    ClassKey key;
    uint32_t job;  // index into the job array
};

inline void plan_items(std::vector<Planned> &planned, const std::vector<BJob> &jobs,
                       std::vector<BItem> &items, std::vector<BSeg> &segs) {
    std::stable_sort(planned.begin(), planned.end(),
                     [](const Planned &x, const Planned &y) { return x.key < y.key; });
    size_t pi = 0;
    while (pi < planned.size()) {
        const ClassKey key = planned[pi].key;
        size_t pe = pi;
        while (pe < planned.size() && planned[pe].key == key) ++pe;
        const uint32_t lanes = batch_lanes((int)key.tl);
        const uint32_t maxseg = batch_max_segs(key.kind, (int)key.tl);
        const uint32_t meta = key.kind | (key.tl << 8) | (key.acc << 16);
        for (size_t t = pi; t < pe; ++t) {
            const uint32_t jid = planned[t].job;
            const BJob &bj = jobs[jid];
            if (!(maxseg == 1 || bj.quads >= lanes)) continue;
            const uint32_t tiles = (bj.quads + lanes - 1) / lanes;
            for (uint32_t u = 0; u < tiles; u += kRepMax) {
                const uint32_t rep = std::min(kRepMax, tiles - u);
                items.push_back({meta, key.arg, (uint32_t)segs.size(), 1u | (rep << 16), bj.in, bj.out});
                segs.push_back({jid, u * lanes, 0u, lanes});
            }
        }
        int32_t open = -1;  // packed item being filled
        uint32_t used = 0;
        for (size_t t = pi; t < pe; ++t) {
            const uint32_t jid = planned[t].job;
            const BJob &bj = jobs[jid];
            const uint32_t quads = bj.quads;
            if (maxseg == 1 || quads >= lanes) continue;
            bool fits = open >= 0 && used + quads <= lanes && (items[open].nseg_rep & 0xffffu) < maxseg;
            if (fits) {
                const BItem &it = items[open];
                fits = bj.in >= it.in_base && bj.out >= it.out_base &&
                       (uint64_t)(bj.in - it.in_base) + bj.in_bytes <= 0xffffffffull &&
                       (uint64_t)(bj.out - it.out_base) + bj.out_bytes <= 0xffffffffull;
            }
            if (!fits) {
                open = (int32_t)items.size();
                used = 0;
                items.push_back({meta, key.arg, (uint32_t)segs.size(), 0u | (1u << 16), bj.in, bj.out});
            }
            segs.push_back({jid, 0u, used, quads});
            items[open].nseg_rep += 1u;
            used += quads;
        }
        pi = pe;
    }
}

// Region of an item (batch_region): its transform size, body variant and
// addressing.  The kernel walks regions largest transform first, then by
// variant, then single-codeword before packed items.
inline uint32_t item_region(const BItem &it) {
    const uint32_t kind = it.meta & 0xffu, tl = (it.meta >> 8) & 0xffu, acc = (it.meta >> 16) & 1u;
    const uint32_t variant = kind == kEncRef ? acc : kind == kEncKpow2 ? 2u : 3u;
    return batch_region(tl, variant, (it.nseg_rep & 0xffffu) > 1u ? 1u : 0u);
}
inline uint32_t region_rank(uint32_t g) {  // position of region g in the kernel's walk
    const uint32_t tl = g / (2u * kBatchVariants), rest = g % (2u * kBatchVariants);
    return (uint32_t)(kBatchClasses - 1 - tl) * (2u * kBatchVariants) + rest;
}

// Reorders items into the kernel's region order (stable: within a region the
// planner's order stays), then fills each region's item range; an empty
// region gets begin == end.  Items keep their segment indices, so the
// segment array is unchanged.
inline void order_regions(std::vector<BItem> &items, uint32_t *begin, uint32_t *end) {
    std::stable_sort(items.begin(), items.end(), [](const BItem &a, const BItem &b) {
        return region_rank(item_region(a)) < region_rank(item_region(b));
    });
    for (int g = 0; g < kBatchRegions; ++g) begin[g] = end[g] = 0;
    for (size_t i = 0; i < items.size(); ++i) {
        const uint32_t g = item_region(items[i]);
        if (begin[g] == end[g]) begin[g] = (uint32_t)i;
        end[g] = (uint32_t)i + 1;
    }
}

// Walks the items exactly as the persistent kernel maps lanes to codeword
// quads and checks that every quad of every planned job is covered exactly
// once, that every item stays within its shape's lanes and segment limit,
// and that packed codewords fit the item's 32-bit offset window.  Returns an
// empty string or a description of the first violation.
inline const char *verify_plan(const std::vector<Planned> &planned, const std::vector<BJob> &jobs,
                               const std::vector<BItem> &items, const std::vector<BSeg> &segs) {
    std::vector<std::vector<uint8_t>> hit(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) hit[j].assign(jobs[j].quads, 0);
    for (const BItem &it : items) {
        const uint32_t tl = (it.meta >> 8) & 0xffu, kind = it.meta & 0xffu;
        const uint32_t lanes = batch_lanes((int)tl);
        const uint32_t nseg = it.nseg_rep & 0xffffu, rep = it.nseg_rep >> 16;
        if (nseg == 0 || rep == 0) return "empty item";
        if (nseg > batch_max_segs(kind, (int)tl)) return "too many segments";
        if (nseg > 1 && rep != 1) return "packed item with repetitions";
        if (it.seg0 + nseg > segs.size()) return "segment range out of bounds";
        uint32_t lane_end = 0;
        for (uint32_t s = 0; s < nseg; ++s) {
            const BSeg &sg = segs[it.seg0 + s];
            if (sg.job >= jobs.size()) return "segment job out of range";
            if (sg.lane0 < lane_end) return "segments overlap";
            lane_end = sg.lane0 + sg.nlanes;
            if (lane_end > lanes) return "segment exceeds the item's lanes";
            const BJob &bj = jobs[sg.job];
            if (nseg > 1 && (bj.in < it.in_base || bj.out < it.out_base ||
                             (uint64_t)(bj.in - it.in_base) + bj.in_bytes > 0xffffffffull ||
                             (uint64_t)(bj.out - it.out_base) + bj.out_bytes > 0xffffffffull))
                return "packed codeword outside the 32-bit window";
            for (uint32_t r = 0; r < rep; ++r)
                for (uint32_t l = 0; l < sg.nlanes; ++l) {
                    const uint32_t q = sg.quad0 + l + r * lanes;
                    if (q >= bj.quads) continue;  // lanes past the shard store nothing
                    if (hit[sg.job][q]++) return "quad covered twice";
                }
        }
    }
    for (const Planned &p : planned)
        for (uint8_t h : hit[p.job])
            if (h != 1) return "quad not covered";
    return "";
}

}  // namespace ars
