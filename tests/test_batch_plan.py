"""CPU unit test of the batch work-item planner (csrc/ars_batch_plan.hpp):
for random mixes of codeword shapes and shard lengths, every column quad of
every job is covered exactly once by the lane mapping the persistent kernel
uses, items respect their shape's lane and segment limits, packed codewords
fit the 32-bit offset window, and the region order the kernel walks (one
contiguous item range per transform size, body variant and addressing)
covers every item.  Also checks that the verifier catches planted
mistakes."""
import os
import shutil
import subprocess

import pytest

import _pkg

CSRC = os.path.join(_pkg.PKG_DIR, "csrc")

DRIVER = r'''
#include "ars_batch_plan.hpp"
#include <cstdio>
#include <cstring>
#include <random>
using namespace ars;

static int run(unsigned seed, int count, bool plant) {
    std::mt19937_64 rng(seed);
    std::vector<BJob> jobs;
    std::vector<Planned> planned;
    const uint32_t shapes[][2] = {{2,1},{4,2},{8,4},{16,8},{32,16},{64,32},{128,64},{256,128},{128,96},{256,224},{256,16},{64,48}};
    uint64_t addr = 0x100000000ull;
    for (int i = 0; i < count; ++i) {
        const uint32_t *sh = shapes[rng() % 12];
        const uint32_t n = sh[0], k = sh[1], T = n - k;
        const uint32_t L = 4u * (1u + (uint32_t)(rng() % ((rng() % 4 == 0) ? 20000 : 64)));
        const int op = rng() % 2;
        BJob j{};
        j.L = L; j.quads = L / 4;
        ClassKey key;
        if (op == 0) {
            j.in_bytes = k * L; j.out_bytes = T * L;
            if ((T & (T - 1)) == 0) key = {kEncRef, (uint32_t)__builtin_ctz(T), (uint32_t)(rng() % 2), k / T};
            else key = {kEncKpow2, (uint32_t)__builtin_ctz(k), 0u, (n / k - 1) | ((n - k) << 16)};
        } else {
            j.in_bytes = j.out_bytes = n * L;
            key = {kRecover, (uint32_t)__builtin_ctz(n), 0u, 0u};
        }
        // mostly increasing addresses, sometimes a jump down or far away
        const int r = rng() % 10;
        if (r == 0) addr -= 1ull << 30; else if (r == 1) addr += 5ull << 30;
        j.in = reinterpret_cast<const uint8_t *>(addr); addr += j.in_bytes + 256;
        j.out = reinterpret_cast<uint8_t *>(addr); addr += j.out_bytes + 256;
        planned.push_back({key, (uint32_t)jobs.size()});
        jobs.push_back(j);
    }
    std::vector<BItem> items;
    std::vector<BSeg> segs;
    plan_items(planned, jobs, items, segs);
    // region order the persistent kernel walks: every item in its region's
    // range, ranges in the kernel's walk order, nothing left out
    uint32_t rb[kBatchRegions], re[kBatchRegions];
    order_regions(items, rb, re);
    size_t covered = 0;
    uint32_t last_rank = 0, last_end = 0;
    for (uint32_t rank = 0; rank < (uint32_t)kBatchRegions; ++rank) {
        for (uint32_t g = 0; g < (uint32_t)kBatchRegions; ++g) {
            if (region_rank(g) != rank || rb[g] == re[g]) continue;
            if (rb[g] != last_end) { std::printf("seed %u: region %u starts at %u, expected %u\n", seed, g, rb[g], last_end); return 1; }
            for (uint32_t i = rb[g]; i < re[g]; ++i)
                if (item_region(items[i]) != g) { std::printf("seed %u: item %u outside its region\n", seed, i); return 1; }
            covered += re[g] - rb[g];
            last_end = re[g];
            last_rank = rank;
        }
    }
    (void)last_rank;
    if (covered != items.size()) { std::printf("seed %u: regions cover %zu of %zu items\n", seed, covered, items.size()); return 1; }
    if (plant && !items.empty()) {
        // planted mistake: drop the last item, or shift one segment's lanes
        if (seed % 2) items.pop_back(); else segs[items[0].seg0].quad0 += 1;
    }
    const char *err = verify_plan(planned, jobs, items, segs);
    if (err[0]) { std::printf("seed %u: %s\n", seed, err); return 1; }
    return 0;
}

int main(int argc, char **argv) {
    const bool plant = argc > 1 && !std::strcmp(argv[1], "plant");
    int bad = 0;
    for (unsigned s = 0; s < 200; ++s) bad += run(s, 1 + (int)(s % 97) * 7, plant);
    std::printf("bad=%d\n", bad);
    return plant ? (bad == 200 ? 0 : 1) : (bad ? 1 : 0);
}
'''


def test_batch_plan_covers_every_quad_once(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ missing")
    src = tmp_path / "plan.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "plan"
    # under AddressSanitizer and UndefinedBehaviorSanitizer: every finding aborts
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-I", CSRC, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    # the verifier must catch planted mistakes in every run
    r = subprocess.run([str(exe), "plant"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
