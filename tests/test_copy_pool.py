"""CPU unit test of the host worker pool
(csrc/ars_work_pool.hpp), built with g++ under ThreadSanitizer: every index of
every job runs exactly once, jobs run back to back with no overlap, and pools
of 0..8 workers start and stop cleanly."""
import os
import shutil
import subprocess

import pytest

import _pkg

CSRC = os.path.join(_pkg.PKG_DIR, "csrc")

DRIVER = r'''
#include "ars_work_pool.hpp"
#include <cstdio>
#include <vector>
using namespace ars;

int main() {
    int bad = 0;
    for (unsigned workers = 0; workers <= 8; ++workers) {
        WorkPool pool(workers);
        for (size_t job = 0; job < 300; ++job) {
            const size_t count = (job * 37) % 97;
            std::vector<std::atomic<int>> hits(count);
            for (auto &h : hits) h.store(0);
            std::atomic<int> inside{0};
            size_t plain_sum = 0;  // written by one index each, read after run()
            std::vector<size_t> out(count, 0);
            pool.run(count, [&](size_t i) {
                inside.fetch_add(1);
                hits[i].fetch_add(1);
                out[i] = i * 3 + job;
                inside.fetch_sub(1);
            });
            if (inside.load() != 0) ++bad;
            for (size_t i = 0; i < count; ++i) {
                if (hits[i].load() != 1) ++bad;
                plain_sum += out[i];
            }
            if (plain_sum != 3 * (count * (count - (count ? 1 : 0)) / 2) + job * count) ++bad;
        }
    }
    std::printf("bad=%d\n", bad);
    return bad ? 1 : 0;
}
'''


def test_copy_pool_runs_every_index_once(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ missing")
    src = tmp_path / "pool.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "pool"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread", "-I", CSRC,
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0 and "tsan" in r.stderr.lower():
        pytest.skip("ThreadSanitizer runtime unavailable")
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-3000:]
