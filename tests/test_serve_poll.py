"""CPU unit test of the serving loop's slot retirement (csrc/ars_serve_poll.hpp):
a failed event query ends the slot and fails its requests (it must not leave
the slot busy, which would hang the dispatcher and every waiter), and the
batch thresholds never exceed max_batch_jobs."""
import os
import shutil
import subprocess

import pytest

import _pkg

DRIVER = r'''
#include "ars_serve_poll.hpp"
#include <cstdio>
#include <initializer_list>
using namespace ars;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL %s\n", #c); return 1; } } while (0)
int main() {
    // codes as the HIP runtime defines them: success 0, not ready 600; any other value is a failure
    const int ok = 0, nr = 600;
    CHECK(classify_slot_poll(ok, ok, nr) == SlotPoll::kDone);
    CHECK(classify_slot_poll(nr, ok, nr) == SlotPoll::kRunning);
    for (int e : {1, 2, 98, 700, 719, 999}) CHECK(classify_slot_poll(e, ok, nr) == SlotPoll::kFailed);
    // a failed slot fails every request that had not already failed
    CHECK(retired_status(SlotPoll::kFailed, 0, -5) == -5);
    CHECK(retired_status(SlotPoll::kFailed, 1, -5) == -5);
    CHECK(retired_status(SlotPoll::kFailed, -3, -5) == -3);
    CHECK(retired_status(SlotPoll::kDone, 1, -5) == 1);
    CHECK(retired_status(SlotPoll::kDone, 0, -5) == 0);
    // thresholds: configured value wins; defaults depend on load; capped by the slot size
    CHECK(min_batch_jobs(0, true, 4, 16, 1024) == 4);
    CHECK(min_batch_jobs(0, false, 4, 16, 1024) == 16);
    CHECK(min_batch_jobs(0, false, 4, 16, 8) == 8);
    CHECK(min_batch_jobs(0, true, 4, 16, 2) == 2);
    CHECK(min_batch_jobs(7, false, 4, 16, 1024) == 7);
    CHECK(min_batch_jobs(7, false, 4, 16, 5) == 5);
    std::printf("ok\n");
    return 0;
}
'''


def test_slot_retirement_and_thresholds(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ missing")
    src = tmp_path / "poll.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "poll"
    inc = os.path.join(_pkg.PKG_DIR, "csrc")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-fsanitize=address,undefined",
                        "-I", inc, str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


def test_hip_codes_used_by_the_server():
    """The HIP status codes the C test program above assumes are the runtime's."""
    hdr = "/opt/rocm/include/hip/hip_runtime_api.h"
    if not os.path.exists(hdr):
        pytest.skip("HIP headers missing")
    txt = open(hdr).read()
    assert "hipSuccess = 0" in txt and "hipErrorNotReady = 600" in txt
