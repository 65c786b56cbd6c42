"""ars_recover's erasure plan buffer, kept per (context, stream): recovers
queued back to back on one stream, interleaved over several streams, issued
from two host threads on one stream, on hipStreamPerThread and on more
streams than the context keeps buffers for all match the oracle."""
import ctypes
import threading

import numpy as np
import pytest

import _pkg
import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HIP_STREAM_PER_THREAD = 2  # hipStreamPerThread
MAX_PLAN_SLOTS = 64        # streams per context with a kept plan buffer


def split_len(n):
    """A shard length whose recover launch has more workgroups than the GPU
    has compute units, so ars_recover takes the plan-buffer path (smaller
    launches compute the plan inside the recover kernel and need no buffer)."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    cols = 1024 if n <= 16 else (128 if n == 256 else 256)  # byte columns per workgroup
    return cols * cus + 64


@pytest.fixture(scope="module")
def ars():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _pkg.load()


def make_jobs(n, k, L, count, seed):
    """`count` erased codewords on the device, each with its own erasure
    pattern and count, plus the expected codewords."""
    T = n - k
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        msg = rng.integers(0, 256, k * L, dtype=np.uint8)
        cw = np.concatenate([O.encode(n, T, msg, L), msg])
        e = int(rng.integers(1, T + 1))
        pos = rng.permutation(n)[:e].astype(np.uint8)
        rx = cw.copy().reshape(n, L)
        rx[pos] = 0
        jobs.append((torch.from_numpy(rx.reshape(-1).copy()).cuda(), pos, cw))
    torch.cuda.synchronize()
    return jobs


def check(jobs):
    torch.cuda.synchronize()
    for i, (d, _, cw) in enumerate(jobs):
        assert np.array_equal(d.cpu().numpy(), cw), f"job {i}"


def test_back_to_back_on_one_stream(ars):
    ctx = ars.Context(0)
    n, k = 256, 128
    L = split_len(n)
    jobs = make_jobs(n, k, L, 12, 1)
    s = torch.cuda.Stream()
    for d, pos, _ in jobs:
        assert ctx.recover(n, k, d.data_ptr(), L, pos, s) == 1
    s.synchronize()
    check(jobs)
    ctx.close()


def test_interleaved_streams(ars):
    ctx = ars.Context(0)
    n, k = 128, 64
    L = split_len(n)
    jobs = make_jobs(n, k, L, 16, 2)
    streams = [torch.cuda.Stream() for _ in range(3)] + [torch.cuda.current_stream()]
    for i, (d, pos, _) in enumerate(jobs):
        assert ctx.recover(n, k, d.data_ptr(), L, pos, streams[i % len(streams)]) == 1
    check(jobs)
    ctx.close()


def test_two_threads_one_stream(ars):
    ctx = ars.Context(0)
    n, k = 256, 128
    L = split_len(n)
    jobs = make_jobs(n, k, L, 24, 3)
    s = torch.cuda.Stream()
    rcs = []

    def issue(part):
        for d, pos, _ in part:
            rcs.append(ctx.recover(n, k, d.data_ptr(), L, pos, s))

    th = [threading.Thread(target=issue, args=(jobs[t::2],)) for t in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert rcs == [1] * len(jobs)
    s.synchronize()
    check(jobs)
    ctx.close()


def test_per_thread_stream_and_many_streams(ars):
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime torch and the library share
    ctx = ars.Context(0)
    n, k = 64, 32
    L = split_len(n)
    count = MAX_PLAN_SLOTS + 8
    jobs = make_jobs(n, k, L, count + 2, 4)
    raw = []
    try:
        for _ in range(count):
            h = ctypes.c_void_p()
            assert hip.hipStreamCreate(ctypes.byref(h)) == 0
            raw.append(h.value)
        # every stream past the kept-buffer limit allocates per call
        for (d, pos, _), h in zip(jobs, raw):
            assert ctx.recover(n, k, d.data_ptr(), L, pos, h) == 1
        for d, pos, _ in jobs[count:]:
            assert ctx.recover(n, k, d.data_ptr(), L, pos, HIP_STREAM_PER_THREAD) == 1
        for h in raw:
            assert hip.hipStreamSynchronize(ctypes.c_void_p(h)) == 0
        assert hip.hipStreamSynchronize(ctypes.c_void_p(HIP_STREAM_PER_THREAD)) == 0
        check(jobs)
    finally:
        for h in raw:
            hip.hipStreamDestroy(ctypes.c_void_p(h))
    ctx.close()
