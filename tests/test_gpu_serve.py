"""GPU tests of the serving loop: results are the single-call results whether
a request runs alone or in a mix submitted concurrently (in the batched kernel,
as plain calls, both in one slot, or by the default rules), the HBM budget
admits requests as memory frees up, and over-budget requests are rejected."""
import threading

import numpy as np
import pytest

import _pkg
import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ars():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _pkg.load()


SHAPES = [(64, 32), (128, 96), (256, 128), (256, 224), (32, 16), (256, 16)]


def make_request_data(rng, n, k, L, op):
    T = n - k
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    if op == "encode":
        exp = O.encode_any(n, k, msg, L) if (T & (T - 1)) == 0 else O.ext_encode_kpow2(n, k, msg, L)
        return msg, np.zeros(T * L, np.uint8), exp, ()
    cw = np.concatenate([O.encode_any(n, k, msg, L), msg])
    e = int(rng.integers(1, T + 1))
    pos = rng.permutation(n)[:e].astype(np.uint8)
    rx = cw.copy().reshape(n, L)
    rx[pos] = 0
    return rx.reshape(-1).copy(), np.zeros(n * L, np.uint8), cw, pos


@pytest.mark.parametrize("direct_min_bytes,min_batch_jobs", [(2 ** 64 - 1, 1), (1, 0), (128 << 10, 1), (0, 0)])
def test_server_alone_equals_in_mix(ars, direct_min_bytes, min_batch_jobs):
    """Each request alone and in a mix submitted from 4 threads: every job in
    the batched kernel (even a slot of one); every job as a plain call; a
    slot split by codeword size between the two; the default rules."""
    rng = np.random.default_rng(45)
    specs = []
    for i in range(48):
        n, k = SHAPES[rng.integers(len(SHAPES))]
        L = int(64 * rng.integers(1, 64))
        op = "encode" if rng.random() < 0.7 else "recover"
        specs.append((op, n, k, L) + make_request_data(rng, n, k, L, op))
    srv = ars.Server(hbm_budget=256 << 20, max_batch_jobs=16, pipeline_depth=2, direct_min_bytes=direct_min_bytes,
                     min_batch_jobs=min_batch_jobs)
    # alone: one request at a time
    alone = []
    for op, n, k, L, inp, out, exp, pos in specs:
        o = np.zeros_like(out)
        keep = []
        r = ars.Server.request(op, n, k, L, pos, host_in=inp, host_out=o, keep=keep)
        st, _ = srv.wait(srv.submit(r))
        assert st == (0 if op == "encode" else 1)
        assert np.array_equal(o, exp), (op, n, k, L)
        alone.append(o)
    # in the mix: submitted from 4 threads at once
    outs = [np.zeros_like(sp[5]) for sp in specs]
    keep = []
    reqs = [ars.Server.request(sp[0], sp[1], sp[2], sp[3], sp[7], host_in=sp[4], host_out=outs[i], keep=keep)
            for i, sp in enumerate(specs)]
    tickets = [None] * len(reqs)

    def worker(idx):
        for i in idx:
            tickets[i] = srv.submit(reqs[i])

    th = [threading.Thread(target=worker, args=(range(t, len(reqs), 4),)) for t in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, tk in enumerate(tickets):
        st, lat = srv.wait(tk)
        assert st >= 0 and lat > 0
        assert np.array_equal(outs[i], alone[i]), i
    stats = srv.stats()
    total = 2 * len(specs)
    assert stats["completed"] == total
    assert stats["batched_jobs"] + stats["direct_jobs"] == total
    if direct_min_bytes == 1:  # every job a plain call: the batched kernel never ran
        assert stats["batches"] == 0 and stats["direct_jobs"] == total
    elif min_batch_jobs == 1 and direct_min_bytes == 2 ** 64 - 1:  # every job in the batched kernel
        assert stats["direct_jobs"] == 0 and stats["batched_jobs"] == total
        assert stats["batches"] < total  # the mix shared launches
        assert stats["mean_batch_jobs"] == pytest.approx(total / stats["batches"])
    elif min_batch_jobs == 1:  # split by codeword size: both paths ran
        assert stats["batched_jobs"] > 0 and stats["direct_jobs"] > 0
    srv.close()


def test_server_budget_admission_and_rejection(ars):
    rng = np.random.default_rng(46)
    n, k, L = 256, 128, 64 * 1024  # 16 MiB codewords
    budget = 40 << 20             # two at a time
    srv = ars.Server(hbm_budget=budget, max_batch_jobs=8, pipeline_depth=2)
    keep, outs, exps, reqs = [], [], [], []
    for _ in range(8):
        inp, out, exp, pos = make_request_data(rng, n, k, L, "encode")
        keep.append(inp)  # host buffers must outlive the request
        outs.append(out)
        exps.append(exp)
        reqs.append(ars.Server.request("encode", n, k, L, host_in=inp, host_out=out, keep=keep))
    tickets = [srv.submit(r) for r in reqs]
    for t, o, e in zip(tickets, outs, exps):
        st, _ = srv.wait(t)
        assert st == 0 and np.array_equal(o, e)
    big = ars.Server.request("encode", 256, 128, 1 << 20)
    with pytest.raises(ars.ArsError) as ei:
        srv.submit(big)
    assert ei.value.code == ars.ARS_E_BUDGET
    srv.close()


def test_server_replay_statuses(ars):
    srv = ars.Server(hbm_budget=1 << 30, max_batch_jobs=64)
    keep = []
    reqs = (ars.Request * 6)()
    reqs[0] = ars.Server.request("encode", 64, 32, 4096)
    reqs[1] = ars.Server.request("recover", 64, 32, 4096, positions=list(range(33)), keep=keep)
    reqs[2] = ars.Server.request("recover", 64, 32, 4096, positions=[5, 5], keep=keep)
    reqs[3] = ars.Server.request("recover", 64, 32, 4096, positions=[1, 2, 3], keep=keep)
    reqs[4] = ars.Server.request("encode", 256, 16, 1024)
    reqs[5] = ars.Server.request("recover", 128, 96, 1024, positions=[], keep=keep)
    st, lat = srv.replay(reqs, [0.0, 0.001, 0.001, 0.002, 0.002, 0.003])
    assert list(st) == [0, 0, ars.ARS_E_DUPLICATE, 1, 0, 1]
    assert (lat > 0).all()
    srv.close()


def test_server_two_waiters_one_ticket(ars):
    """Two threads waiting on one ticket: one collects the status, the other
    is told the ticket is gone (never a freed request)."""
    srv = ars.Server(hbm_budget=256 << 20, max_batch_jobs=8)
    rng = np.random.default_rng(47)
    for rep in range(4):
        inp, out, exp, _ = make_request_data(rng, 256, 128, 1 << 16, "encode")
        keep = [inp, out]
        t = srv.submit(ars.Server.request("encode", 256, 128, 1 << 16, host_in=inp, host_out=out, keep=keep))
        got = []

        def waiter():
            try:
                got.append(("ok",) + tuple(srv.wait(t)))
            except ars.ArsError as err:
                got.append(("err", err.code))

        th = [threading.Thread(target=waiter) for _ in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join(timeout=60)
        assert sorted(g[0] for g in got) == ["err", "ok"], got
        assert [g[1] for g in got if g[0] == "err"] == [ars.ARS_E_STATE]
        assert [g[1] for g in got if g[0] == "ok"] == [0]
        assert np.array_equal(out, exp)
    srv.close()


def test_plain_replay_matches_oracle(ars):
    """ars_plain_replay, the served stream's plain-call baseline: a trace of
    host-payload requests with arrivals that overlap is served one call per
    request on one stream through one device buffer, and every output equals
    the oracle; invalid requests report their error in `statuses`."""
    rng = np.random.default_rng(47)
    specs = []
    for i in range(40):
        n, k = SHAPES[rng.integers(len(SHAPES))]
        L = int(64 * rng.integers(1, 64))
        op = "encode" if rng.random() < 0.7 else "recover"
        specs.append((op, n, k, L) + make_request_data(rng, n, k, L, op))
    keep = []
    outs = [np.zeros_like(sp[5]) for sp in specs]
    reqs = (ars.Request * (len(specs) + 1))()
    for i, sp in enumerate(specs):
        reqs[i] = ars.Server.request(sp[0], sp[1], sp[2], sp[3], sp[7], host_in=sp[4], host_out=outs[i], keep=keep)
    reqs[len(specs)] = ars.Server.request("recover", 64, 32, 64, [3, 3], keep=keep)  # duplicate position
    arrivals = np.concatenate([np.sort(rng.uniform(0, 0.004, len(specs))), [0.005]])
    biggest = max(sp[1] * sp[3] for sp in specs)
    buf = torch.empty(biggest, dtype=torch.uint8, device="cuda")
    ctx = ars.Context(0)
    st, lat = ctx.plain_replay(reqs, arrivals, buf.data_ptr(), biggest)
    for i, sp in enumerate(specs):
        assert st[i] == (0 if sp[0] == "encode" else 1), i
        assert lat[i] > 0
        assert np.array_equal(outs[i], sp[6]), (i, sp[:4])
    assert st[len(specs)] == ars.ARS_E_DUPLICATE
    # a request larger than the buffer is rejected before anything runs
    with pytest.raises(ars.ArsError) as ei:
        ctx.plain_replay(reqs, arrivals, buf.data_ptr(), biggest - 1)
    assert ei.value.code == ars.ARS_E_INVAL
    ctx.close()
