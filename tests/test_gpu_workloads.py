"""Result checks on the workloads the batched and served figures come from.

* The mixed batch (bench.c3_workload: 4096 codewords, n ~ U{32..256}, k = n/2,
  shard_len log-uniform 64 B - 64 KiB, seed 44), run as one batch: every
  codeword's encode and recover equal single ars_encode / ars_recover calls
  on a copy of the same pool, a seeded sample of encodes equals the oracle,
  and with the erased shards zeroed every recover returns its codeword.
* The served Poisson stream (bench.c4_trace: 1 KiB - 4 MiB shards, 70 %
  encode / 30 % recover) with host payloads through ars_server_replay: every
  output equals a single call on the same input, and a seeded sample equals
  the oracle."""
import numpy as np
import pytest

import _pkg
import bench
import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ars():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _pkg.load()


def _zero_erased(pool, w):
    for i, pos in enumerate(w["positions"]):
        if len(pos):
            n, L, o = int(w["ns"][i]), int(w["Ls"][i]), int(w["offs"][i])
            pool[o:o + n * L].view(n, L).index_fill_(0, torch.from_numpy(pos.astype(np.int64)).cuda(), 0)


def _single_calls(ctx, w, pool, op):
    for i in range(len(w["ns"])):
        n, L, o = int(w["ns"][i]), int(w["Ls"][i]), int(w["offs"][i])
        k = n // 2
        T = n - k
        base = pool.data_ptr() + o
        if op == "encode":
            assert ctx.encode(n, k, base + T * L, base, L) == 0
        else:
            assert ctx.recover(n, k, base, L, w["positions"][i]) == 1


def test_c3_mixed_batch_equals_single_calls_and_oracle(ars):
    w = bench.c3_workload(np)
    count = len(w["ns"])
    gen = torch.Generator(device="cuda")
    gen.manual_seed(w["seed"])
    A = torch.randint(0, 256, (w["total"],), dtype=torch.uint8, device="cuda", generator=gen)
    B = A.clone()
    earr, rarr, keep, _ = bench.c3_jobs(ars, w, A)
    b = ars.Batch(count)
    ctx = b.ctx
    # encode: one batch on A, one call per codeword on B
    b.run_jobs(earr, count)
    _single_calls(ctx, w, B, "encode")
    torch.cuda.synchronize()
    assert all(earr[i].status == 0 for i in range(count))
    assert torch.equal(A, B), "batched encodes differ from single calls"
    # a seeded sample of encodes against the oracle (CPU)
    rng = np.random.default_rng(440)
    for i in rng.choice(count, 24, replace=False):
        n, L, o = int(w["ns"][i]), int(w["Ls"][i]), int(w["offs"][i])
        T = n - n // 2
        cw = A[o:o + n * L].cpu().numpy()
        assert np.array_equal(cw[:T * L], O.encode(n, T, cw[T * L:], L)), (i, n, L)
    encoded = A.clone()
    # recover with the erased shards zeroed: every codeword comes back
    _zero_erased(A, w)
    _zero_erased(B, w)
    b.run_jobs(rarr, count)
    _single_calls(ctx, w, B, "recover")
    torch.cuda.synchronize()
    assert all(rarr[i].status == 1 for i in range(count))
    assert torch.equal(A, B), "batched recovers differ from single calls"
    assert torch.equal(A, encoded), "a recover did not return its codeword"
    b.close()


def test_c4_served_stream_equals_single_calls_and_oracle(ars):
    trace, arrivals = bench.c4_trace(np, 2000, 80, 451)
    rng = np.random.default_rng(452)
    ctx = ars.Context(0)
    reqs = (ars.Request * len(trace))()
    keep, ins, outs, expect = [], [], [], []
    for i, (op, n, k, L, pos) in enumerate(trace):
        T = n - k
        msg = torch.randint(0, 256, (k * L,), dtype=torch.uint8, device="cuda")
        if op == "encode":
            d_par = torch.zeros(T * L, dtype=torch.uint8, device="cuda")
            assert ctx.encode(n, k, msg.data_ptr(), d_par.data_ptr(), L) == 0
            inp, out, exp = msg.cpu().numpy(), np.zeros(T * L, np.uint8), d_par
        else:
            cw = torch.zeros(n * L, dtype=torch.uint8, device="cuda")
            assert ctx.encode(n, k, msg.data_ptr(), cw.data_ptr(), L) == 0
            cw[T * L:] = msg
            rx = cw.view(n, L).clone()
            # zeroed or garbage erased shards (the result is c_p ^ g_p)
            fill = torch.randint(0, 256, (len(pos), L), dtype=torch.uint8, device="cuda") if rng.random() < 0.3 \
                else torch.zeros(len(pos), L, dtype=torch.uint8, device="cuda")
            rx[torch.from_numpy(pos.astype(np.int64)).cuda()] = fill
            rx = rx.reshape(-1)
            inp, out = rx.cpu().numpy(), np.zeros(n * L, np.uint8)
            exp = rx.clone()
            assert ctx.recover(n, k, exp.data_ptr(), L, pos) == 1
        torch.cuda.synchronize()
        ins.append(inp)
        outs.append(out)
        expect.append(exp.cpu().numpy())
        reqs[i] = ars.Server.request(op, n, k, L, positions=pos if pos is not None else (), host_in=inp,
                                     host_out=out, keep=keep)
    srv = ars.Server(hbm_budget=64 << 30, max_batch_jobs=1024, pipeline_depth=2)
    st, lat = srv.replay(reqs, arrivals)
    srv.close()
    for i, (op, n, k, L, pos) in enumerate(trace):
        assert st[i] == (0 if op == "encode" else 1), (i, op, st[i])
        assert lat[i] > 0
        assert np.array_equal(outs[i], expect[i]), (i, op, n, k, L)
    # oracle on a seeded sample of the smaller requests
    small = [i for i, t in enumerate(trace) if t[1] * t[3] <= (8 << 20)]
    for i in rng.choice(small, min(6, len(small)), replace=False):
        op, n, k, L, pos = trace[i]
        T = n - k
        if op == "encode":
            assert np.array_equal(outs[i], O.encode(n, T, ins[i], L)), i
        else:
            exp = ins[i].copy()
            O.recover(n, T, exp, L, pos.astype(np.uint8))
            assert np.array_equal(outs[i], exp), i
    ctx.close()
