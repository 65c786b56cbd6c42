"""GPU tests of the batched persistent kernel: every job of a mixed batch
must give exactly what the single-call entry point (and the oracle) gives
for it alone."""
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


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def make_mixed(rng, count, shapes, lens):
    """Encode + recover job specs with host-side expected results."""
    specs, checks = [], []
    for _ in range(count):
        n, k = shapes[rng.integers(len(shapes))]
        T = n - k
        L = int(lens[rng.integers(len(lens))])
        msg = rng.integers(0, 256, k * L, dtype=np.uint8)
        if rng.random() < 0.5:
            acc = bool((T & (T - 1)) == 0 and rng.random() < 0.3)
            pin = rng.integers(0, 256, T * L, dtype=np.uint8) if acc else np.zeros(T * L, np.uint8)
            if (T & (T - 1)) == 0:
                exp = O.encode(n, T, msg, L, parity=pin.copy())
            else:
                exp = O.ext_encode_kpow2(n, k, msg, L)
            dm, dp = dev(msg), dev(pin)
            specs.append({"op": "encode", "n": n, "k": k, "shard_len": L, "message": dm,
                          "parity": dp, "accumulate": acc})
            checks.append((dp, exp, 0))
        else:
            cw = np.concatenate([O.encode_any(n, k, msg, L), msg])
            e = int(rng.integers(0, T + 1))
            pos = rng.permutation(n)[:e].astype(np.uint8)
            rx = cw.copy().reshape(n, L)
            rx[pos] = rng.integers(0, 256, (e, L), dtype=np.uint8) if rng.random() < 0.3 else 0
            exp = rx.reshape(-1).copy()
            if e:
                if (T & (T - 1)) == 0:
                    O.recover(n, T, exp, L, pos)
                else:
                    O.ext_recover_lagrange(n, k, exp, L, pos)
            d = dev(rx.reshape(-1))
            specs.append({"op": "recover", "n": n, "k": k, "shard_len": L, "codeword": d,
                          "positions": pos})
            checks.append((d, exp, 1))
    return specs, checks


@pytest.mark.parametrize("graph", [True, False])
def test_batch_mixed_matches_oracle(ars, graph):
    rng = np.random.default_rng(44 if graph else 45)
    shapes = [(32, 16), (64, 32), (128, 64), (256, 128), (128, 96), (256, 224), (16, 8), (8, 4),
              (256, 16), (4, 2), (2, 1)]
    lens = [64, 128, 192, 1024, 4096, 65536 // 8]
    specs, checks = make_mixed(rng, 120, shapes, lens)
    b = ars.Batch(256)
    st = b.run(specs, graph=graph)
    torch.cuda.synchronize()
    info = b.info()
    assert info["fallback_jobs"] == 0
    for i, ((buf, exp, want), s) in enumerate(zip(checks, st)):
        assert s == want, (i, specs[i]["op"], s)
        assert np.array_equal(buf.cpu().numpy(), exp), (i, specs[i]["op"], specs[i]["n"], specs[i]["k"])


def test_batch_equals_single_calls_and_replays(ars):
    """The same batch object run three times (graph replays with different
    contents) gives, each time, what one-codeword calls give."""
    b = ars.Batch(64)
    for rep in range(3):
        rng = np.random.default_rng(100 + rep)
        specs, checks = make_mixed(rng, 40, [(64, 32), (256, 128), (32, 16), (128, 96)], [64, 256, 2048])
        # single-call results on copies
        singles = []
        for sp in specs:
            c = ars.Codec(sp["n"], sp["k"])
            if sp["op"] == "encode":
                p2 = sp["parity"].clone()
                c.encode_systematic_sharded(sp["message"], p2, None, sp["shard_len"],
                                            accumulate=sp["accumulate"])
                singles.append(p2)
            else:
                c2 = sp["codeword"].clone()
                c.recover_erasure_shards(c2, None, sp["shard_len"], sp["positions"])
                singles.append(c2)
        torch.cuda.synchronize()
        b.run(specs)
        torch.cuda.synchronize()
        for (buf, exp, _), single in zip(checks, singles):
            assert torch.equal(buf, single)
            assert np.array_equal(buf.cpu().numpy(), exp)


def test_batch_tiny_codewords_packed(ars):
    """Many (32,16) codewords with 64-byte shards share work items."""
    rng = np.random.default_rng(46)
    specs, checks = make_mixed(rng, 600, [(32, 16), (16, 8), (8, 4)], [64])
    b = ars.Batch(1024)
    st = b.run(specs)
    torch.cuda.synchronize()
    info = b.info()
    assert info["items"] < info["jobs"], info
    for (buf, exp, want), s in zip(checks, st):
        assert s == want
        assert np.array_equal(buf.cpu().numpy(), exp)


def test_batch_statuses_and_fallback(ars):
    rng = np.random.default_rng(47)
    n, k, L = 64, 32, 102  # L % 4 != 0 -> plain-call fallback
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    cw = np.concatenate([O.encode(n, n - k, msg, L), msg])
    rx = cw.copy().reshape(n, L)
    rx[[1, 5, 9]] = 0
    d = dev(rx.reshape(-1))
    good = dev(cw)
    specs = [
        {"op": "recover", "n": n, "k": k, "shard_len": L, "codeword": d, "positions": [1, 5, 9]},
        {"op": "recover", "n": n, "k": k, "shard_len": L, "codeword": good, "positions": list(range(33))},
        {"op": "recover", "n": n, "k": k, "shard_len": L, "codeword": good, "positions": [3, 3]},
        {"op": "recover", "n": n, "k": k, "shard_len": L, "codeword": good, "positions": [64]},
        {"op": "encode", "n": 48, "k": 16, "shard_len": L, "message": good, "parity": good},
    ]
    b = ars.Batch(8)
    st = b.run(specs)
    torch.cuda.synchronize()
    assert st == [1, 0, ars.ARS_E_DUPLICATE, ars.ARS_E_POSITION, ars.ARS_E_INVAL]
    assert b.info()["fallback_jobs"] == 1
    assert np.array_equal(d.cpu().numpy(), cw)
    assert np.array_equal(good.cpu().numpy(), cw)


def test_batch_object_runs_on_two_streams(ars):
    """Runs of one batch object issued back-to-back on two streams, with no
    host synchronisation between them: each run's descriptor upload waits on
    the device for the previous run, so every run reads its own jobs."""
    b = ars.Batch(128)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    runs = []
    for rep, count in enumerate((48, 12, 40, 8)):
        rng = np.random.default_rng(700 + rep)
        lens = [65536, 16384] if rep == 0 else [64, 256, 1024]
        specs, checks = make_mixed(rng, count, [(256, 128), (64, 32), (128, 96), (32, 16)], lens)
        torch.cuda.synchronize()  # inputs are on the device before the runs start
        st = b.run(specs, stream=streams[rep % 2])
        runs.append((specs, checks, st))
    torch.cuda.synchronize()
    for rep, (specs, checks, st) in enumerate(runs):
        for i, ((buf, exp, want), s) in enumerate(zip(checks, st)):
            assert s == want, (rep, i)
            assert np.array_equal(buf.cpu().numpy(), exp), (rep, i, specs[i]["op"], specs[i]["n"])
