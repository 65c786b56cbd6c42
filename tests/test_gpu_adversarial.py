"""Adversarial shapes and erasure patterns: (256,224) and (256,16) (T = 240,
outside the reference's contract, pinned by Lagrange) with all-parity,
all-leading-data, contiguous-run and alternating erasures; the oracle
comparison runs at small shard lengths, the full 4 MiB size checks the
round trip (uniqueness) on the device."""
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


def patterns(n, k, rng):
    T = n - k
    out = {
        "all_parity": list(range(T)),
        "leading_data": list(range(T, T + min(T, k))),
        "alternating": list(range(1, min(2 * T, n), 2))[:T],
    }
    runs = []
    run = max(1, T // 4)
    starts = rng.permutation(n // run)[:4] * run
    for s0 in starts:
        runs.extend(range(s0, min(n, s0 + run)))
    out["runs"] = sorted(set(runs))[:T]
    return out


@pytest.mark.parametrize("n,k", [(256, 224), (256, 16), (256, 128), (64, 32)])
def test_patterns_match_oracle(ars, n, k):
    rng = np.random.default_rng(46 + k)
    L = 192
    T = n - k
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    cw = np.concatenate([O.encode_any(n, k, msg, L), msg])
    codec = ars.Codec(n, k)
    dm = torch.from_numpy(msg).cuda()
    dp = torch.zeros(T * L, dtype=torch.uint8, device="cuda")
    codec.encode_systematic_sharded(dm, dp, None, L, accumulate=False)
    assert np.array_equal(dp.cpu().numpy(), cw[:T * L])
    for name, pos in patterns(n, k, rng).items():
        rx = cw.copy().reshape(n, L)
        rx[pos] = 0
        d = torch.from_numpy(rx.reshape(-1).copy()).cuda()
        assert codec.recover_erasure_shards(d, None, L, pos), name
        assert np.array_equal(d.cpu().numpy(), cw), (n, k, name)


@pytest.mark.parametrize("n,k", [(256, 224), (256, 16)])
def test_full_size_4mib_patterns_round_trip(ars, n, k):
    L = 4 << 20
    T = n - k
    g = torch.Generator(device="cuda")
    g.manual_seed(46)
    cw = torch.empty(n * L, dtype=torch.uint8, device="cuda")
    cw[T * L:] = torch.randint(0, 256, (k * L,), dtype=torch.uint8, device="cuda", generator=g)
    codec = ars.Codec(n, k)
    codec.encode_systematic_sharded(cw[T * L:], cw[:T * L], None, L, accumulate=False)
    ref = cw.clone()
    rng = np.random.default_rng(7)
    for name, pos in patterns(n, k, rng).items():
        v = cw.view(n, L)
        v[torch.tensor(pos, dtype=torch.int64, device="cuda")] = 0
        assert codec.recover_erasure_shards(cw, None, L, pos)
        torch.cuda.synchronize()
        assert torch.equal(cw, ref), (n, k, name)
    del ref, cw
    torch.cuda.empty_cache()


def test_65536_tiny_codewords_batched_equals_plain(ars):
    """65536 (32,16) codewords with 64-byte shards: encode and recover of
    random e <= 16 in one batch equals one call per codeword."""
    count, n, k, L = 65536, 32, 16, 64
    T = n - k
    rng = np.random.default_rng(46)
    g = torch.Generator(device="cuda")
    g.manual_seed(47)
    pool = torch.randint(0, 256, (count * n * L,), dtype=torch.uint8, device="cuda", generator=g)
    cws = pool.view(count, n * L)
    enc = [{"op": "encode", "n": n, "k": k, "shard_len": L, "message": cws[i, T * L:], "parity": cws[i, :T * L]}
           for i in range(count)]
    b = ars.Batch(count)
    st = b.run(enc)
    assert all(s == 0 for s in st)
    torch.cuda.synchronize()
    # plain calls on a copy for a sample of codewords
    sample = rng.choice(count, 64, replace=False)
    for i in sample:
        msg = cws[i, T * L:].cpu().numpy()
        assert np.array_equal(cws[i, :T * L].cpu().numpy(), O.encode(n, T, msg, L))
    ref = cws.clone()
    es = rng.integers(0, T + 1, count)
    rec = []
    for i in range(count):
        pos = rng.permutation(n)[:es[i]]
        rec.append({"op": "recover", "n": n, "k": k, "shard_len": L, "codeword": cws[i], "positions": pos})
        if len(pos):
            cws[i].view(n, L)[torch.from_numpy(pos.astype(np.int64)).cuda()] = 0
    st = b.run(rec)
    torch.cuda.synchronize()
    assert all(s == 1 for s in st)
    assert torch.equal(cws, ref)
    assert b.info()["items"] < count // 8


@pytest.mark.parametrize("n,k", [(256, 128), (128, 64)])
def test_whole_wave_patterns_round_trip(ars, n, k):
    # erasure patterns where whole waves of the streaming recover hold only
    # erased rows, or none (a wave covers 32 rows at n = 256, 16 below): those
    # waves skip their phase-A IFFT or FFT; 256 KiB shards run the streaming
    # kernel with more than one workgroup per CU, the garbage-free round trip
    # must give back the codeword
    L = 256 << 10
    T = n - k
    w = 32 if n == 256 else 16
    g = torch.Generator(device="cuda")
    g.manual_seed(47)
    cw = torch.empty(n * L, dtype=torch.uint8, device="cuda")
    cw[T * L:] = torch.randint(0, 256, (k * L,), dtype=torch.uint8, device="cuda", generator=g)
    codec = ars.Codec(n, k)
    codec.encode_systematic_sharded(cw[T * L:], cw[:T * L], None, L, accumulate=False)
    ref = cw.clone()
    pats = {
        "first_half": list(range(T)),
        "alternate_waves": [r for r in range(n) if (r // w) % 2 == 0][:T],
        "waves_plus_partial": (list(range(0, w)) + list(range(3 * w, 3 * w + w // 2)) +
                               list(range(n - w // 2, n)))[:T],
    }
    for name, pos in pats.items():
        v = cw.view(n, L)
        v[torch.tensor(pos, dtype=torch.int64, device="cuda")] = 0
        assert codec.recover_erasure_shards(cw, None, L, pos)
        torch.cuda.synchronize()
        assert torch.equal(cw, ref), (n, k, name)
    del ref, cw
    torch.cuda.empty_cache()
