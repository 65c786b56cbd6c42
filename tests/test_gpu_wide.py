"""Codewords of 4 GiB and more (n * shard_len >= 2^32): access mode 3, where
each row is addressed through a resource based at that row, so no byte offset
passes 2^32.

Columns are independent, so every wide result is checked column-slice by
column-slice: a slice of w columns of every shard is a codeword of its own
(shard_len w), which the oracle encodes on the CPU and the ordinary (narrow)
GPU path encodes and recovers.  Slices sit at the first columns, across the
middle and at the last columns, so rows whose byte offset passes 2^32 are
covered.  Recover is checked on the whole codeword (zeroed erased rows must
come back) and on garbage-filled rows through the slices (result c ^ g)."""
import numpy as np
import pytest

import _pkg
import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MiB = 1 << 20
W = 4099  # slice width: odd, so the narrow path takes its byte form


@pytest.fixture(scope="module")
def ars():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _pkg.load()


def _slices(L):
    return (0, L // 2 - W // 2, L - W)


def _rand(nbytes, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randint(0, 256, (nbytes,), dtype=torch.uint8, device="cuda", generator=g)


def _cols(buf, rows, L, c0, w=W):
    """columns [c0, c0 + w) of every row of a (rows x L) device buffer, packed."""
    return buf.view(rows, L)[:, c0:c0 + w].contiguous()


def _erase(cw, n, L, pos, fill=None):
    idx = torch.from_numpy(pos.astype(np.int64)).cuda()
    v = cw.view(n, L)
    v[idx] = 0 if fill is None else fill


# (n, k, shard_len): T = n - k a power of two (the tile and the small encode,
# one chunk and many), k a power of two with T not (three parity blocks),
# the smallest code, and a 32-point code with a long shard.
CASES = [(256, 128, 24 * MiB), (256, 252, 17 * MiB), (256, 64, 17 * MiB), (32, 16, 136 * MiB),
         (2, 1, 2048 * MiB + 4096)]


@pytest.mark.parametrize("n,k,L", CASES)
def test_wide_encode_and_recover(ars, n, k, L):
    T = n - k
    assert n * L >= 1 << 32 and L % 4 == 0
    print(f"\nwide encode/recover n={n} k={k} L={L}", flush=True)
    ctx = ars.Context(0)
    cw = torch.empty(n * L, dtype=torch.uint8, device="cuda")
    cw[T * L:] = _rand(k * L, n * 7 + k)
    assert ctx.encode(n, k, cw.data_ptr() + T * L, cw.data_ptr(), L) == 0
    torch.cuda.synchronize()
    for c0 in _slices(L):
        sl = _cols(cw, n, L, c0)
        msg = sl[T:].cpu().numpy().reshape(-1)
        got = sl[:T].cpu().numpy().reshape(-1)
        assert np.array_equal(got, O.encode_any(n, k, msg, W)), ("oracle", c0)
        # the narrow GPU path on the same slice
        par = torch.zeros(T * W, dtype=torch.uint8, device="cuda")
        m = sl[T:].reshape(-1).contiguous()
        assert ctx.encode(n, k, m.data_ptr(), par.data_ptr(), W) == 0
        torch.cuda.synchronize()
        assert np.array_equal(got, par.cpu().numpy()), ("narrow", c0)
    # recover T zeroed rows (the first and last rows among them): the whole codeword comes back
    rng = np.random.default_rng(n + L)
    pos = np.array(([0, n - 1] + [int(p) for p in rng.permutation(n) if p not in (0, n - 1)])[:T], np.uint8)
    ref = cw.clone()
    _erase(cw, n, L, pos)
    assert ctx.recover(n, k, cw.data_ptr(), L, pos) == 1
    torch.cuda.synchronize()
    assert torch.equal(cw, ref)
    # garbage in the erased rows: every slice equals the narrow recover of that slice
    _erase(cw, n, L, pos, _rand(len(pos) * L, L).view(len(pos), L))
    rx = {c0: _cols(cw, n, L, c0) for c0 in _slices(L)}
    assert ctx.recover(n, k, cw.data_ptr(), L, pos) == 1
    for c0, s in rx.items():
        d = s.reshape(-1).contiguous()
        assert ctx.recover(n, k, d.data_ptr(), W, pos) == 1
        torch.cuda.synchronize()
        assert torch.equal(_cols(cw, n, L, c0).reshape(-1), d), ("garbage", c0)
    del cw, ref, rx
    torch.cuda.empty_cache()
    ctx.close()


def test_wide_accumulate_plan_noalloc_and_recompute(ars):
    """(256,128), 6 GiB: ARS_ENC_ACCUMULATE, ars_recover_with_plan,
    ars_recover_noalloc and ars_recompute_data on a wide codeword against the
    narrow path and the oracle on column slices."""
    n, k, L = 256, 128, 24 * MiB
    T = n - k
    print(f"\nwide accumulate/plan/noalloc/recompute n={n} k={k} L={L}", flush=True)
    ctx = ars.Context(0)
    msg = _rand(k * L, 901)
    par = _rand(T * L, 902)
    p0 = {c0: _cols(par, T, L, c0) for c0 in _slices(L)}
    assert ctx.encode(n, k, msg.data_ptr(), par.data_ptr(), L, flags=1) == 0
    torch.cuda.synchronize()
    for c0, s in p0.items():
        m = _cols(msg, k, L, c0).reshape(-1).contiguous()
        exp = O.encode(n, T, m.cpu().numpy(), W, parity=s.reshape(-1).cpu().numpy().copy())
        assert np.array_equal(_cols(par, T, L, c0).cpu().numpy().reshape(-1), exp), ("accumulate", c0)
    # a codeword: parity of the message, overwritten
    cw = torch.empty(n * L, dtype=torch.uint8, device="cuda")
    cw[T * L:] = msg
    del msg, par, p0
    assert ctx.encode(n, k, cw.data_ptr() + T * L, cw.data_ptr(), L) == 0
    ref = cw.clone()
    rng = np.random.default_rng(903)
    pos = rng.permutation(n)[:T].astype(np.uint8)
    plan = torch.zeros(ars.lib().ars_recover_plan_size(), dtype=torch.uint8, device="cuda")
    assert ctx.recover_plan(n, k, pos, plan.data_ptr()) == 1
    _erase(cw, n, L, pos)
    assert ctx.recover_with_plan(n, k, cw.data_ptr(), L, plan.data_ptr()) == 1
    torch.cuda.synchronize()
    assert torch.equal(cw, ref), "recover_with_plan"
    pos2 = rng.permutation(n)[:T - 5].astype(np.uint8)
    _erase(cw, n, L, pos2)
    assert ctx.recover_noalloc(n, k, cw.data_ptr(), L, pos2) == 1
    torch.cuda.synchronize()
    assert torch.equal(cw, ref), "recover_noalloc"
    # recompute_data (n = 2T): the message rows are rebuilt from the parity rows
    cw[T * L:] = 0
    assert ctx.recompute_data(n, k, cw.data_ptr(), L) == 0
    torch.cuda.synchronize()
    assert torch.equal(cw, ref), "recompute_data"
    del cw, ref
    torch.cuda.empty_cache()
    ctx.close()


def test_wide_host_buffers(ars):
    """The host-buffer pipeline stages column chunks, so codewords of 4 GiB and
    more in host memory work on any alignment: (32,16) with 136 MiB + 1 shards."""
    n, k, L = 32, 16, 136 * MiB + 1
    T = n - k
    print(f"\nwide host encode/recover n={n} k={k} L={L}", flush=True)
    codec = ars.Codec(n, k)
    rng = np.random.default_rng(904)
    cw = np.empty(n * L, np.uint8)
    cw[T * L:] = rng.integers(0, 256, k * L, dtype=np.uint8)
    cw[:T * L] = 0
    codec.encode_systematic_sharded(cw[T * L:], cw[:T * L], None, L, accumulate=False)
    rows = cw.reshape(n, L)
    for c0 in _slices(L):
        sl = np.ascontiguousarray(rows[:, c0:c0 + W])
        assert np.array_equal(sl[:T].reshape(-1), O.encode(n, T, sl[T:].reshape(-1), W)), c0
    pos = np.array([0, 3, 17, 31], np.uint8)
    keep = rows[pos].copy()
    rows[pos] = 0
    assert codec.recover_erasure_shards(cw, None, L, pos)
    assert np.array_equal(rows[pos], keep)


def test_wide_rejections(ars):
    """A wide codeword needs 4-byte aligned device buffers and shard_len % 4
    == 0; calls without the wide form (decode, batches, serving) keep the
    4 GiB limit."""
    n, k, L = 256, 128, 16 * MiB + 4
    ctx = ars.Context(0)
    cw = torch.empty(n * L + 8, dtype=torch.uint8, device="cuda")
    base = cw.data_ptr()
    assert ars.lib().ars_recover(ctx.handle, n, k, base + 1, L, None, 0, None) == ars.ARS_E_UNSUPPORTED
    assert ars.lib().ars_encode(ctx.handle, n, k, base + 1 + 128 * L, base + 1, L, 0, None) == ars.ARS_E_UNSUPPORTED
    assert ars.lib().ars_encode(ctx.handle, n, k, base + 128 * (L - 2), base, L - 2, 0, None) == ars.ARS_E_UNSUPPORTED
    assert ars.lib().ars_decode(ctx.handle, n, k, base, L, 0, None, None, None) == ars.ARS_E_INVAL
    assert "4 GiB" in ars.lib().ars_last_error().decode()
    del cw
    torch.cuda.empty_cache()
    ctx.close()
