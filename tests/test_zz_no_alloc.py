"""ars_recover_noalloc (the allocation-free recover) against ars_recover,
through the C ABI: byte-equal outputs on consecutive calls that reuse the
library's per-GPU plan scratch, the second and third with smaller inputs."""
import numpy as np
import pytest

import _pkg

pytestmark = pytest.mark.gpu


def _codeword(torch, ars, ctx, n, k, L, rng):
    msg = torch.from_numpy(rng.integers(0, 256, k * L, dtype=np.uint8)).cuda()
    cw = torch.zeros(n * L, dtype=torch.uint8, device="cuda")
    ctx.encode(n, k, msg.data_ptr(), cw.data_ptr(), L)
    cw[(n - k) * L:] = msg
    return cw


def test_noalloc_matches_recover_on_consecutive_calls():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ars = _pkg.load()
    ctx = ars.Context(0)
    rng = np.random.default_rng(91)
    # (n, k, L, e): the first two launch more workgroups than compute units,
    # so their plans go through the scratch; the second is smaller than the
    # first (fewer shards and a shorter shard) with a different pattern; the
    # third is small enough for the one-launch path, which needs no plan buffer
    for n, k, L, e in ((256, 128, 1 << 20, 128), (128, 64, 300 * 1024, 40), (64, 32, 1024, 32)):
        cw = _codeword(torch, ars, ctx, n, k, L, rng)
        pos = rng.permutation(n)[:e].astype(np.uint8)
        rx = cw.clone().view(n, L)
        # garbage in the erased shards: both calls must XOR it into the result
        rx[torch.from_numpy(pos.astype(np.int64)).cuda()] = torch.from_numpy(
            rng.integers(0, 256, (e, L), dtype=np.uint8)).cuda()
        rx = rx.reshape(-1)
        a, b = rx.clone(), rx.clone()
        assert ctx.recover(n, k, a.data_ptr(), L, pos) == 1
        assert ctx.recover_noalloc(n, k, b.data_ptr(), L, pos) == 1
        torch.cuda.synchronize()
        assert torch.equal(a, b), (n, k, L, e)
        # and zero-filled erasures come back as the codeword
        z = cw.clone().view(n, L)
        z[torch.from_numpy(pos.astype(np.int64)).cuda()] = 0
        z = z.reshape(-1)
        assert ctx.recover_noalloc(n, k, z.data_ptr(), L, pos) == 1
        torch.cuda.synchronize()
        assert torch.equal(z, cw), (n, k, L, e)
    # argument checks and the reference's e > T answer, as ars_recover
    buf = torch.zeros(64 * 64, dtype=torch.uint8, device="cuda")
    assert ctx.recover_noalloc(64, 32, buf.data_ptr(), 64, np.arange(33, dtype=np.uint8)) == 0
    with pytest.raises(ars.ArsError):
        ctx.recover_noalloc(64, 32, buf.data_ptr(), 64, np.array([1, 1], np.uint8))
    with pytest.raises(ars.ArsError):
        ctx.recover_noalloc(512, 256, buf.data_ptr(), 8, np.array([1], np.uint8))
    ctx.close()
