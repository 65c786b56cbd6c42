"""Autotuner: every tile variant it can choose is bit-exact, and tuning keeps
results unchanged."""
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


@pytest.mark.parametrize("quads,buffers", [(16, 1), (16, 2), (32, 1), (32, 2), (64, 1), (64, 2)])
def test_every_variant_is_bit_exact(ars, quads, buffers):
    ctx = ars.Context(0)
    rng = np.random.default_rng(quads + buffers)
    # small shards, then one large enough for thousands of resident
    # workgroups (the exchange barriers are exercised under full occupancy)
    for n, T, L in [(64, 32, 640), (128, 64, 640), (256, 128, 640), (256, 32, 640), (128, 32, 640),
                    (256, 128, 65536)]:
        k = n - T
        tl, nl = T.bit_length() - 1, n.bit_length() - 1
        if tl >= 5:
            ctx.tune_set(ars.ARS_OP_ENCODE, tl, quads, buffers)
        ctx.tune_set(ars.ARS_OP_RECOVER, nl, quads, buffers)
        msg = rng.integers(0, 256, k * L, dtype=np.uint8)
        exp = O.encode(n, T, msg, L)
        dm = torch.from_numpy(msg).cuda()
        dp = torch.zeros(T * L, dtype=torch.uint8, device="cuda")
        ctx.encode(n, k, dm.data_ptr(), dp.data_ptr(), L)
        torch.cuda.synchronize()
        assert np.array_equal(dp.cpu().numpy(), exp), (n, T, quads, buffers)
        cw = np.concatenate([exp, msg])
        pos = rng.permutation(n)[:T].astype(np.uint8)
        rx = cw.copy().reshape(n, L)
        rx[pos] = 0
        d = torch.from_numpy(rx.reshape(-1).copy()).cuda()
        assert ctx.recover(n, k, d.data_ptr(), L, pos) == 1
        torch.cuda.synchronize()
        assert np.array_equal(d.cpu().numpy(), cw), (n, T, quads, buffers)
    ctx.close()


def test_autotune_picks_and_stays_exact(ars):
    ctx = ars.Context(0)
    enc, rec = ctx.autotune(256, 128, 1 << 16)
    assert enc["quads"] in (16, 32, 64) and rec["quads"] in (16, 32, 64)
    assert enc["best_us"] > 0 and enc["best_us"] <= enc["default_us"] + 1e-9
    rng = np.random.default_rng(9)
    n, k, L = 256, 128, 4096
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    dm = torch.from_numpy(msg).cuda()
    dp = torch.zeros((n - k) * L, dtype=torch.uint8, device="cuda")
    ctx.encode(n, k, dm.data_ptr(), dp.data_ptr(), L)
    torch.cuda.synchronize()
    assert np.array_equal(dp.cpu().numpy(), O.encode(n, n - k, msg, L))
    with pytest.raises(ars.ArsError):
        ctx.tune_set(ars.ARS_OP_RECOVER, 8, 48, 1)
    ctx.close()
