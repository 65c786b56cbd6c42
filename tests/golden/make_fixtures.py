"""Writes tests/golden/fixtures.json: small seeded input/output vectors of the
hot path, produced by the oracle and checked here, as they are written,
against what the mathematics fixes:
  * recover returns the transmitted codeword (uniqueness, e <= T);
  * decode with the formal derivative returns it for columns with <= T/2
    errors (bounded-distance uniqueness);
  * encode output satisfies the erasure round trip above.
They are regression pins for the oracle (tests/test_golden.py) and direct
pins for the GPU path (tests/test_gpu_golden.py).

    python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def hx(a):
    return np.asarray(a, np.uint8).tobytes().hex()


def codeword(n, T, L, rng):
    msg = rng.integers(0, 256, (n - T) * L, dtype=np.uint8)
    return msg, np.concatenate([O.encode_any(n, n - T, msg, L), msg])


def main():
    out = {"encode": [], "recover": [], "decode": []}
    rng = np.random.default_rng(2024)
    for n, T, L in [(4, 2, 3), (16, 8, 5), (64, 32, 8), (256, 128, 4), (256, 32, 4), (256, 240, 2), (32, 16, 1)]:
        msg, cw = codeword(n, T, L, rng)
        out["encode"].append({"n": n, "k": n - T, "shard_len": L, "message": hx(msg), "parity": hx(cw[:T * L])})
        for e in sorted({1, T // 2, T}):
            pos = np.sort(rng.permutation(n)[:e]).astype(np.uint8)
            rx = cw.copy().reshape(n, L)
            rx[pos] = 0
            got = rx.reshape(-1).copy()
            assert O.recover_any(n, n - T, got, L, pos) == 1
            assert np.array_equal(got, cw), (n, T, e)  # uniqueness
            out["recover"].append({"n": n, "k": n - T, "shard_len": L, "positions": [int(p) for p in pos],
                                   "received": hx(rx), "recovered": hx(got)})
    for n, T, L in [(8, 4, 6), (16, 8, 6), (64, 32, 6), (256, 128, 3), (256, 16, 4)]:
        _, cw = codeword(n, T, L, rng)
        rx = cw.copy().reshape(n, L)
        errs = []
        for j in range(L):
            ne = int(rng.integers(0, T + 2))
            rx[rng.permutation(n)[:ne], j] ^= rng.integers(1, 256, ne, dtype=np.uint8)
            errs.append(ne)
        rec = {"n": n, "k": n - T, "shard_len": L, "received": hx(rx)}
        for mode, formal in (("reference", False), ("formal", True)):
            got = rx.reshape(-1).copy()
            rc, st = O.decode_sharded(n, T, got, L, formal=formal)
            assert rc in (0, 1)
            if formal:
                for j, ne in enumerate(errs):
                    if ne <= T // 2:
                        assert np.array_equal(got.reshape(n, L)[:, j], cw.reshape(n, L)[:, j]), (n, T, j)
            rec[mode] = {"ok": int(rc), "status": [int(v) for v in st], "decoded": hx(got)}
        out["decode"].append(rec)
    path = os.path.join(HERE, "fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}: {sum(len(v) for v in out.values())} vectors")


if __name__ == "__main__":
    main()
