"""The committed fixtures (tests/golden/fixtures.json, written and checked
against closed forms by tests/golden/make_fixtures.py) still come out of the
oracle byte for byte: a regression pin on the checker itself."""
import json
import os

import numpy as np

import oracle as O

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))


def b(h):
    return np.frombuffer(bytes.fromhex(h), np.uint8).copy()


def test_encode_fixtures():
    for r in FIX["encode"]:
        assert O.encode_any(r["n"], r["k"], b(r["message"]), r["shard_len"]).tobytes().hex() == r["parity"]


def test_recover_fixtures():
    for r in FIX["recover"]:
        got = b(r["received"])
        assert O.recover_any(r["n"], r["k"], got, r["shard_len"], r["positions"]) == 1
        assert got.tobytes().hex() == r["recovered"]


def test_decode_fixtures():
    for r in FIX["decode"]:
        n, T, L = r["n"], r["n"] - r["k"], r["shard_len"]
        if T & (T - 1):
            continue  # error decoding follows the reference: T a power of two
        for mode in ("reference", "formal"):
            got = b(r["received"])
            rc, st = O.decode_sharded(n, T, got, L, formal=mode == "formal")
            assert rc == r[mode]["ok"] and list(st) == r[mode]["status"]
            assert got.tobytes().hex() == r[mode]["decoded"]
