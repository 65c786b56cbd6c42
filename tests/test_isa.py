"""Static checks on the gfx950 assembly of the kernels (CPU only: hipcc
cross-compiles).  They guard against code shapes that cost more than the
arithmetic, each of which has happened in this code base:
  * a buffer load or store wrapped in a readfirstlane / exec loop, which the
    compiler emits when a buffer descriptor sits in VGPRs;
  * function calls in a kernel (the descriptors cross the call in VGPRs and
    the callee saves its callee-saved VGPRs to scratch on every call);
  * flat loads / stores in the one-codeword tile kernels, i.e. LDS
    addressed through generic pointers;
  * scratch in the kernels the library launches by default."""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import _pkg

CSRC = os.path.join(_pkg.PKG_DIR, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-S", "--cuda-device-only"]
SOURCES = ["ars_kernels.hip", "ars_batch.hip", "ars_decode.hip", "ars_host.cpp"]


def functions(asm):
    """{symbol: list of lines} for every function in a device assembly file."""
    lines = asm.split("\n")
    starts = [i for i, l in enumerate(lines) if re.match(r"^_Z[A-Za-z0-9_]+:", l)]
    starts.append(len(lines))
    return {lines[a].split(":")[0]: lines[a:b] for a, b in zip(starts, starts[1:])}


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc missing")
    out = tmp_path_factory.mktemp("isa")

    def one(src):
        dst = os.path.join(out, src + ".s")
        cmd = [HIPCC] + FLAGS + (["-x", "hip"] if src.endswith(".cpp") else []) + \
            [os.path.join(CSRC, src), "-o", dst]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        with open(dst) as f:
            return src, functions(f.read())

    with ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(one, SOURCES))


def waterfalls(body):
    return sum(1 for i, l in enumerate(body[:-1])
               if "s_and_saveexec_b64" in l and re.search(r"buffer_(load|store)", body[i + 1]))


def test_no_readfirstlane_loops_around_buffer_access(asm):
    bad = {f"{src}:{name[:80]}": n for src, fns in asm.items() for name, body in fns.items()
           if (n := waterfalls(body))}
    assert not bad, bad


def test_no_calls_in_kernels(asm):
    bad = [f"{src}:{name[:80]}" for src, fns in asm.items() for name, body in fns.items()
           if any("s_swappc_b64" in l for l in body)]
    assert not bad, bad


def test_tile_kernels_address_lds_directly(asm):
    # (the batched kernels read their arena descriptors, which are global,
    # through generic pointers; flat loads there are not LDS accesses)
    bad = [name[:80] for name, body in asm["ars_kernels.hip"].items()
           if "tile" in name and any(re.search(r"\bflat_(load|store)", l) for l in body)]
    assert not bad, bad


def test_default_kernels_do_not_spill(asm):
    # the variants the launchers pick without tuning: PQ 64 / NBUF 1 for
    # encode (all four access modes); recover: the streaming kernel with
    # 32 quads at 256 points and 64 below (either load form; 64 quads for wide
    # rows), and on the byte path PQ 32 / NBUF 2 at 256 points, PQ 64 / NBUF 2
    # below; the one-launch recovers use the same variants.
    # A spill matters doubly in the streaming kernel: a scratch reload waits for
    # every memory operation issued before it, the prefetched rows included.
    # (encode and the streaming recover in access modes 1 and 2: aligned dwords
    # and dwords shifted past alignment)
    want = [r"k_encode_tileILi7ELi64ELi[0123]ELb[01]ELi1ELb[01]E", r"k_encode_tileILi[56]ELi64ELi[0123]ELb[01]ELi1E",
            r"k_recover_streamILi8ELi64ELb1ELi[12]E", r"k_recover_streamILi8ELi64ELb0ELi[12]E",
            r"k_recover_streamILi8ELi32ELb1ELi[12]E", r"k_recover_streamILi8ELi32ELb0ELi[12]E",
            r"k_recover_stream_fusedILi8ELi32ELb1ELi[12]E", r"k_recover_streamILi8ELi64ELb[01]ELi3E",
            r"k_recover_streamILi7ELi64ELb1ELi[12]E", r"k_recover_streamILi5ELi64ELb1ELi[12]E",
            r"k_recover_tileILi8ELi32ELb0ELi2E",
            r"k_recover_tile_fusedILi8ELi32ELb0ELi2E", r"k_recover_tile_fusedILi5ELi64ELb0ELi2E",
            r"k_recover_small_fusedILi4ELi[012]E", r"k_syndrome_tile", r"k_decode_columns"]
    seen = {}
    unmatched = []
    for w in want:
        hits = {name[:90]: body for fns in asm.values() for name, body in fns.items() if re.search(w, name)}
        if not hits:
            unmatched.append(w)
        for name, body in hits.items():
            m = re.search(r"; ScratchSize: (\d+)", "\n".join(body))
            seen[name] = int(m.group(1)) if m else 0
    assert not unmatched, f"no kernel matches {unmatched}"
    assert all(v == 0 for v in seen.values()), {k: v for k, v in seen.items() if v}
