"""Benchmark: one (n=256, k=128) codeword with 1 MiB shards on one MI355X.

A step is one encode (128 parity shards from 128 message shards) followed by
one recover of 128 uniformly random erasures (seeded), both through the C ABI
on device-resident buffers.  Throughput counts algorithmic bytes:
encode n*L (k*L read + T*L written), recover (n+e)*L (n*L read + e*L written).

Prints ONE JSON line (rank 0).  Usage:
    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBPS = 8000.0
METRIC = "encode/recover GB/s vs HBM roofline; batched p50/p99 latency vs plain calls"


def load_pkg():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _pkg
    return _pkg.load()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_info():
    """Host CPU as /proc/cpuinfo and the scheduler report it: model string,
    logical CPUs online, CPUs this process may run on, and the flags the
    restatement needs."""
    flags = set()
    ident = {}
    try:
        for line in open("/proc/cpuinfo"):
            key = line.split(":", 1)[0].strip()
            if key == "flags":
                flags = set(line.split(":", 1)[1].split())
            elif key in ("cpu family", "model", "stepping", "cache size", "cpu MHz") and key not in ident:
                ident[key] = line.split(":", 1)[1].strip()
            if line.strip() == "":
                break  # the first processor's block is enough
    except OSError:
        pass
    # a virtualised host may report only "AMD EPYC" as its model string; the
    # family / model / stepping numbers still identify the generation
    return {"model": cpu_model(), "cpuid": {k.replace(" ", "_"): v for k, v in ident.items()},
            "nproc_online": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0)),
            "avx512bw": "avx512bw" in flags, "gfni": "gfni" in flags}


def cpu_baseline(n, k, L, e, seconds):
    """Oracle restatement of the reference's CPU path on the same (n, k, L)
    codeword: encode + recover steps until `seconds` have passed, on one core
    (the reference is single-threaded) and then on every CPU this process may
    use, one pinned thread and one independent codeword each.  Also times the
    reference README's (64, 32) rows (README.md:29-35, AMD EPYC 9575F) for a
    sanity check, through the reference's beta = 0 fixed sequences (its
    default dispatch) and through the generic recursion."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    T = n - k
    gfni = O.gfni_supported()
    step_bytes = n * L + (n + e) * L

    def worker(seed, budget, cpu=None):
        if cpu is not None:
            os.sched_setaffinity(0, {cpu})  # this thread only (Linux: 0 = the caller)
        rng = np.random.default_rng(seed)
        msg = rng.integers(0, 256, k * L, dtype=np.uint8)
        pos = np.sort(rng.permutation(n)[:e]).astype(np.uint8)
        # buffers allocated and touched once, outside the timed calls, as the
        # reference's benches do (rs/benches/*.rs iter_batched set-up)
        ws = np.ones(n * L, np.uint8)
        par = np.ones(T * L, np.uint8)
        rx = np.ones(n * L, np.uint8)
        steps, t_enc, t_rec = 0, 0.0, 0.0
        t_end = time.perf_counter() + budget
        while steps == 0 or time.perf_counter() < t_end:
            par[:] = 0
            t0 = time.perf_counter()
            O.encode(n, T, msg, L, parity=par, gfni=gfni, workspace=ws)
            t1 = time.perf_counter()
            rx2 = rx.reshape(n, L)
            rx2[:T] = par.reshape(T, L)
            rx2[T:] = msg.reshape(k, L)
            rx2[pos] = 0
            t2 = time.perf_counter()
            O.recover(n, T, rx, L, pos, gfni=gfni, workspace=ws)
            t3 = time.perf_counter()
            t_enc += t1 - t0
            t_rec += t3 - t2
            steps += 1
        return steps, t_enc, t_rec

    cpus = sorted(os.sched_getaffinity(0))
    main_aff = set(cpus)
    steps, t_enc, t_rec = worker(7, seconds, cpus[0])
    os.sched_setaffinity(0, main_aff)
    threads = len(cpus)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:  # ctypes calls release the GIL
        res = list(ex.map(lambda i: worker(100 + i, seconds / 2, cpus[i]), range(threads)))
    wall = time.perf_counter() - t0
    all_steps = sum(r[0] for r in res)
    readme = {}
    for Lr, enc_ref, rec_ref in ((64, "385ns, 10GiB/s", "7.7us, 510MiB/s"), (1024, "2.4us, 25GiB/s", "14us, 4.1GiB/s"),
                                 (65536, "192us, 20GiB/s", "740us, 5.3GiB/s")):
        pos = np.sort(np.random.default_rng(Lr).permutation(64)[:32]).astype(np.uint8)
        row = {"readme_encode": enc_ref, "readme_recover": rec_ref}
        for form, unrolled in (("", True), ("generic_", False)):
            O.gfni_set_unrolled(unrolled)
            t = O.gfni_time_codec(64, 32, Lr, pos, max(5, int(4e7 / (64 * Lr))))
            if t is None:
                break
            me, mr = t[0] * 1e-9, t[1] * 1e-9
            row.update({f"{form}encode_us": round(me * 1e6, 2), f"{form}encode_GiBps": round(64 * Lr / me / 2**30, 2),
                        f"{form}recover_us": round(mr * 1e6, 2),
                        f"{form}recover_GiBps": round(64 * Lr / mr / 2**30, 2)})
        O.gfni_set_unrolled(True)
        readme[str(Lr)] = row
    return {
        "value": round(steps * step_bytes / (t_enc + t_rec) / 1e9, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{steps} encode+recover steps of the same (n={n},k={k}) codeword, L={L}, "
                   f"e={e}; {'AVX-512BW+GFNI' if gfni else 'LUT'} restatement of "
                   f"rs/src/kernel/gfni_kernel.rs + rs/src/codec.rs on {cpu_model()}"),
        "encode_GBps": round(steps * n * L / t_enc / 1e9, 3),
        "recover_GBps": round(steps * (n + e) * L / t_rec / 1e9, 3),
        "recover_GBps_nL_convention": round(steps * n * L / t_rec / 1e9, 3),
        "host": cpu_info(),
        "all_cores": {"threads": threads, "pinned": True, "steps": all_steps,
                      "GBps": round(all_steps * step_bytes / wall / 1e9, 2),
                      "sample": "one independent codeword per usable CPU, each thread pinned to its CPU, same shape"},
        "readme_64x32_check": readme,
    }


def device_source_sha():
    """SHA-256 over the device sources (csrc/*.hip and *.hpp, sorted by name):
    a committed profile applies only to kernels built from the same text."""
    import hashlib
    csrc = os.path.join(ROOT, "additive-fft-reed-solomon_amd", "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(csrc)):
        if name.endswith((".hip", ".hpp")):
            h.update(name.encode())
            with open(os.path.join(csrc, name), "rb") as f:
                h.update(f.read())
    return h.hexdigest()


def _profile_entry(fname, kernel):
    """The committed rocprofv3 record (profiles/<fname>) of exactly `kernel`,
    provided the record was made from the device sources this run uses
    (device_source_sha); (name, record) or (None, None)."""
    try:
        with open(os.path.join(ROOT, "profiles", fname)) as f:
            doc = json.load(f)
        recs = doc["kernels"]
    except (OSError, ValueError, KeyError):
        return None, None
    if kernel not in recs or doc.get("device_source_sha") != device_source_sha():
        return None, None
    return kernel, recs[kernel]


def _profiled_run_launch_us(fname, kernel):
    """The per-launch time (HIP events) that the profiled run's own bench line
    measured for `kernel`, on the box the profile comes from; None unless the
    profile holds `kernel` and matches this run's device sources."""
    if _profile_entry(fname, kernel)[0] is None:
        return None
    with open(os.path.join(ROOT, "profiles", fname)) as f:
        line = json.load(f).get("bench_line") or {}
    return line.get("launch_us") if line.get("kernel") == kernel else None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (profiles/traffic_256x128_1MiB.json), or None."""
    _, rec = _profile_entry("traffic_256x128_1MiB.json", kernel)
    return rec["traffic_bytes"] if rec else None


def probe_valu_rates():
    """Live VALU issue rates (wave64 instructions / s) from libars_probe.so:
    independent v_xor_b32 chains, v_perm_b32 chains, the fused kernels'
    butterfly in a small loop on registers, and the same butterflies in
    straight-line code the size of a tile body.  None if the probe library
    is missing."""
    import ctypes
    path = os.path.join(ROOT, "additive-fft-reed-solomon_amd", "libars_probe.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.ars_probe_valu_rates.restype = ctypes.c_int
    lib.ars_probe_valu_rates.argtypes = [ctypes.POINTER(ctypes.c_double)]
    out = (ctypes.c_double * 5)()
    if lib.ars_probe_valu_rates(out) != 0:
        return None
    return {"simple": out[0], "perm": out[1], "butterfly_loop": out[2], "butterfly_straight_line": out[3],
            "butterfly_burst": out[4]}


def bench_phaseb(torch):
    """The bit-sliced phase-B prototype against the product's v_perm form
    (libars_probe.so, ars_probe_phaseb): 16 rows of 4 MiB, a few maps of phase
    B (IFFT levels 4..7, high-bit derivative terms, FFT levels 7..4 of the
    256-point recover) per launch on register-resident data, the bit-sliced
    form transposing in and out every map as a fused kernel would.  Both
    outputs are compared; VALU instructions per wave come from the committed
    rocprofv3 pass (profiles/phaseb_prototype.json) when it was made from
    these device sources.  None if the probe library is missing."""
    import ctypes
    path = os.path.join(ROOT, "additive-fft-reed-solomon_amd", "libars_probe.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.ars_probe_phaseb.restype = ctypes.c_int
    lib.ars_probe_phaseb.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double)]
    maps = lib.ars_probe_phaseb_maps()
    L = 1 << 22
    g = torch.Generator(device="cuda")
    g.manual_seed(49)
    src = torch.randint(0, 256, (16 * L,), dtype=torch.uint8, device="cuda", generator=g)
    try:
        with open(os.path.join(ROOT, "profiles", "phaseb_prototype.json")) as f:
            doc = json.load(f)
        prof = doc["kernels"] if doc.get("device_source_sha") == device_source_sha() else {}
    except (OSError, ValueError, KeyError):
        prof = {}
    out = {"workload": f"phase B of the (256,k) recover on 16 rows x {L} columns, {maps} maps per launch",
           "maps_per_launch": maps}
    res = []
    for v, name, kern, cols in ((0, "v_perm_bytes", "{anon}::k_phaseb_bytes", 4),
                                (1, "bit_sliced", "{anon}::k_phaseb_bits", 32)):
        o = torch.empty_like(src)
        us = ctypes.c_double()
        if lib.ars_probe_phaseb(v, src.data_ptr(), o.data_ptr(), L, 20, ctypes.byref(us)) != 0:
            return None
        res.append(o)
        rec = prof.get(kern)
        out[name] = {"us": round(us.value, 2), "ps_per_column_map": round(us.value * 1e6 / (L * maps), 3),
                     "columns_per_lane": cols,
                     "valu_instr_per_column_map": (round(rec["valu_instr_per_wave"] / (64 * cols * maps), 2)
                                                   if rec else None)}
    torch.cuda.synchronize()
    out["outputs_equal"] = bool(torch.equal(res[0], res[1]))
    out["bit_sliced_over_v_perm_time"] = round(out["bit_sliced"]["us"] / out["v_perm_bytes"]["us"], 3)
    return out


def valu_roofline(kernel, seconds, rates):
    """Secondary roofline: the kernel's VALU instruction rate (committed
    SQ_INSTS_VALU per wave x waves, profiles/valu_256x128_1MiB.json, over the
    launch time measured now) against the attainable rate for its instruction
    mix.  gfx950 issues one wave's v_perm_b32 beside other waves' simple ops,
    so the peak is the overlapped bound: the simple ops at the live v_xor_b32
    rate or the lookups at the live v_perm_b32 rate, whichever takes longer.
    The kernel is also set against the register-only probes of its butterfly:
    default wave priorities, and the lookup bursts the tiles use."""
    try:
        with open(os.path.join(ROOT, "profiles", "valu_256x128_1MiB.json")) as f:
            perm = json.load(f)["perm_fraction"]
    except (OSError, ValueError, KeyError):
        return None
    name, rec = _profile_entry("valu_256x128_1MiB.json", kernel)
    if rec is None or rates is None:
        return None
    fam = perm[name]["value"] if name in perm else None
    if fam is None:
        return None
    achieved = rec["valu_instr_per_wave"] * rec["waves"] / seconds
    peak = 1.0 / max(fam / rates["perm"], (1.0 - fam) / rates["simple"])
    return {"bound": "valu", "kernel": name, "achieved": round(achieved / 1e12, 3), "peak": round(peak / 1e12, 3),
            "unit": "T wave64-instr/s", "frac": round(achieved / peak, 4), "perm_fraction": fam,
            # the kernel against the same multiply-add butterflies issued in
            # straight-line code (its instruction mix is that butterfly's)
            "frac_of_straight_line_butterfly_rate": round(achieved / rates["butterfly_straight_line"], 4),
            "frac_of_burst_butterfly_rate": round(achieved / rates["butterfly_burst"], 4),
            "peak_from": "live v_xor_b32 and v_perm_b32 issue rates, overlapped: max(simple share / v_xor rate, "
                         "v_perm share / v_perm rate)"}


# ---------------------------------------------------------------------------
# rocprofv3 summaries.  A profiled run writes CSV directories (one per
# PROFILE_RUNS entry, `profile_command`); `--summarize-rocprof DIR` turns them
# into the profiles/*.md tables and the profiles/*.json records that the
# roofline fields above read, stamped with the device-source SHA the run was
# built from and the git commit.

PROFILE_HEAD = "--no-cpu --no-batched --no-served --no-host --no-decode --no-grid --no-c5 --no-wide"
PROFILE_RUNS = {
    "trace": ("--kernel-trace --stats", "--steps 5 --warmup 2 " + PROFILE_HEAD),
    "fetch": ("--pmc FETCH_SIZE", "--steps 2 --warmup 1 " + PROFILE_HEAD),
    "write": ("--pmc WRITE_SIZE", "--steps 2 --warmup 1 " + PROFILE_HEAD),
    "valu": ("--pmc SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_WAVES SQ_BUSY_CYCLES GRBM_GUI_ACTIVE",
             "--steps 2 --warmup 1 " + PROFILE_HEAD),
    "lds": ("--pmc SQ_LDS_IDX_ACTIVE SQ_LDS_BANK_CONFLICT SQ_LDS_DATA_FIFO_FULL SQ_LDS_CMD_FIFO_FULL "
            "SQ_INSTS_LDS SQ_WAVES SQ_WAIT_INST_LDS SQ_ACTIVE_INST_LDS GRBM_GUI_ACTIVE",
            "--steps 2 --warmup 1 " + PROFILE_HEAD),
    # read requests by size (FETCH_SIZE's 128-byte term reads TCC_BUBBLE, which
    # gfx950 does not count: hence its calibration), atomics, and DRAM requests
    "rdsize": ("--pmc TCC_EA0_RDREQ TCC_EA0_RDREQ_32B TCC_EA0_RDREQ_64B TCC_EA0_RDREQ_128B",
               "--steps 2 --warmup 1 " + PROFILE_HEAD),
    "atomic": ("--pmc TCC_ATOMIC TCC_EA0_ATOMIC TCC_EA0_WRREQ TCC_EA0_WRREQ_64B", "--steps 2 --warmup 1 " + PROFILE_HEAD),
    "dram": ("--pmc TCC_EA0_RDREQ_DRAM TCC_EA0_WRREQ_DRAM TCC_EA0_WRREQ_ATOMIC_DRAM TCC_EA0_RDREQ_DRAM_32B",
             "--steps 2 --warmup 1 " + PROFILE_HEAD),
    "phaseb": ("--pmc SQ_INSTS_VALU SQ_WAVES SQ_BUSY_CYCLES GRBM_GUI_ACTIVE", "--phaseb"),
    "batched": ("--kernel-trace --stats",
                "--steps 2 --warmup 1 --no-cpu --no-served --no-host --no-decode --no-grid --no-c5 --no-wide"),
    "decode": ("--kernel-trace --stats",
               "--steps 2 --warmup 1 --no-cpu --no-batched --no-served --no-host --no-grid --no-c5 --no-wide"),
}


def profile_command(name, outdir, repo="$R"):
    prof, args = PROFILE_RUNS[name]
    return f"rocprofv3 {prof} -f csv -d {outdir}/{name} -o run -- python3 {repo}/bench.py {args}"


def _kname(raw):
    """rocprofv3's kernel name without return type, argument list and the
    library's namespace."""
    s = raw.strip().replace("(anonymous namespace)", "{anon}")
    if s.startswith("void "):
        s = s[5:]
    depth = 0
    for i, ch in enumerate(s):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            s = s[:i]
            break
    return s.replace("ars::", "").strip()


def _rocprof_csv(d, suffix):
    import csv
    import glob
    out = []
    for f in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        with open(f, newline="") as fh:
            out += [dict(r, _file=f) for r in csv.DictReader(fh)]
    return out


def _dispatches(d):
    """{(file, dispatch id): {"kernel", "grid", counter: value}} of a --pmc run."""
    per = {}
    for r in _rocprof_csv(d, "counter_collection.csv"):
        key = (r["_file"], int(r["Dispatch_Id"]))
        e = per.setdefault(key, {"kernel": _kname(r["Kernel_Name"]), "grid": int(float(r.get("Grid_Size") or 0))})
        e[r["Counter_Name"]] = e.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return per


def _by_kernel(per, counters):
    """{kernel: {"dispatches": c, counter: mean per dispatch}}."""
    acc = {}
    for e in per.values():
        if not all(c in e for c in counters):
            continue
        a = acc.setdefault(e["kernel"], {"dispatches": 0, **{c: 0.0 for c in counters}})
        a["dispatches"] += 1
        for c in counters:
            a[c] += e[c]
    for a in acc.values():
        for c in counters:
            a[c] /= a["dispatches"]
    return acc


def _mangle(kernel):
    """Itanium name of a namespace-ars kernel template with int / bool
    arguments, e.g. k_recover_stream<8, 64, true, 1>."""
    base, _, args = kernel.partition("<")
    enc = ""
    for a in args.rstrip(">").split(","):
        a = a.strip()
        enc += "Lb1E" if a == "true" else "Lb0E" if a == "false" else f"Li{int(a)}E"
    return f"_ZN3ars{len(base)}{base}I{enc}E"


def _perm_fractions(kernels):
    """v_perm_b32 share of the VALU instructions in each kernel's gfx950
    assembly (static count over the kernel body, built here with hipcc)."""
    import re
    import subprocess
    import tempfile
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    src = os.path.join(ROOT, "additive-fft-reed-solomon_amd", "csrc", "ars_kernels.hip")
    with tempfile.TemporaryDirectory() as td:
        dst = os.path.join(td, "k.s")
        subprocess.run([hipcc, "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-S", "--cuda-device-only", src, "-o", dst],
                       check=True, capture_output=True)
        asm = open(dst).read().split("\n")
    starts = [i for i, l in enumerate(asm) if re.match(r"^_Z[A-Za-z0-9_]+:", l)] + [len(asm)]
    bodies = {asm[a].split(":")[0]: asm[a:b] for a, b in zip(starts, starts[1:])}
    out = {}
    for k in kernels:
        # the symbol continues with the parameter types after the template arguments
        body = next((b for sym, b in bodies.items() if sym.startswith(_mangle(k))), None)
        if body is None:
            continue
        # (inline-assembly lines may start at column 0)
        valu = sum(1 for l in body if re.match(r"^\s*v_", l))
        perm = sum(1 for l in body if re.match(r"^\s*v_perm_b32", l))
        if valu:
            out[k] = {"value": round(perm / valu, 4),
                      "source": f"{perm} v_perm_b32 of {valu} VALU instructions in the gfx950 assembly of {k}"}
    return out


def _headline_kernel(k):
    return k.startswith(("k_encode_tile<7,", "k_recover_tile<8,", "k_recover_stream<8,"))


def summarize_rocprof(d):
    """Writes the profiles/ records of the runs under d (see PROFILE_RUNS)."""
    import subprocess
    sha = device_source_sha()
    with open(os.path.join(d, "sha.txt")) as f:
        run_sha = f.read().strip()
    if run_sha != sha:
        raise SystemExit(f"profiles in {d} were made from device sources {run_sha[:12]}, this tree has {sha[:12]}")
    commit = subprocess.run(["git", "-C", ROOT, "rev-parse", "HEAD"], capture_output=True, text=True).stdout.strip()
    prof = os.path.join(ROOT, "profiles")
    n, k, L, e = 256, 128, 1 << 20, 128
    alg = {"encode": n * L, "recover": (n + e) * L}
    stamp = {"device_source_sha": sha, "commit": commit}
    wl = "(n=256,k=128), L=1 MiB, e=128"
    how = ("Run on one MI355X from /tmp with TMPDIR=/tmp, repo at $R; each counter set in a run of its own, "
           "every pass under a `timeout -s KILL` of its own (60-90 s).  Summarised by `python bench.py --summarize-rocprof DIR`, "
           f"device sources {sha[:16]} (bench.device_source_sha), commit {commit[:12]}.")

    def md(fname, title, cmds, body):
        with open(os.path.join(prof, fname), "w") as f:
            f.write(f"# {title}\n\n{how}\n\n```\n" + "\n".join(cmds) + "\n```\n\n" + body.rstrip() + "\n")

    # kernel stats of the headline run, the batched run and the decode run
    def live(run):
        """The bench line the profiled run printed: its own event timing of the
        dominant kernel, encode and recover on the same box, for comparison."""
        try:
            line = [x for x in open(os.path.join(d, run + ".log")) if x.startswith("{")][-1]
            b = json.loads(line)
        except (OSError, IndexError, ValueError):
            return None
        return {"kernel": b["roofline"]["kernel"], "launch_us": b["roofline"].get("launch_us"),
                "encode_us": b["encode"]["us"], "recover_us": b["recover"]["us"]}

    def stats(run, fname, title, note, keep):
        rows = _rocprof_csv(os.path.join(d, run), "kernel_stats.csv")
        recs = {}
        for r in rows:
            kn = _kname(r["Name"])
            recs[kn] = {"calls": int(r["Calls"]), "avg_us": round(float(r["AverageNs"]) / 1e3, 2),
                        "min_us": round(float(r["MinNs"]) / 1e3, 2), "max_us": round(float(r["MaxNs"]) / 1e3, 2),
                        "total_ms": round(float(r["TotalDurationNs"]) / 1e6, 3)}
        order = sorted(recs, key=lambda x: -recs[x]["total_ms"])
        t = "| Kernel | Calls | Avg (us) | Min (us) | Max (us) | Total (ms) |\n|---|---|---|---|---|---|\n"
        t += "".join(f"| `{x}` | {recs[x]['calls']} | {recs[x]['avg_us']} | {recs[x]['min_us']} | "
                     f"{recs[x]['max_us']} | {recs[x]['total_ms']} |\n" for x in order[:keep])
        lv = live(run)
        if lv:
            note += (f"\n\nThe same run's bench line timed `{lv['kernel']}` at {lv['launch_us']} us per launch "
                     f"(HIP events around back-to-back launches on its stream; encode {lv['encode_us']} us, recover "
                     f"{lv['recover_us']} us).  The rocprof average below also holds the autotuner's and the "
                     "correctness gate's launches.")
        md(fname + ".md", title, [profile_command(run, "/tmp/prof")], note + "\n\n" + t)
        with open(os.path.join(prof, fname + ".json"), "w") as f:
            json.dump({"workload": title, **stamp, "bench_line": lv, "kernels": {x: recs[x] for x in order}}, f,
                      indent=1)
        return recs

    stats("trace", "kernel_stats_256x128_1MiB", f"Kernel stats: {wl}, encode + recover on one MI355X",
          "The run launches the VALU issue-rate probes (`k_xor`, `k_perm`, `k_butterfly`, `k_butterfly_long`), "
          "the autotuner (every tile variant, 6 launches each), the correctness gate, warm-up, 5 timed steps, "
          "and 3 x max(steps, 20) + 1 launches each of encode alone and recover-with-plan alone. torch kernels are "
          "the bench's data set-up, correctness gate and copy reference.", 30)
    if os.path.isdir(os.path.join(d, "batched")):
        stats("batched", "kernel_stats_batched_mixed_4096", "Kernel stats: batched mixed-size workload vs one "
              "launch per codeword", "Workload: 4096 codewords, n ~ U{32,64,128,256}, k = n/2, shard_len "
              "log-uniform 64 B - 64 KiB, encode then recover of e ~ U{0..T} erasures, as one batch and as one "
              "call per codeword; the run also holds the (256,128) headline steps and the autotuner.", 30)
    if os.path.isdir(os.path.join(d, "decode")):
        stats("decode", "kernel_stats_decode_256x128", "Kernel stats: sharded error decoding, (n=256,k=128)",
              "The run holds the bench's decode line (clean scrub, corrupted columns, formal-derivative mode, "
              "key equations on the GPU) beside the headline steps.", 30)

    # HBM traffic: FETCH_SIZE and WRITE_SIZE passes, calibrated on the
    # largest device-to-device copy (n*L bytes read and written); read
    # requests counted by size (no calibration) as the cross-check; atomic
    # requests, which the L2 forwards to the memory side, where each reads
    # the line it modifies
    fetch, write = _dispatches(os.path.join(d, "fetch")), _dispatches(os.path.join(d, "write"))
    cal = {}
    for name, per, ctr in (("fetch", fetch, "FETCH_SIZE"), ("write", write, "WRITE_SIZE")):
        cps = [x for x in per.values() if "copyBuffer" in x["kernel"] and ctr in x]
        g = max(x["grid"] for x in cps)
        vals = sorted(x[ctr] for x in cps if x["grid"] == g)
        cal[name] = {"kernel": cps[0]["kernel"], "grid": g, "dispatches": len(vals),
                     f"{ctr}_KiB": round(vals[len(vals) // 2], 1), "true_KiB": n * L / 1024,
                     "factor": round(n * L / 1024 / vals[len(vals) // 2], 4)}
    if abs(cal["write"]["factor"] - 1.0) > 0.02:
        raise SystemExit(f"the largest copy is not n*L bytes: {cal['write']}")
    fk, wk = _by_kernel(fetch, ["FETCH_SIZE"]), _by_kernel(write, ["WRITE_SIZE"])
    rq = _by_kernel(_dispatches(os.path.join(d, "rdsize")),
                    ["TCC_EA0_RDREQ_32B", "TCC_EA0_RDREQ_64B", "TCC_EA0_RDREQ_128B"])
    at = _by_kernel(_dispatches(os.path.join(d, "atomic")), ["TCC_ATOMIC", "TCC_EA0_ATOMIC"])
    dr = _by_kernel(_dispatches(os.path.join(d, "dram")), ["TCC_EA0_WRREQ_ATOMIC_DRAM"])
    tr = {}
    for kn in sorted(set(fk) & set(wk) & set(rq) & set(at)):
        if not _headline_kernel(kn):
            continue
        a = alg["encode" if kn.startswith("k_encode") else "recover"]
        rd = fk[kn]["FETCH_SIZE"] * cal["fetch"]["factor"] * 1024
        wr = wk[kn]["WRITE_SIZE"] * cal["write"]["factor"] * 1024
        rmw = 64 * at[kn]["TCC_EA0_ATOMIC"]
        by_size = (32 * rq[kn]["TCC_EA0_RDREQ_32B"] + 64 * rq[kn]["TCC_EA0_RDREQ_64B"]
                   + 128 * rq[kn]["TCC_EA0_RDREQ_128B"])
        t = rd + wr + rmw
        tr[kn] = {"FETCH_SIZE_KiB": round(fk[kn]["FETCH_SIZE"], 1), "WRITE_SIZE_KiB": round(wk[kn]["WRITE_SIZE"], 1),
                  "read_bytes_by_request_size": int(by_size), "atomic_requests": int(at[kn]["TCC_EA0_ATOMIC"]),
                  "atomic_requests_at_dram": int(dr[kn]["TCC_EA0_WRREQ_ATOMIC_DRAM"]) if kn in dr else None,
                  "dispatches": fk[kn]["dispatches"], "read_bytes": int(rd), "write_bytes": int(wr),
                  "atomic_rmw_read_bytes": int(rmw), "traffic_bytes": int(t), "algorithmic_bytes": a,
                  "traffic_over_algorithmic": round(t / a, 4)}
    with open(os.path.join(prof, "traffic_256x128_1MiB.json"), "w") as f:
        json.dump({"workload": wl, **stamp, "fetch_write_units": "KiB per dispatch", "calibration": cal,
                   "traffic": "FETCH_SIZE x fetch factor + WRITE_SIZE x write factor + 64 B per atomic request "
                              "(read by the memory side that executes it)", "kernels": tr}, f, indent=1)
    t = ("| Kernel | Disp. | FETCH_SIZE (KiB) | read bytes (FETCH x f) | read bytes by request size | WRITE_SIZE (KiB) | "
         "atomic requests (at DRAM) | traffic (bytes) | algorithmic (bytes) | ratio |\n"
         "|---|---|---|---|---|---|---|---|---|---|\n")
    t += "".join(f"| `{x}` | {v['dispatches']} | {v['FETCH_SIZE_KiB']} | {v['read_bytes']} | "
                 f"{v['read_bytes_by_request_size']} | {v['WRITE_SIZE_KiB']} | {v['atomic_requests']} "
                 f"({v['atomic_requests_at_dram']}) | {v['traffic_bytes']} | {v['algorithmic_bytes']} | "
                 f"{v['traffic_over_algorithmic']} |\n" for x, v in tr.items())
    cf, cw = cal["fetch"], cal["write"]
    md("hbm_traffic_256x128_1MiB.md", f"HBM traffic per launch: {wl}",
       [profile_command(x, "/tmp/prof") for x in ("fetch", "write", "rdsize", "atomic", "dram")],
       f"**Calibration** on the largest device-to-device copy of the run (`{cf['kernel']}`, {n * L} bytes read "
       f"and written, median of {cf['dispatches']} dispatches): FETCH_SIZE {cf['FETCH_SIZE_KiB']} KiB, factor "
       f"{cf['factor']}; WRITE_SIZE {cw['WRITE_SIZE_KiB']} KiB, factor {cw['factor']}.  FETCH_SIZE's 128-byte "
       "term counts TCC_BUBBLE, which stays 0 on gfx950 while its reads are 128-byte requests "
       "(TCC_EA0_RDREQ_128B), so FETCH_SIZE shows half of them; the read bytes counted by request size "
       "(32 x RDREQ_32B + 64 x RDREQ_64B + 128 x RDREQ_128B) need no factor and agree with FETCH_SIZE x factor.\n\n"
       "**Atomics.** The streaming recover adds each recovered quad into its erased shard with "
       "`buffer_atomic_xor`.  TCC_EA0_ATOMIC shows the L2 forwarding every one of them to the memory side as "
       "a 64-byte request (TCC_EA0_WRREQ_ATOMIC_DRAM: all to DRAM), where the line is read, modified and "
       "written; WRITE_SIZE counts the request, no TCC counter sees the read, so it is added as 64 bytes per "
       "request.  The streaming kernel reads only the n - e shards it keeps (the erased ones are neither "
       "loaded nor multiplied).\n\n"
       "traffic = FETCH_SIZE x fetch factor + WRITE_SIZE x write factor + 64 x atomic requests.  Algorithmic "
       "bytes: n*L for encode (k*L read, T*L written), (n+e)*L for recover (n*L read, e*L written).  Every tile "
       "variant the autotuner launches is listed.\n\n" + t)

    # VALU issue and LDS activity
    vk = _by_kernel(_dispatches(os.path.join(d, "valu")),
                    ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_WAVES", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE"])
    heads = [x for x in sorted(vk) if _headline_kernel(x)]
    perm = _perm_fractions(heads)
    vr = {x: {"waves": int(vk[x]["SQ_WAVES"]), "valu_instr_per_wave": round(vk[x]["SQ_INSTS_VALU"] / vk[x]["SQ_WAVES"], 1),
              "busy_cycles_per_xcd": int(vk[x]["GRBM_GUI_ACTIVE"] / 8), "dispatches": vk[x]["dispatches"]}
          for x in heads}
    with open(os.path.join(prof, "valu_256x128_1MiB.json"), "w") as f:
        json.dump({"workload": wl, **stamp, "perm_fraction": perm, "kernels": vr}, f, indent=1)
    t = "| Kernel | Dispatches | Waves | VALU instr / wave | busy cycles / XCD | v_perm share (static) |\n|---|---|---|---|---|---|\n"
    t += "".join(f"| `{x}` | {v['dispatches']} | {v['waves']} | {v['valu_instr_per_wave']} | {v['busy_cycles_per_xcd']} "
                 f"| {perm.get(x, {}).get('value', '')} |\n" for x, v in vr.items())
    md("valu_256x128_1MiB.md", f"VALU issue: {wl}", [profile_command("valu", "/tmp/prof")],
       "SQ_INSTS_VALU counts wave64 VALU instructions over the device; GRBM_GUI_ACTIVE is summed over the 8 "
       "XCDs (busy cycles per XCD = value / 8).  The v_perm share is counted in the kernel's gfx950 assembly "
       "(each GF(2^8) multiply of four bytes is three v_perm_b32 table lookups, counted in the inline-assembly "
       "bursts too); bench.py's VALU roofline takes the overlapped bound of the live v_xor_b32 and v_perm_b32 "
       "issue rates for that share.\n\n" + t)
    lc = ["SQ_LDS_IDX_ACTIVE", "SQ_LDS_BANK_CONFLICT", "SQ_LDS_DATA_FIFO_FULL", "SQ_LDS_CMD_FIFO_FULL", "SQ_INSTS_LDS",
          "SQ_WAVES", "SQ_WAIT_INST_LDS", "SQ_ACTIVE_INST_LDS", "GRBM_GUI_ACTIVE"]
    lk = _by_kernel(_dispatches(os.path.join(d, "lds")), lc)
    t = ("| Kernel | Dispatches | Waves | Cycles / XCD | LDS busy / CU | LDS busy | Bank-conflict cycles / CU | "
         "LDS instr / wave |\n|---|---|---|---|---|---|---|---|\n")
    for x in sorted(lk):
        if not _headline_kernel(x):
            continue
        v = lk[x]
        cyc = v["GRBM_GUI_ACTIVE"] / 8
        busy = v["SQ_LDS_IDX_ACTIVE"] / 256
        t += (f"| `{x}` | {v['dispatches']} | {int(v['SQ_WAVES'])} | {int(cyc)} | {int(busy)} | {busy / cyc:.3f} | "
              f"{int(v['SQ_LDS_BANK_CONFLICT'] / 256)} | {v['SQ_INSTS_LDS'] / v['SQ_WAVES']:.1f} |\n")
    md("lds_256x128_1MiB.md", f"LDS activity per launch: {wl}", [profile_command("lds", "/tmp/prof")],
       "Per-CU figures are chip sums / 256 CUs; LDS busy = SQ_LDS_IDX_ACTIVE per CU over the cycles per XCD "
       "(rocprofiler-sdk's LdsUtil expression for gfx950).\n\n" + t)
    # the bit-sliced phase-B prototype against the v_perm form (bench_phaseb)
    if os.path.isdir(os.path.join(d, "phaseb")):
        pc = ["SQ_INSTS_VALU", "SQ_WAVES", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE"]
        pk = _by_kernel(_dispatches(os.path.join(d, "phaseb")), pc)
        try:
            line = json.loads([x for x in open(os.path.join(d, "phaseb.log")) if x.startswith("{")][-1])
        except (OSError, IndexError, ValueError):
            line = None
        maps = line["maps_per_launch"] if line else 4
        cols = {"k_phaseb_bytes": 4, "k_phaseb_bits": 32}
        recs = {}
        for x, v in sorted(pk.items()):
            c = next((n for k2, n in cols.items() if x.endswith(k2)), None)
            if c is None:
                continue
            vpw = v["SQ_INSTS_VALU"] / v["SQ_WAVES"]
            recs[x] = {"dispatches": v["dispatches"], "waves": int(v["SQ_WAVES"]), "valu_instr_per_wave": round(vpw, 1),
                       "columns_per_lane": c, "valu_instr_per_column_map": round(vpw / (64 * c * maps), 2),
                       "busy_cycles_per_xcd": int(v["GRBM_GUI_ACTIVE"] / 8)}
        with open(os.path.join(prof, "phaseb_prototype.json"), "w") as f:
            json.dump({"workload": "phase B of the 256-point recover, 16 rows x 4 MiB", **stamp, "bench_line": line,
                       "kernels": recs}, f, indent=1)
        t = ("| Kernel | Dispatches | Waves | Columns / lane | VALU instr / wave | VALU instr / column / map | "
             "busy cycles / XCD |\n|---|---|---|---|---|---|---|\n")
        t += "".join(f"| `{x}` | {v['dispatches']} | {v['waves']} | {v['columns_per_lane']} | {v['valu_instr_per_wave']} | "
                     f"{v['valu_instr_per_column_map']} | {v['busy_cycles_per_xcd']} |\n" for x, v in recs.items())
        if line:
            t += (f"\nThe same run's timing (HIP events, 20 launches): v_perm form "
                  f"{line['v_perm_bytes']['ps_per_column_map']} ps per column and map, bit-sliced "
                  f"{line['bit_sliced']['ps_per_column_map']} ps (ratio {line['bit_sliced_over_v_perm_time']}); "
                  f"outputs equal: {line['outputs_equal']}.\n")
        md("phaseb_prototype.md", "Bit-sliced phase B against the v_perm form",
           [profile_command("phaseb", "/tmp/prof")],
           "Phase B of the 256-point recover (IFFT levels 4..7, the derivative's high-bit terms, FFT levels 7..4; "
           "compile-time twiddles) on register-resident data, "
           f"{maps} maps per launch.  `k_phaseb_bytes` is the product's form (one column quad per lane, three "
           "v_perm_b32 lookups per four byte products); `k_phaseb_bits` holds 32 columns per lane as 8 bit "
           "planes, multiplies with compile-time XOR networks and transposes each row in and out every map.\n\n"
           + t)
    print(f"wrote profiles from {d} (device sources {sha[:12]}, commit {commit[:12]})")


def c3_workload(np, count=4096, seed=44):
    """The mixed batch (C3): `count` codewords with n ~ U{32,64,128,256},
    k = n/2, shard_len log-uniform in [64 B, 64 KiB] (multiple of 64), laid
    out back to back (256-byte aligned) in one pool; each is encoded, then
    recovers e ~ U{0..T} random erasures.  Returns the shapes, pool offsets and
    erasure positions; the pool's bytes come from a torch generator seeded
    with `seed`."""
    rng = np.random.default_rng(seed)
    ns = rng.choice([32, 64, 128, 256], count)
    Ls = (64 * np.maximum(1, np.round(2.0 ** rng.uniform(6, 16, count) / 64))).astype(np.int64)
    sizes = ns * Ls
    offs = np.zeros(count, np.int64)
    offs[1:] = np.cumsum((sizes + 255) // 256 * 256)[:-1]
    total = int(offs[-1] + sizes[-1])
    positions = []
    for i in range(count):
        T = int(ns[i]) - int(ns[i]) // 2
        e = int(rng.integers(0, T + 1))
        positions.append(rng.permutation(int(ns[i]))[:e])
    return {"ns": ns, "Ls": Ls, "offs": offs, "total": total, "positions": positions, "seed": seed}


def c3_jobs(ars, w, pool):
    """ars.Batch job arrays (encode, then recover) of a c3_workload over `pool`."""
    enc, rec = [], []
    enc_bytes = rec_bytes = 0
    for i in range(len(w["ns"])):
        n, L, o = int(w["ns"][i]), int(w["Ls"][i]), int(w["offs"][i])
        k = n // 2
        T = n - k
        enc.append({"op": "encode", "n": n, "k": k, "shard_len": L,
                    "message": pool[o + T * L:o + n * L], "parity": pool[o:o + T * L]})
        e = len(w["positions"][i])
        rec.append({"op": "recover", "n": n, "k": k, "shard_len": L, "codeword": pool[o:o + n * L],
                    "positions": w["positions"][i]})
        enc_bytes += n * L
        rec_bytes += (n + e) * L if e else 0
    earr, ekeep = ars.Batch.make_jobs(enc)
    rarr, rkeep = ars.Batch.make_jobs(rec)
    return earr, rarr, (ekeep, rkeep), enc_bytes + rec_bytes


def bench_batched(ars, torch, np, stream, count=4096, seed=44, reps=3):
    """Mixed sizes (c3_workload), one persistent launch vs one launch per
    codeword.  Before timing, one batched pass and one plain pass from the
    same starting bytes must leave identical pools."""
    w = c3_workload(np, count, seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    pool = torch.randint(0, 256, (w["total"],), dtype=torch.uint8, device="cuda", generator=gen)
    earr, rarr, keep, byts = c3_jobs(ars, w, pool)
    sizes = w["ns"] * w["Ls"]
    b = ars.Batch(count)
    ctx = b.ctx

    def batched():
        b.run_jobs(earr, count, stream)
        b.run_jobs(rarr, count, stream)

    def plain():
        ars.lib().ars_run_jobs_plain(ctx.handle, earr, count, stream.cuda_stream)
        ars.lib().ars_run_jobs_plain(ctx.handle, rarr, count, stream.cuda_stream)

    def timed(fn, r):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        z = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        a.record(stream)
        for _ in range(r):
            fn()
        z.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(z) / r * 1e-3, (time.perf_counter() - w0) / r

    # (the copies run on torch's current stream, the jobs on `stream`)
    start = pool.clone()
    torch.cuda.synchronize()
    batched()
    torch.cuda.synchronize()
    got = pool.clone()
    pool.copy_(start)
    torch.cuda.synchronize()
    plain()
    torch.cuda.synchronize()
    if not torch.equal(got, pool):
        raise RuntimeError("batched run differs from plain calls on the mixed workload")
    del got, start
    s_b, w_b = timed(batched, reps)
    s_p, w_p = timed(plain, reps)
    info = b.info()
    b.close()
    del pool
    return {
        "workload": f"{count} codewords, n~U{{32,64,128,256}}, k=n/2, L log-uniform 64 B-64 KiB, "
                    "encode then recover e~U{0..T}",
        "bytes": byts, "data_bytes": int(sizes.sum()),
        "batched_GBps": round(byts / max(s_b, w_b) / 1e9, 1),
        "batched_ms": round(max(s_b, w_b) * 1e3, 3),
        "plain_GBps": round(byts / max(s_p, w_p) / 1e9, 1),
        "plain_ms": round(max(s_p, w_p) * 1e3, 3),
        "speedup": round(max(s_p, w_p) / max(s_b, w_b), 2),
        "items_last_run": info["items"], "grid": info["grid"],
        "checked": "one batched pass equals one plain-call pass, whole pool compared",
    }


C4_SHAPES = [(64, 32), (128, 96), (256, 128), (256, 224)]


def c4_trace(np, rate, count, seed):
    """The served stream (C4): Poisson arrivals at `rate` requests/s, 70 %
    encode / 30 % recover, (n,k) ~ U{(64,32),(128,96),(256,128),(256,224)},
    shard_len log-uniform 1 KiB-4 MiB (multiple of 64), recover e ~ U{1..T}.
    Returns [(op, n, k, L, positions)] and the arrival times (s)."""
    rng = np.random.default_rng(seed)
    trace = []
    for i in range(count):
        n, k = C4_SHAPES[rng.integers(4)]
        L = int(64 * max(1, round(2.0 ** rng.uniform(10, 22) / 64)))
        if rng.random() < 0.7:
            trace.append(("encode", n, k, L, None))
        else:
            e = int(rng.integers(1, n - k + 1))
            trace.append(("recover", n, k, L, rng.permutation(n)[:e]))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    return trace, arrivals


def check_served(ars, torch, np, srv, ctx, count=32, seed=4501):
    """A short C4 trace with host payloads through the server: every output
    must equal a single ars_encode / ars_recover call on the same input.
    Returns the number of requests checked."""
    trace, arrivals = c4_trace(np, 4000, count, seed)
    reqs = (ars.Request * count)()
    keep, outs, expect = [], [], []
    for i, (op, n, k, L, pos) in enumerate(trace):
        T = n - k
        msg = torch.randint(0, 256, (k * L,), dtype=torch.uint8, device="cuda")
        if op == "encode":
            exp = torch.zeros(T * L, dtype=torch.uint8, device="cuda")
            ctx.encode(n, k, msg.data_ptr(), exp.data_ptr(), L)
            inp = msg.cpu().numpy()
        else:
            cw = torch.zeros(n * L, dtype=torch.uint8, device="cuda")
            ctx.encode(n, k, msg.data_ptr(), cw.data_ptr(), L)
            cw[T * L:] = msg
            cw.view(n, L)[torch.from_numpy(pos.astype(np.int64)).cuda()] = 0
            inp = cw.cpu().numpy()
            exp = cw
            ctx.recover(n, k, exp.data_ptr(), L, pos)
        torch.cuda.synchronize()
        expect.append(exp.cpu().numpy())
        outs.append(np.zeros_like(expect[-1]))
        reqs[i] = ars.Server.request(op, n, k, L, positions=pos if pos is not None else (), host_in=inp,
                                     host_out=outs[-1], keep=keep)
        keep.append(inp)
    st, _ = srv.replay(reqs, arrivals)
    for i, (op, n, k, L, pos) in enumerate(trace):
        if st[i] != (0 if op == "encode" else 1) or not np.array_equal(outs[i], expect[i]):
            raise RuntimeError(f"served request {i} ({op} n={n} k={k} L={L}) differs from a single call")
    return count


def bench_served(ars, torch, np, rates=(2000, 8000, 16000, 24000, 32000, 48000), per_rate=1500, seed=45):
    """Poisson arrivals (c4_trace) under a 64 GiB HBM admission budget;
    payloads are device-resident (the pool slot's bytes), so latency is
    queueing + GPU.  Each trace is then replayed through plain calls
    (ars_plain_replay: one call per request, in arrival order on one stream,
    no batching) for the plain-call latency and throughput beside the
    server's."""
    srv = ars.Server(hbm_budget=64 << 30, max_batch_jobs=1024, pipeline_depth=2)
    ctx = ars.Context(0)
    checked = check_served(ars, torch, np, srv, ctx)
    out = []
    for ri, rate in enumerate(rates):
        trace, arrivals = c4_trace(np, rate, per_rate, seed + ri)
        reqs = (ars.Request * per_rate)()
        keep = []
        alg = 0
        biggest = 0
        for i, (op, n, k, L, pos) in enumerate(trace):
            biggest = max(biggest, n * L)
            reqs[i] = ars.Server.request(op, n, k, L, positions=pos, keep=keep)
            alg += n * L if op == "encode" else (n + len(pos)) * L
        srv.reset_stats()
        st, lat = srv.replay(reqs, arrivals)
        stt = srv.stats()
        span = stt["last_completion_s"] - stt["first_arrival_s"]
        buf = torch.empty(biggest, dtype=torch.uint8, device="cuda")
        stream = torch.cuda.Stream()
        pst, plat = ctx.plain_replay(reqs, arrivals, buf.data_ptr(), biggest, stream)
        pspan = float(np.max(arrivals + plat) - arrivals[0])
        del buf
        out.append({"rate_req_per_s": rate, "requests": per_rate,
                    "ok": int((st >= 0).sum()),
                    "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                    "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                    "throughput_GBps": round(stt["algorithmic_bytes"] / span / 1e9, 1),
                    "achieved_req_per_s": round(per_rate / span, 1),
                    "mean_batch_jobs": round(stt["mean_batch_jobs"], 2),
                    "batched_jobs": int(stt["batched_jobs"]), "direct_jobs": int(stt["direct_jobs"]),
                    "plain": {"ok": int((pst >= 0).sum()),
                              "p50_ms": round(float(np.percentile(plat, 50)) * 1e3, 3),
                              "p99_ms": round(float(np.percentile(plat, 99)) * 1e3, 3),
                              "throughput_GBps": round(alg / pspan / 1e9, 1),
                              "achieved_req_per_s": round(per_rate / pspan, 1)}})
        print(f"served rate {rate}: {out[-1]}", file=sys.stderr, flush=True)
    srv.close()
    ctx.close()
    return {"workload": "Poisson arrivals, 70% encode / 30% recover, (n,k) in {(64,32),(128,96),"
                        "(256,128),(256,224)}, L log-uniform 1 KiB-4 MiB, e~U{1..T}, 64 GiB budget",
            "plain": "the same trace, one ars_encode / ars_recover call per request in arrival order on "
                     "one stream (ars_plain_replay); latency = completion - arrival",
            "checked_requests": checked,
            "sustained_req_per_s": sustained(out),
            "sweep": out}


def sustained(sweep, keep_up=0.95):
    """Highest offered rate each path kept up with (achieved >= 95 % of offered)."""
    best = {"server": 0, "plain": 0}
    for r in sweep:
        if r["achieved_req_per_s"] >= keep_up * r["rate_req_per_s"]:
            best["server"] = max(best["server"], r["rate_req_per_s"])
        if r["plain"]["achieved_req_per_s"] >= keep_up * r["rate_req_per_s"]:
            best["plain"] = max(best["plain"], r["rate_req_per_s"])
    return best


def bench_served_small(ars, torch, np, rates=(20000, 50000, 100000, 200000), per_rate=4000, seed=48):
    """The served loop on small codewords, where launch cost matters: Poisson
    arrivals of the batched workload's shapes (n ~ U{32,64,128,256}, k = n/2,
    shard_len log-uniform 64 B-64 KiB, multiple of 64), 70 % encode / 30 %
    recover with e ~ U{1..T}, device-resident payloads; the server against
    the same trace through plain calls (ars_plain_replay)."""
    srv = ars.Server(hbm_budget=64 << 30, max_batch_jobs=1024, pipeline_depth=2)
    ctx = ars.Context(0)
    out = []
    for ri, rate in enumerate(rates):
        rng = np.random.default_rng(seed + ri)
        reqs = (ars.Request * per_rate)()
        keep = []
        alg = biggest = 0
        for i in range(per_rate):
            n = int(2 ** rng.integers(5, 9))
            k = n // 2
            L = int(64 * max(1, round(2.0 ** rng.uniform(6, 16) / 64)))
            biggest = max(biggest, n * L)
            if rng.random() < 0.7:
                reqs[i] = ars.Server.request("encode", n, k, L)
                alg += n * L
            else:
                e = int(rng.integers(1, n - k + 1))
                reqs[i] = ars.Server.request("recover", n, k, L, positions=rng.permutation(n)[:e], keep=keep)
                alg += (n + e) * L
        arrivals = np.cumsum(rng.exponential(1.0 / rate, per_rate))
        srv.reset_stats()
        st, lat = srv.replay(reqs, arrivals)
        stt = srv.stats()
        span = stt["last_completion_s"] - stt["first_arrival_s"]
        buf = torch.empty(biggest, dtype=torch.uint8, device="cuda")
        pst, plat = ctx.plain_replay(reqs, arrivals, buf.data_ptr(), biggest, torch.cuda.Stream())
        pspan = float(np.max(arrivals + plat) - arrivals[0])
        del buf
        out.append({"rate_req_per_s": rate, "requests": per_rate, "ok": int((st >= 0).sum()),
                    "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                    "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                    "throughput_GBps": round(stt["algorithmic_bytes"] / span / 1e9, 1),
                    "achieved_req_per_s": round(per_rate / span, 1),
                    "mean_batch_jobs": round(stt["mean_batch_jobs"], 2),
                    "batched_jobs": int(stt["batched_jobs"]), "direct_jobs": int(stt["direct_jobs"]),
                    "plain": {"ok": int((pst >= 0).sum()),
                              "p50_ms": round(float(np.percentile(plat, 50)) * 1e3, 3),
                              "p99_ms": round(float(np.percentile(plat, 99)) * 1e3, 3),
                              "throughput_GBps": round(alg / pspan / 1e9, 1),
                              "achieved_req_per_s": round(per_rate / pspan, 1)}})
        print(f"served small rate {rate}: {out[-1]}", file=sys.stderr, flush=True)
    srv.close()
    ctx.close()
    return {"workload": "Poisson arrivals, 70% encode / 30% recover, n in {32,64,128,256}, k=n/2, "
                        "L log-uniform 64 B-64 KiB, e~U{1..T}; server vs the same trace through plain calls",
            "sustained_req_per_s": sustained(out),
            "sweep": out}


# (N, T) pairs of the reference's criterion benches, each at shard_len 64,
# 1024 and 65536, aligned and one byte past alignment
# (rs/benches/encode_systematic_sharded.rs:103-160,
# rs/benches/recover_erasure_shards.rs:158-215)
GRID_NT = [(2, 1), (4, 1), (4, 2), (8, 2), (8, 4), (16, 4), (16, 8), (32, 8), (32, 16), (64, 16), (64, 32),
           (128, 32), (128, 64), (256, 64), (256, 128)]
GRID_L = (64, 1024, 65536)


def _per_call(torch, stream, fn, budget_s=0.004, sync_reps=15, min_reps=5):
    """(mean us per call issued back to back on `stream`, median us of one
    call waited for on the host).  The first is the device-side cost once
    launches overlap; the second what a caller that waits for each call sees."""
    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    one = max(time.perf_counter() - t, 1e-6)
    reps = int(min(200, max(min_reps, budget_s / one)))
    a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    z.record(stream)
    torch.cuda.synchronize()
    stream_us = a.elapsed_time(z) / reps * 1e3
    lat = [one]
    for _ in range(sync_reps - 1):
        t = time.perf_counter()
        fn()
        stream.synchronize()
        lat.append(time.perf_counter() - t)
    return stream_us, sorted(lat)[len(lat) // 2] * 1e6


def bench_reference_grid(ars, torch, np, stream, with_cpu=True, seed=42):
    """The reference's benchmark grid on the GPU: per-call encode and recover
    (T random erasures, zero-filled, as rs/benches/recover_erasure_shards.rs:
    110-134) at every (N, T, shard_len) point, device buffers aligned and one
    byte past a 256-byte boundary, beside the CPU restatement at the same
    points (median of C-timed calls, one thread).  Every GPU point is checked:
    the recovered codeword equals the encoded one."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ctx = ars.default_context(0)
    cap = 256 * 65536 + 512
    mbuf = torch.empty(cap, dtype=torch.uint8, device="cuda")
    cbuf = torch.empty(cap, dtype=torch.uint8, device="cuda")
    rows = []
    gfni = with_cpu and O.gfni_supported()
    for L in GRID_L:
        for N, T in GRID_NT:
            k = N - T
            rng = np.random.default_rng(seed + N * 1000 + T * 10 + L)
            pos = np.sort(rng.permutation(N)[:T]).astype(np.uint8)
            row = {"N": N, "T": T, "L": L}
            for aligned in (True, False):
                off = 0 if aligned else 1
                msg = mbuf[off:off + k * L]
                cw = cbuf[off:off + N * L]
                msg.copy_(torch.from_numpy(rng.integers(0, 256, k * L, dtype=np.uint8)))
                cw.zero_()
                ctx.encode(N, k, msg.data_ptr(), cw.data_ptr(), L, 0, stream)
                cw[T * L:] = msg
                torch.cuda.synchronize()
                good = cw.clone()
                cw.view(N, L)[torch.from_numpy(pos.astype(np.int64)).cuda()] = 0
                torch.cuda.synchronize()
                ctx.recover(N, k, cw.data_ptr(), L, pos, stream)
                torch.cuda.synchronize()
                if not torch.equal(cw, good):
                    raise RuntimeError(f"grid point N={N} T={T} L={L} aligned={aligned}: recover differs")
                par = cbuf[off:off + T * L]
                e_s, e_l = _per_call(torch, stream, lambda: ctx.encode(N, k, msg.data_ptr(), par.data_ptr(), L, 0,
                                                                       stream))
                # repeated in place: each call XORs its result into the shards again
                # (data-independent time)
                r_s, r_l = _per_call(torch, stream, lambda: ctx.recover(N, k, cw.data_ptr(), L, pos, stream))
                tag = "aligned" if aligned else "unaligned"
                row[tag] = {"encode_us": round(e_s, 2), "encode_sync_us": round(e_l, 2),
                            "recover_us": round(r_s, 2), "recover_sync_us": round(r_l, 2)}
                if gfni:
                    reps = int(min(400, max(5, 3e7 / (N * L))))
                    t = O.gfni_time_codec(N, T, L, pos, reps, off)
                    if t is not None:
                        row[tag]["cpu_encode_us"] = round(t[0] / 1e3, 2)
                        row[tag]["cpu_recover_us"] = round(t[1] / 1e3, 2)
            rows.append(row)
    del mbuf, cbuf

    def ratio(key, gkey):
        v = [r[a][key] / r[a][gkey] for r in rows for a in ("aligned", "unaligned") if key in r[a]]
        return round(float(np.median(v)), 2) if v else None
    return {"workload": "the reference's criterion grid: 15 (N,T) x shard_len {64, 1024, 65536} x {aligned, "
                        "+1 byte}; encode of random messages, recover of T random zero-filled erasures",
            "cpu": (f"AVX-512BW+GFNI restatement, 1 thread, {cpu_model()}" if gfni else None),
            "median_cpu_over_gpu": {"encode": ratio("cpu_encode_us", "encode_us"),
                                    "recover": ratio("cpu_recover_us", "recover_us")},
            "points": rows}


def bench_c5(ars, torch, np, stream, seed=46):
    """The adversarial shapes: (256,224) and (256,16) with 4 MiB shards --
    encode and recover GB/s and HBM fraction under all-parity,
    all-leading-data, contiguous-run and alternating erasures -- and 65536
    (32,16) codewords with 64 B shards, one call each vs one batch."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_gpu_adversarial import patterns
    ctx = ars.default_context(0)
    rng = np.random.default_rng(seed)
    L = 4 << 20
    shapes = {}
    for n, k in ((256, 224), (256, 16)):
        T = n - k
        cw = torch.randint(0, 256, (n * L,), dtype=torch.uint8, device="cuda")
        par, msg = cw[:T * L], cw[T * L:]
        ctx.encode(n, k, msg.data_ptr(), par.data_ptr(), L, 0, stream)
        torch.cuda.synchronize()
        good = cw.clone()
        enc_s, _ = _per_call(torch, stream, lambda: ctx.encode(n, k, msg.data_ptr(), par.data_ptr(), L, 0, stream),
                             budget_s=0.05, sync_reps=3)
        res = {"encode": {"us": round(enc_s, 1), "GBps": round(n * L / enc_s / 1e3, 1),
                          "frac_hbm": round(n * L / enc_s / 1e3 / HBM_PEAK_GBPS, 4)}}
        for name, pos in patterns(n, k, rng).items():
            pos = np.asarray(pos, np.uint8)
            cw.copy_(good)
            cw.view(n, L)[torch.from_numpy(pos.astype(np.int64)).cuda()] = 0
            torch.cuda.synchronize()
            ctx.recover(n, k, cw.data_ptr(), L, pos, stream)
            torch.cuda.synchronize()
            if not torch.equal(cw, good):
                raise RuntimeError(f"C5 ({n},{k}) {name}: recover differs")
            r_s, _ = _per_call(torch, stream, lambda: ctx.recover(n, k, cw.data_ptr(), L, pos, stream),
                               budget_s=0.05, sync_reps=3)
            b = (n + len(pos)) * L
            res[f"recover_{name}"] = {"e": len(pos), "us": round(r_s, 1), "GBps": round(b / r_s / 1e3, 1),
                                      "frac_hbm": round(b / r_s / 1e3 / HBM_PEAK_GBPS, 4)}
        shapes[f"({n},{k})"] = res
        del cw, good
    # 65536 tiny codewords: (32,16), 64 B shards
    count, n, k, Ls = 65536, 32, 16, 64
    T = n - k
    pool = torch.randint(0, 256, (count * n * Ls,), dtype=torch.uint8, device="cuda")
    enc, rec, alg_e, alg_r = [], [], 0, 0
    for i in range(count):
        o = i * n * Ls
        enc.append({"op": "encode", "n": n, "k": k, "shard_len": Ls, "message": pool[o + T * Ls:o + n * Ls],
                    "parity": pool[o:o + T * Ls]})
        pos = rng.permutation(n)[:int(rng.integers(1, T + 1))]
        rec.append({"op": "recover", "n": n, "k": k, "shard_len": Ls, "codeword": pool[o:o + n * Ls],
                    "positions": pos})
        alg_e += n * Ls
        alg_r += (n + len(pos)) * Ls
    earr, ekeep = ars.Batch.make_jobs(enc)
    rarr, rkeep = ars.Batch.make_jobs(rec)
    b = ars.Batch(count)
    tiny = {}
    for name, arr, alg in (("encode", earr, alg_e), ("recover", rarr, alg_r)):
        bt, _ = _per_call(torch, stream, lambda: b.run_jobs(arr, count, stream), budget_s=0.05, sync_reps=3)
        pl, _ = _per_call(torch, stream, lambda: ars.lib().ars_run_jobs_plain(ctx.handle, arr, count,
                                                                            stream.cuda_stream),
                          budget_s=0.0, sync_reps=1, min_reps=1)
        tiny[name] = {"batched_ms": round(bt / 1e3, 3), "plain_ms": round(pl / 1e3, 3),
                      "batched_GBps": round(alg / bt / 1e3, 1), "plain_GBps": round(alg / pl / 1e3, 1),
                      "speedup": round(pl / bt, 1)}
    b.close()
    del pool
    return {"workload": "(256,224) and (256,16) at 4 MiB shards with the four erasure patterns; 65536 x (32,16) "
                        "at 64 B shards, e~U{1..16}, one batch vs one call per codeword",
            "shapes": shapes, "tiny_32x16_64B": tiny}


def bench_wide(ars, torch, np, stream, seed=47):
    """(256,128) with 24 MiB shards, a 6 GiB codeword (n * shard_len >= 2^32:
    each row addressed through its own resource), beside 16 MiB - 64 shards
    just under 4 GiB (32-bit offsets): encode and recover (e = 128) GB/s.
    Each recover is checked to return its codeword before it is timed."""
    ctx = ars.default_context(0)
    rng = np.random.default_rng(seed)
    n, k = 256, 128
    T = n - k
    res = {}
    for name, L in (("narrow_4GiB_minus", (16 << 20) - 64), ("wide_6GiB", 24 << 20)):
        cw = torch.randint(0, 256, (n * L,), dtype=torch.uint8, device="cuda")
        par, msg = cw[:T * L], cw[T * L:]
        ctx.encode(n, k, msg.data_ptr(), par.data_ptr(), L, 0, stream)
        torch.cuda.synchronize()
        good = cw.clone()
        pos = np.sort(rng.permutation(n)[:T]).astype(np.uint8)
        cw.view(n, L)[torch.from_numpy(pos.astype(np.int64)).cuda()] = 0
        torch.cuda.synchronize()
        ctx.recover(n, k, cw.data_ptr(), L, pos, stream)
        torch.cuda.synchronize()
        if not torch.equal(cw, good):
            raise RuntimeError(f"{name}: recover differs")
        del good
        enc_s, _ = _per_call(torch, stream, lambda: ctx.encode(n, k, msg.data_ptr(), par.data_ptr(), L, 0, stream),
                             budget_s=0.1, sync_reps=3)
        rec_s, _ = _per_call(torch, stream, lambda: ctx.recover(n, k, cw.data_ptr(), L, pos, stream),
                             budget_s=0.1, sync_reps=3)
        res[name] = {"shard_len": L, "codeword_bytes": n * L,
                     "encode": {"ms": round(enc_s / 1e3, 3), "GBps": round(n * L / enc_s / 1e3, 1)},
                     "recover": {"ms": round(rec_s / 1e3, 3), "GBps": round((n + T) * L / rec_s / 1e3, 1)}}
        del cw, par, msg
        torch.cuda.empty_cache()
    return {"workload": "(256,128), e = 128: 16 MiB - 64 B shards (n*L < 4 GiB) vs 24 MiB shards (6 GiB)", **res}


def bench_host(ars, torch, np, n, k, L, e, reps=5, seed=46):
    """The host-slice entry points (the Rust binding's path): the same
    (n, k, L) codeword in pageable host memory, streamed through the GPU in
    column chunks.  PCIe-inclusive wall-clock rate, beside the pinned-copy
    rate of the link in each direction for scale."""
    T = n - k
    rng = np.random.default_rng(seed)
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    par = np.zeros(T * L, np.uint8)
    c = ars.Codec(n, k)
    c.encode_systematic_sharded(msg, par, None, L, accumulate=False)  # warm-up (pipeline set-up)
    cw = np.concatenate([par, msg])
    pos = rng.permutation(n)[:e].astype(np.uint8)
    t_enc = t_rec = 0.0
    for _ in range(reps):
        t = time.perf_counter()
        c.encode_systematic_sharded(msg, par, None, L, accumulate=False)
        t_enc += time.perf_counter() - t
        rx = cw.copy().reshape(n, L)
        rx[pos] = 0  # the reference's contract: erased shards arrive zeroed
        rx = rx.reshape(-1)
        t = time.perf_counter()
        ok = c.recover_erasure_shards(rx, None, L, pos)
        t_rec += time.perf_counter() - t
        assert ok
    assert np.array_equal(rx, cw)
    enc_b, rec_b = n * L, (n + e) * L
    big = 256 << 20
    h = torch.empty(big, dtype=torch.uint8, pin_memory=True)
    d = torch.empty(big, dtype=torch.uint8, device="cuda")
    d.copy_(h, non_blocking=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(3):
        d.copy_(h, non_blocking=True)
    torch.cuda.synchronize()
    h2d = 3 * big / (time.perf_counter() - t) / 1e9
    t = time.perf_counter()
    for _ in range(3):
        h.copy_(d, non_blocking=True)
    torch.cuda.synchronize()
    d2h = 3 * big / (time.perf_counter() - t) / 1e9
    del h, d
    return {"workload": f"host slices (n={n},k={k}), L={L}, e={e}, pageable numpy buffers",
            "encode_GBps": round(enc_b * reps / t_enc / 1e9, 2),
            "recover_GBps": round(rec_b * reps / t_rec / 1e9, 2),
            "encode_ms": round(t_enc / reps * 1e3, 2), "recover_ms": round(t_rec / reps * 1e3, 2),
            "pcie_bytes_encode": n * L, "pcie_bytes_recover": (n + e) * L,
            "pinned_h2d_GBps": round(h2d, 1), "pinned_d2h_GBps": round(d2h, 1)}


def bench_decode(ars, torch, np, n, k, L, reps=5, seed=47):
    """Sharded error decoding of the same (n, k) codeword (formal derivative):
    a clean scrub at shard_len L (syndrome pass only), 1 % of the columns with
    <= T/2 errors each at L, and whole-shard corruption (4 shards, so every
    column needs the host key equation) at L/16.  Wall-clock per call
    (the call is synchronous); bytes counted as n*L read."""
    T = n - k
    c = ars.Codec(n, k)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    msg = torch.randint(0, 256, (k * L,), dtype=torch.uint8, device="cuda", generator=g)
    par = torch.zeros(T * L, dtype=torch.uint8, device="cuda")
    c.encode_systematic_sharded(msg, par, None, L, accumulate=False)
    cw = torch.cat([par, msg]).view(n, L)
    rng = np.random.default_rng(seed)

    def timed(rx_fn, length):
        best = None
        for _ in range(reps):
            rx = rx_fn()
            torch.cuda.synchronize()
            t = time.perf_counter()
            ok, counts = c.decode_systematic_sharded(rx.view(-1), length, formal_derivative=True)
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        return best, ok, counts

    t_clean, ok, counts = timed(lambda: cw.clone(), L)
    assert ok and counts == (L, 0, 0)
    cols = torch.from_numpy(np.sort(rng.permutation(L)[:L // 100])).cuda()
    sub = cw[:, cols].cpu().numpy()
    for j in range(sub.shape[1]):
        ne = int(rng.integers(1, T // 2 + 1))
        sub[rng.permutation(n)[:ne], j] ^= rng.integers(1, 256, ne, dtype=np.uint8)
    bad = torch.from_numpy(sub).cuda()

    def sparse():
        rx = cw.clone()
        rx[:, cols] = bad
        return rx

    t_sparse, ok, counts = timed(sparse, L)
    assert ok and counts[1] == len(cols)
    Ld = L // 16
    cwd = cw[:, :Ld].contiguous()
    shards = torch.from_numpy(rng.permutation(n)[:4]).cuda()
    noise = torch.randint(1, 256, (4, Ld), dtype=torch.uint8, device="cuda", generator=g)

    def dense():
        rx = cwd.clone()
        rx[shards] ^= noise
        return rx

    t_dense, ok, counts = timed(dense, Ld)
    assert ok and counts == (0, Ld, 0)
    return {"workload": f"(n={n},k={k}) formal-derivative decode",
            "clean_scrub": {"shard_len": L, "ms": round(t_clean * 1e3, 3),
                            "GBps": round(n * L / t_clean / 1e9, 1)},
            "sparse_1pct_columns": {"shard_len": L, "bad_columns": int(len(cols)),
                                    "ms": round(t_sparse * 1e3, 3),
                                    "GBps": round(n * L / t_sparse / 1e9, 1)},
            "dense_4_shards": {"shard_len": Ld, "bad_columns": Ld, "ms": round(t_dense * 1e3, 3),
                               "columns_per_s": round(Ld / t_dense, 1),
                               "GBps": round(n * Ld / t_dense / 1e9, 2)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-decode", action="store_true")
    ap.add_argument("--no-grid", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-wide", action="store_true")
    ap.add_argument("--summarize-rocprof", metavar="DIR",
                    help="write profiles/ from the rocprofv3 CSV directories under DIR and exit")
    ap.add_argument("--print-source-sha", action="store_true", help="print the device-source SHA and exit")
    ap.add_argument("--phaseb", action="store_true",
                    help="run only the bit-sliced phase-B prototype against the v_perm form and print its JSON")
    ap.add_argument("--print-profile-commands", metavar="DIR",
                    help="print the rocprofv3 command of each profile pass, writing under DIR, and exit")
    args = ap.parse_args()
    if args.print_source_sha:
        print(device_source_sha())
        return
    if args.print_profile_commands:
        for name in PROFILE_RUNS:
            print(profile_command(name, args.print_profile_commands, ROOT))
        return
    if args.summarize_rocprof:
        summarize_rocprof(args.summarize_rocprof)
        return
    if args.phaseb:
        import torch
        print(json.dumps(bench_phaseb(torch)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); every figure is "
              "per-GPU and the path shards by byte-column range with no exchange.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    ars = load_pkg()
    n, k, L = 256, 128, 1 << 20
    T = n - k
    e = T
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(43)
    cw = torch.empty(n * L, dtype=torch.uint8, device=dev)
    cw[T * L:] = torch.randint(0, 256, (k * L,), dtype=torch.uint8, device=dev, generator=gen)
    par, msg = cw[:T * L], cw[T * L:]
    pos = np.sort(np.random.default_rng(43).permutation(n)[:e]).astype(np.uint8)
    ctx = ars.default_context(0)
    stream = torch.cuda.current_stream()
    # the autotuner picks the tile variant for this shape (timed on scratch)
    tune_enc, tune_rec = ctx.autotune(n, k, L, stream)

    def encode():
        ctx.encode(n, k, msg.data_ptr(), par.data_ptr(), L, 0, stream)

    def recover():
        ctx.recover(n, k, cw.data_ptr(), L, pos, stream)

    # correctness gate (untimed): encode, erase, recover, compare
    encode()
    ref = cw.clone()
    cwv = cw.view(n, L)
    cwv[torch.from_numpy(pos.astype(np.int64)).to(dev)] = 0
    recover()
    torch.cuda.synchronize()
    if not torch.equal(cw, ref):
        print("bench: recovered codeword differs from the encoded one", file=sys.stderr)
        sys.exit(1)

    for _ in range(args.warmup):
        encode()
        recover()
    torch.cuda.synchronize()

    # whole step: encode + recover (per-call erasure plan included)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t0.record(stream)
    for _ in range(args.steps):
        encode()
        recover()
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_step = t0.elapsed_time(t1) / args.steps
    step_bytes = n * L + (n + e) * L
    value = step_bytes / (ms_step * 1e-3) / 1e9

    # per-kernel timing on the launch stream: encode kernel alone, and the
    # fused recover kernel alone (erasure plan built once, outside)
    plan = torch.empty(ars.lib().ars_recover_plan_size(), dtype=torch.uint8, device=dev)
    ctx.recover_plan(n, k, pos, plan.data_ptr(), stream)

    def timed(fn, reps, windows=3):
        """Seconds per launch: the median over `windows` windows of `reps`
        back-to-back launches, each bracketed by events on the launch stream."""
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        fn()
        torch.cuda.synchronize()
        per = []
        for _ in range(windows):
            a.record(stream)
            for _ in range(reps):
                fn()
            b.record(stream)
            torch.cuda.synchronize()
            per.append(a.elapsed_time(b) / reps * 1e-3)
        return sorted(per)[len(per) // 2]

    reps = max(args.steps, 20)
    s_enc = timed(encode, reps)
    s_rec = timed(lambda: ctx.recover_with_plan(n, k, cw.data_ptr(), L, plan.data_ptr(), stream), reps)
    # attainable HBM bandwidth: a plain device-to-device copy of the same
    # size as the recover's input (read + write counted)
    src = cw
    dst = torch.empty_like(cw)
    s_copy = timed(lambda: dst.copy_(src), reps)
    copy_gbps = 2 * n * L / s_copy / 1e9
    del dst
    enc_gbps = n * L / s_enc / 1e9
    rec_gbps = (n + e) * L / s_rec / 1e9
    # the tile variants the autotuner kept (quads per workgroup, LDS buffers)
    # (names as rocprofv3 prints them: aligned 4-byte buffers, access mode 1)
    enc_name = f"k_encode_tile<7, {tune_enc['quads']}, 1, false, {tune_enc['buffers']}, true>"
    rec_name = (f"k_recover_stream<8, {tune_rec['quads']}, true, 1>" if tune_rec["buffers"] == 1
                else f"k_recover_tile<8, {tune_rec['quads']}, true, {tune_rec['buffers']}>")
    if s_rec >= s_enc:
        dom, dom_bytes, dom_s = rec_name, (n + e) * L, s_rec
    else:
        dom, dom_bytes, dom_s = enc_name, n * L, s_enc
    achieved = dom_bytes / dom_s / 1e9
    traffic = measured_traffic(dom)
    _, stat = _profile_entry("kernel_stats_256x128_1MiB.json", dom)
    valu_rates = probe_valu_rates()
    phaseb = bench_phaseb(torch)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (uniform random bytes, seed 43; erasures: 128 seeded uniform positions)",
        "config": {"workload": "single codeword (n=256,k=128), 1 MiB shards: encode 128 parity "
                               "shards then recover 128 random erasures",
                   "n": n, "k": k, "shard_len": L, "erasures": e, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": traffic, "algorithmic_bytes": dom_bytes, "launch_us": round(dom_s * 1e6, 2),
                     "rocprof_avg_us": stat["avg_us"] if stat else None,
                     # the same kernel timed with HIP events in the profiled run itself: rocprof's
                     # average is compared with this, since the clock under these kernels differs
                     # from box to box
                     "rocprof_run_launch_us": _profiled_run_launch_us("kernel_stats_256x128_1MiB.json", dom)},
        "valu_roofline": valu_roofline(dom, dom_s, valu_rates),
        "valu_issue_rates": {k: round(v / 1e12, 3) for k, v in valu_rates.items()} if valu_rates else None,
        "encode": {"GBps": round(enc_gbps, 1), "us": round(s_enc * 1e6, 2),
                   "frac_hbm": round(enc_gbps / HBM_PEAK_GBPS, 4), "bytes": n * L,
                   "valu_roofline": valu_roofline(enc_name, s_enc, valu_rates)},
        "recover": {"GBps": round(rec_gbps, 1), "us": round(s_rec * 1e6, 2),
                    "frac_hbm": round(rec_gbps / HBM_PEAK_GBPS, 4), "bytes": (n + e) * L,
                    "valu_roofline": valu_roofline(rec_name, s_rec, valu_rates)},
        "autotune": {"encode": tune_enc, "recover": tune_rec},
        "phaseb_prototype": phaseb,
        "copy_reference": {"GBps": round(copy_gbps, 1), "bytes_read_plus_written": 2 * n * L,
                           "frac_hbm": round(copy_gbps / HBM_PEAK_GBPS, 4)},
        "wall_s_timed_region": round(wall, 4),
    }
    if not args.no_batched:
        del plan
        out["batched"] = bench_batched(ars, torch, np, stream)
    if not args.no_served:
        out["served"] = bench_served(ars, torch, np)
        out["served_small"] = bench_served_small(ars, torch, np)
    if not args.no_host:
        out["host_slices"] = bench_host(ars, torch, np, n, k, L, e)
        print(f"host slices: {out['host_slices']}", file=sys.stderr, flush=True)
    if not args.no_decode:
        out["decode"] = bench_decode(ars, torch, np, n, k, L)
        print(f"decode: {out['decode']}", file=sys.stderr, flush=True)
    if not args.no_grid:
        out["reference_grid"] = bench_reference_grid(ars, torch, np, stream, with_cpu=not args.no_cpu)
        print(f"reference grid: {out['reference_grid']['median_cpu_over_gpu']}", file=sys.stderr, flush=True)
    if not args.no_c5:
        out["c5"] = bench_c5(ars, torch, np, stream)
        print(f"c5: {out['c5']}", file=sys.stderr, flush=True)
    if not args.no_wide:
        out["wide"] = bench_wide(ars, torch, np, stream)
        print(f"wide: {out['wide']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, k, L, e, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
