#!/usr/bin/env python3
"""Benchmark: k-mers/s counted bit-exact on one MI355X.

Workload (default): a seeded 5 Mb genome at 50% GC, 2x150 bp pairs at 100x
(1,666,667 pairs), substitution rate rising linearly 0.1% -> 1.0% along each
read; exact counting at k = 21, 33 and 55 with the default 2^29-bit Bloom tiers.
One step = Stage 2 + Stage 3 (Bloom build, count table, sort, cutoff, solid set)
at all three k over the reads already resident in HBM.

Prints one JSON line (see README of the driver contract) with a `roofline`
object for the dominant kernel (HIP events on the engine's stream) and a
`cpu_baseline` object (the oracle's multi-thread port on a bounded sample of
the same reads, timed on this host's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "spades-rs_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
G_BYTES = 64  # one random 64 B access per table RMW
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "hbm_traffic_5mb_100x.json")
COPY_BYTES = 1 << 30  # the calibration copy that --copy-calibration runs first


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="5mb_100x",
                    choices=["5mb_100x", "1mb_30x", "64lib", "learned", "ingest", "unitigs", "meta"])
    ap.add_argument("--meta-scale", type=float, default=1.0,
                    help="meta: genome sizes and pairs x this (1.0 = 20 genomes of 1-6 Mb, 30M pairs)")
    ap.add_argument("--train-seeds", type=int, default=8, help="learned: training seeds (held-out seeds follow)")
    ap.add_argument("--test-seeds", type=int, default=2)
    ap.add_argument("--margin", type=float, default=0.05, help="learned: confidence margin of the first pass")
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--scale", type=float, default=0.1,
                    help="64lib: genome-size scale (1.0 = 5 kb-6 Mb genomes, ~5.4 Gb of reads)")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner on each engine before the timed region (the defaults are its usual choice)")
    ap.add_argument("--no-microbench", dest="microbench", action="store_false",
                    help="skip the random-atomic / copy microbenchmarks")
    ap.add_argument("--graph", action="store_true", help="64lib: value from graph replays instead of eager runs")
    ap.add_argument("--ks", default="21,33,55")
    ap.add_argument("--error-correct", action="store_true",
                    help="run Hammer before counting (AssemblerConfig::error_correct)")
    ap.add_argument("--bloom-bits", type=int, default=512 * 1024 * 1024)
    ap.add_argument("--cpu-sample-reads", type=int, default=100_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--table-load", type=float, default=0.0)
    ap.add_argument("--insert-path", type=int, default=0,
                    help="Stage 2 count: 0 or 3 minimizer-binned where it applies, 1 table atomics")
    ap.add_argument("--fasta", action="store_true", help="ingest: the pairs as multi-line FASTA files")
    ap.add_argument("--sequential", action="store_true",
                    help="default workloads: run the per-k contexts one after another instead of together "
                         "(kc_run_many)")
    ap.add_argument("--blocks-per-cu", type=int, default=0,
                    help="workgroups per CU of the table insert / bin count (0: default shape)")
    ap.add_argument("--setter-path", type=int, default=0,
                    help="Bloom first setters: 0 sorted by partition with minima in LDS where it applies, 1 atomics")
    ap.add_argument("--microbench-sweep", action="store_true",
                    help="only print random-atomic / random-read rates against array size")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy first (scale for rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.microbench_sweep:
        return microbench_sweep()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design note only "
                                   "(--gpus must be 1)"}), flush=True)
        sys.exit(2)

    import numpy as np
    import torch
    from kc import AssemblerConfig, SpectrumEngine, synth

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        sys.exit(1)
    ks = [int(x) for x in args.ks.split(",")]
    if args.workload == "64lib":
        return bench_many_libraries(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "ingest":
        return bench_ingest(args)
    if args.workload == "unitigs":
        return bench_unitigs(args)
    if args.workload == "meta":
        return bench_meta(args)
    if args.copy_calibration:
        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t0 = time.time()
    if args.workload == "5mb_100x":
        store, meta = synth.workload_5mb_100x(seed=2)
        wdesc = "5 Mb genome (50% GC), 2x150 bp at 100x, errors 0.1%->1.0% along the read"
    else:
        store, meta = synth.workload_1mb_30x(seed=1)
        wdesc = "1 Mb genome, 2x150 bp at 30x, 0.5% uniform errors"
    log(f"generated {len(store)} reads in {time.time() - t0:.1f}s")

    engines = []
    upload_ms = 0.0
    for k in ks:
        eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=args.bloom_bits, error_correct=args.error_correct),
                             table_load=args.table_load, timing=True)
        if args.insert_path or args.setter_path or args.blocks_per_cu:
            eng.set_tune(insert_path=args.insert_path, setter_path=args.setter_path, blocks_per_cu=args.blocks_per_cu)
        u0 = time.perf_counter()
        eng.upload(store)  # host -> device copy of the packed reads (outside the timed region)
        upload_ms += (time.perf_counter() - u0) * 1e3
        if args.autotune:
            log(f"k={k} tune: {eng.autotune(reps=1)}")
        engines.append(eng)
    log("uploaded")

    from kc import run_many

    def step_sequential():
        return [e.run() for e in engines]

    def step_concurrent():
        return run_many(engines)

    # one context per k over the same resident reads; without prior contigs the ks are
    # independent, so by default they run together (kc_run_many: a host thread and stream
    # each) and a step ends when all three have
    step = step_sequential if args.sequential else step_concurrent
    cold_ms = None
    for i in range(args.warmup):
        c0 = time.perf_counter()
        infos = step()
        if i == 0:
            cold_ms = (time.perf_counter() - c0) * 1e3
        log(f"warmup {i}: " + ", ".join(f"k{k}: {inf['ms_stage2'] + inf['ms_stage3']:.1f} ms"
                                        for k, inf in zip(ks, infos)))
    torch.cuda.synchronize()
    kernel_ms = {}
    step_ms = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        infos = step()
        step_ms.append((time.perf_counter() - s0) * 1e3)
        for k, e in zip(ks, engines):
            for name, (ms, n) in e.kernel_times().items():
                acc = kernel_ms.setdefault((k, name), [0.0, 0])
                acc[0] += ms
                acc[1] += n
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    n_occ = sum(inf["n_occ"] for inf in infos)
    ms_per_step = elapsed * 1e3 / args.steps
    value = n_occ / (ms_per_step / 1e3)
    log(f"steps: {', '.join(f'{x:.1f}' for x in step_ms)} ms")
    seq_ms = ms_per_step if args.sequential else None
    if not args.sequential:
        # Kernels of contexts running together share the device, so their event times are
        # not one kernel's own duration: the per-kernel figures (and the dominant kernel's
        # roofline) come from the same number of steps run one context after another.
        kernel_ms = {}
        for e in engines:
            e.kernel_times()
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        for i in range(args.steps):
            step_sequential()
            for k, e in zip(ks, engines):
                for name, (ms, n) in e.kernel_times().items():
                    acc = kernel_ms.setdefault((k, name), [0.0, 0])
                    acc[0] += ms
                    acc[1] += n
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - s0) * 1e3 / args.steps
        log(f"sequential steps: {seq_ms:.2f} ms a step")

    # dominant kernel and its roofline.  Algorithmic bytes per unit of work: one k-mer
    # occurrence moves s_k packed-stream bytes plus one random 64 B count-table access
    # (the counting kernels); one distinct k-mer touches 3 seen and 3 solid Bloom words
    # (the first-setter kernels).  "irreducible_bytes" is what the kernel cannot avoid
    # moving (the packed reads once and the distinct entries written out).
    tot = sorted(((v[0], k, name, v[1]) for (k, name), v in kernel_ms.items()), reverse=True)
    for ms, k, name, n in tot[:40]:
        log(f"  k={k:3d} {name:18s} {ms / max(n, 1):9.3f} ms/launch x{n}")
    dom_ms, dom_k, dom_name, dom_n = tot[0]
    per_k = {k: inf for k, inf in zip(ks, infos)}
    lens = store.lengths.astype(np.int64)

    def stream_bytes(k):
        return int(((lens + 3) // 4)[lens >= k].sum())

    inf = per_k[dom_k]
    per_launch_s = dom_ms / dom_n / 1e3
    slot = 16 if dom_k <= 31 else 32  # table slot bytes (one / two key words, 32-bit events)
    setter_kernels = ("k_first_setter", "fs_sort", "k_fs_items", "k_fs_min", "k_promote", "k_seen_from_fs")
    if dom_name == "k_hammer":
        # one random 64 B access of the solid tier per is_solid test (at least its first word)
        alg = inf["hammer_queries"] * G_BYTES
        irreducible = stream_bytes(dom_k) * 2
        unit = "is_solid test: one random 64 B solid-tier access"
    elif dom_name in setter_kernels:
        alg = inf["n_distinct"] * 6 * G_BYTES
        irreducible = inf["n_distinct"] * (slot + 2 * 8)
        unit = "distinct k-mer: 3 seen + 3 solid Bloom words (6 x 64 B)"
    else:
        alg = stream_bytes(dom_k) + inf["n_occ"] * G_BYTES
        irreducible = stream_bytes(dom_k) + inf["n_distinct"] * slot
        unit = "k-mer occurrence: s_k packed bytes + one 64 B count-table access"
    achieved = alg / per_launch_s / 1e9
    roof = {"bound": "hbm", "kernel": f"{dom_name} (k={dom_k})", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "algorithmic_bytes_per_launch": alg, "bytes_unit": unit, "avg_launch_ms": round(dom_ms / dom_n, 4),
            "irreducible_bytes_per_launch": irreducible,
            "irreducible_GBs": round(irreducible / per_launch_s / 1e9, 1),
            "windows_per_s": round(inf["n_occ"] / per_launch_s, 1)}
    if args.error_correct:
        hk = kernel_ms.get((dom_k, "k_hammer"))
        if hk:
            roof["hammer"] = {"k": dom_k, "ms_per_launch": round(hk[0] / hk[1], 4),
                              "queries_per_launch": inf["hammer_queries"],
                              "queries_per_s": round(inf["hammer_queries"] / (hk[0] / hk[1] / 1e3), 1),
                              "corrected_reads": inf["corrected_reads"]}
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated),
    # used only while the recorded launch time still matches this kernel's
    rec = None
    if args.workload == "5mb_100x" and os.path.exists(TRAFFIC_JSON):
        rec = json.load(open(TRAFFIC_JSON)).get(f"{dom_name}:{dom_k}")
    if rec and "bytes_per_launch" in rec and abs(rec["kernel_ms"] - dom_ms / dom_n) <= 0.1 * rec["kernel_ms"]:
        roof["traffic"] = rec["bytes_per_launch"]
        roof["traffic_over_algorithmic"] = round(rec["bytes_per_launch"] / alg, 2)
        roof["traffic_over_irreducible"] = round(rec["bytes_per_launch"] / irreducible, 2)
    # whole-path roofline with the design-independent byte count of the path
    path_b = 0
    for k in ks:
        inf = per_k[k]
        path_b += stream_bytes(k) + inf["n_occ"] * G_BYTES + inf["n_distinct"] * 6 * G_BYTES + inf["n_counts"] * 20
    path = {"bytes_per_step": path_b, "achieved_GBs": round(path_b / (ms_per_step / 1e3) / 1e9, 1),
            "frac": round(path_b / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}

    # bit-exactness of the timed configuration: digests of every output of the last step
    # against the sequential oracle's on the same reads (tests/golden/full_size_digests.json)
    exact = verify_digests(args, engines, ks, infos)
    log(f"exact: {exact}")

    micro = None
    if args.microbench:
        micro = microbenchmarks()
        log(f"microbenchmarks: {micro}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, store, ks)

    out = {
        "metric": "k-mers/s counted bit-exact (Bloom tiers + sorted count table + cutoff + solid set)",
        "value": round(value, 1), "unit": "k-mers/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic (seeded)",
        "exact": exact["all"] if exact else None,
        "config": {"workload": wdesc, "reads": len(store), "read_len": 150, "ks": ks, "bloom_bits": args.bloom_bits,
                   "error_correct": bool(args.error_correct),
                   "n_occ": n_occ, "per_k": {str(k): {"n_occ": per_k[k]["n_occ"], "n_distinct": per_k[k]["n_distinct"],
                                                      "n_counts": per_k[k]["n_counts"], "cutoff": per_k[k]["cutoff"],
                                                      "n_solid": per_k[k]["n_solid"],
                                                      "ms": round(per_k[k]["ms_stage2"] + per_k[k]["ms_stage3"], 3)}
                                             for k in ks},
                   "cold_first_step_ms": round(cold_ms, 3) if cold_ms is not None else None,
                   "schedule": "sequential contexts" if args.sequential else
                               "one context per k, run together (kc_run_many)",
                   "ms_per_step_sequential": round(seq_ms, 3),
                   "upload_ms": round(upload_ms, 3),
                   "value_with_upload": round(n_occ / ((ms_per_step + upload_ms) / 1e3), 1)},
        "roofline": roof, "path_roofline": path, "cpu_baseline": cpu, "exact_detail": exact,
        "tune": {str(k): e.tune() for k, e in zip(ks, engines)},
        "microbench": micro,
    }
    if cpu:
        out["gpu_vs_cpu"] = round(value / cpu["value"], 1)
    for e in engines:
        e.close()
    print(json.dumps(out), flush=True)


def verify_digests(args, engines, ks, infos):
    """Per k: the engine's last outputs against the recorded oracle digests of the same
    workload, k and tier size; None when no digest is recorded for the configuration."""
    from kc.digest import diff, engine_digests
    path = os.path.join(ROOT, "tests", "golden", "full_size_digests.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    res = {}
    for k, e, inf in zip(ks, engines, infos):
        want = rec.get(f"{args.workload}_k{k}" + ("_hammer" if args.error_correct else ""))
        if not want or want["bloom_bits"] != args.bloom_bits:
            res[str(k)] = None
            continue
        bad = diff(engine_digests(e, inf, with_reads=args.error_correct), want)
        res[str(k)] = not bad
        if bad:
            log(f"k={k}: outputs differ from the oracle in {bad}")
    checked = [v for v in res.values() if v is not None]
    res["all"] = bool(checked) and all(checked) if len(checked) == len(ks) else None
    return res


def cpu_info():
    """Host CPU as the run saw it: nproc, the CPUs this process may use, the model name."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


def cpu_baseline(args, store, ks):
    """The oracle's multi-thread mode (the reference's rayon chunking and sharded maps) on
    a bounded sample of the same reads: one warm-up, then the median of 3 runs, Stage 2+3
    only, on every CPU this process may use (at most 16, the host share of one GPU here);
    the single-thread sequential oracle (the reference at -t 1) on a smaller sample beside it."""
    from oracle import oracle as O
    hw = cpu_info()
    threads = args.cpu_threads or min(hw["usable"], 16)
    d, o, l = store.arrays()

    def sample(n):
        n = min(n, len(store))
        end = int(o[n - 1]) + (int(l[n - 1]) + 3) // 4
        return n, d[:end], o[:n], l[:n]

    def timed(n, th):
        n, dd, oo, ll = sample(n)
        occ, secs = 0, 0.0
        for k in ks:
            S = O.Spectrum(dd, oo, ll, k, args.bloom_bits, bool(args.error_correct), 5.0, th)
            occ += S.n_occ
            secs += S.stage2_secs + S.stage3_secs
        return occ, secs

    n_mt = args.cpu_sample_reads
    timed(n_mt, threads)  # warm-up
    runs = [timed(n_mt, threads) for _ in range(3)]
    occ = runs[0][0]
    med = float(np.median([s for _, s in runs]))
    n_1 = max(n_mt // 5, 1)
    occ1, s1 = timed(n_1, 1)
    log(f"cpu: {threads} threads {occ} windows, runs {[round(s, 2) for _, s in runs]} s; 1 thread {occ1} in {s1:.2f} s")
    return {"value": round(occ / med, 1), "unit": "k-mers/s", "cores": threads, "kind": "port",
            "sample": f"first {min(n_mt, len(store))} reads of the same store, k={','.join(map(str, ks))}, "
                      f"{args.bloom_bits}-bit tiers, Hammer {'on' if args.error_correct else 'off'}, Stage 2+3 of the "
                      f"oracle's multi-thread mode (reference "
                      f"chunking + sharded maps); median of 3 runs after one warm-up",
            "runs_s": [round(s, 3) for _, s in runs], "host": hw,
            "single_thread": {"value": round(occ1 / s1, 1), "unit": "k-mers/s", "cores": 1,
                              "sample": f"first {min(n_1, len(store))} reads, sequential oracle (-t 1 order)"}}


def microbenchmarks(device=0):
    """Measured ceilings (libkc_bench.so): random non-returning atomics (u32 add, u64 or)
    and random 8-byte reads over a 2 GiB array (the size of a 2^29-entry u32 array), and a
    4 GiB streaming copy."""
    import ctypes
    path = os.path.join(PKG, "libkc_bench.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.kc_bench_device.restype = ctypes.c_int
    L.kc_bench_device.argtypes = [ctypes.c_int32, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double)]
    out = {}
    n_ops = 1 << 29
    for name, kind, nbytes in (("atomic_add_u32", 0, 2 << 30), ("atomic_or_u64", 1, 2 << 30),
                               ("read_64b", 2, 2 << 30), ("copy", 3, 4 << 30)):
        ms = ctypes.c_double(0)
        rc = L.kc_bench_device(device, kind, nbytes, n_ops, 3, ctypes.byref(ms))
        if rc != 0:
            out[name] = {"error": rc}
            continue
        if kind == 3:
            out[name] = {"GB_s": round(2 * nbytes / (ms.value / 1e3) / 1e9, 1), "ms": round(ms.value, 3)}
        else:
            out[name] = {"G_ops_s": round(n_ops / (ms.value / 1e3) / 1e9, 2), "ms": round(ms.value, 3),
                         "array_bytes": nbytes}
    return out


def microbench_sweep(device=0, n_ops=1 << 28):
    """Random-access rates against array size (1 MiB to 8 GiB): non-returning u32 atomicAdd
    at agent and at workgroup scope, returning u32 atomicAdd, and 8-byte random reads.
    Prints one JSON object per size."""
    import ctypes
    L = ctypes.CDLL(os.path.join(PKG, "libkc_bench.so"))
    L.kc_bench_device.restype = ctypes.c_int
    L.kc_bench_device.argtypes = [ctypes.c_int32, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double)]
    kinds = (("atomic_add_agent", 0), ("atomic_add_workgroup", 4), ("atomic_add_returning", 5), ("read_8b", 2))
    for mib in (1, 4, 16, 64, 128, 256, 512, 2048, 8192):
        row = {"array_mib": mib}
        for name, kind in kinds:
            ms = ctypes.c_double(0)
            rc = L.kc_bench_device(device, kind, mib << 20, n_ops, 3, ctypes.byref(ms))
            row[name + "_G_per_s"] = round(n_ops / (ms.value / 1e3) / 1e9, 2) if rc == 0 else f"error {rc}"
        print(json.dumps(row), flush=True)


def _key_index(keys):
    """Rows of (n, 4) k-mer limbs as one sortable 1-D array (k <= 63: limbs 2, 3 are 0)."""
    a = np.ascontiguousarray(keys[:, :2], dtype=np.uint64)
    return a.view([("lo", "<u8"), ("hi", "<u8")]).reshape(-1)


def learned_seed(seed, scale, k):
    """One repeat-rich library: the exact Hammer-on selection (the judge) and the
    uncorrected table's features.  Returns a dict of arrays and timings."""
    from kc import AssemblerConfig, SpectrumEngine, synth
    g = max(int(12_000_000 * scale), 20_000)
    store, quals, meta = synth.repeat_pairs_with_quals(seed, genome_size=g, coverage=60.0)
    ex = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=1 << 29, error_correct=True))
    ex.upload(store)
    ex.run()
    info_ex = ex.run()  # warm
    kex, _, mex = ex.counts()
    judge = np.sort(_key_index(kex[mex.astype(bool)]))
    ex.close()
    pl = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=1 << 29))
    pl.upload(store)
    pl.run()
    info_pl = pl.run()
    pl.set_quals(quals, meta["qoff"])
    keys, cnt, mask = pl.counts()
    X = pl.features()
    y = np.isin(_key_index(keys), judge)
    return {"engine": pl, "X": X, "y": y, "cnt": cnt, "cutoff": info_pl["cutoff"], "cutoff_exact": info_ex["cutoff"],
            "top": info_pl["top_cov"],
            "judge": judge.shape[0],
            "judge_outside": int(judge.shape[0] - y.sum()), "t_exact": info_ex["ms_stage3"],
            "t_plain": info_pl["ms_stage3"], "reads": len(store), "n_occ": info_pl["n_occ"]}


def bench_learned(args):
    """Learned first pass versus the exact Hammer-on solid-k-mer selection on repeat-rich
    genomes (20% repeats, 60x, 0.3% errors, qualities), k=21: trained on --train-seeds
    seeds, judged on the next --test-seeds seeds."""
    import torch
    from kc.learn import Mlp
    k = 21
    scale = args.scale
    Xs, ys = [], []
    for s in range(args.train_seeds):
        d = learned_seed(1000 + s, scale, k)
        Xs.append(d["X"])
        ys.append(d["y"])
        d["engine"].close()
        log(f"train seed {s}: {d['X'].shape[0]} entries, {d['y'].mean():.3f} solid after correction")
    X, y = np.concatenate(Xs), np.concatenate(ys).astype(np.uint8)
    m = Mlp(seed=1)
    t0 = time.perf_counter()
    loss = m.train(X, y, epochs=args.epochs, batch=4096, lr=2e-3, seed=7)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"trained on {X.shape[0]} rows in {train_s:.1f}s, last loss {loss:.4f}")
    res = []
    for s in range(args.test_seeds):
        d = learned_seed(5000 + s, scale, k)
        eng = d["engine"]
        lab, fb, ms = eng.learned_select(m, args.margin)
        ms_best = min([ms] + [eng.learned_select(m, args.margin)[2] for _ in range(2)])
        plain = d["cnt"] >= d["cutoff"]
        # Library-level margin: a fragile cutoff (kc_library_cutoff_fragile) sends the
        # whole library to the exact selection.
        lib_fb = eng.cutoff_fragile()
        final = d["y"] if lib_fb else lab
        cost = d["t_plain"] + (d["t_exact"] if lib_fb else ms_best)
        res.append({"entries": int(lab.shape[0]), "agreement": float(np.mean(lab == d["y"])),
                    "library_fallback": bool(lib_fb), "agreement_after_fallback": float(np.mean(final == d["y"])),
                    "end_to_end_ms": round(cost, 3),
                    "plain_rule_agreement": float(np.mean(plain == d["y"])),
                    "fallback_fraction": float(fb.mean()), "residual_mismatches": int(np.sum(lab != d["y"])),
                    "judge_solid": d["judge"], "judge_solid_outside_table": d["judge_outside"],
                    "exact_selection_ms": round(d["t_exact"], 3), "learned_pass_ms": round(ms_best, 3),
                    "plain_rule_selection_ms": round(d["t_plain"], 3),
                    "cutoff_uncorrected": int(d["cutoff"]), "cutoff_exact": int(d["cutoff_exact"]),
                    "rule_at_exact_cutoff_agreement": float(np.mean((d["cnt"] >= d["cutoff_exact"]) == d["y"])),
                    "speedup_vs_exact": round(d["t_exact"] / ms_best, 3)})
        eng.close()
        log(f"test seed {s}: {res[-1]}")
    m.close()
    agree = float(np.mean([r["agreement_after_fallback"] for r in res]))
    speed = sum(r["exact_selection_ms"] for r in res) / sum(r["end_to_end_ms"] for r in res)
    out = {
        "metric": "learned first pass vs exact Hammer-on solid-k-mer selection (speedup at label agreement)",
        "value": round(speed, 3), "unit": "x",
        "n_gpus": 1, "higher_is_better": True, "data": "synthetic (seeded)", "dtype": "bf16/fp32",
        "config": {"workload": f"repeat-rich genome 12 Mb x {scale}, 20% repeats, 60x, 0.3% errors, k=21",
                   "train_seeds": args.train_seeds, "test_seeds": args.test_seeds, "margin": args.margin,
                   "epochs": args.epochs, "train_rows": int(X.shape[0]), "train_s": round(train_s, 2)},
        "agreement": agree, "min_agreement": float(min(r["agreement_after_fallback"] for r in res)),
        "per_kmer_only_agreement": float(np.mean([r["agreement"] for r in res])),
        "library_fallbacks": int(sum(r["library_fallback"] for r in res)), "per_test_seed": res,
    }
    print(json.dumps(out), flush=True)


def bench_many_libraries(args):
    """64 independent libraries (uneven read lengths 50-250 bp, 5 kb-6 Mb genomes x scale,
    20-200x, 0.3% errors) counted at k=33 in one context: Stage 2 + 3 of every library per
    step.  value: eager runs; with --graph, HIP graph replays (the same binned count, its
    sizes planned by the capturing call and guarded on the device) -- the other mode is
    timed beside it."""
    import numpy as np
    import torch
    from kc import AssemblerConfig, SpectrumEngine, synth
    k = 33
    t0 = time.time()
    libs = synth.workload_64_libraries(
        seed=3, scale=args.scale,
        progress=lambda i, st: log(f"library {i}: {len(st)} reads ({time.time() - t0:.0f} s)") if i % 8 == 7 else None)
    reads = sum(len(st) for st, _ in libs)
    bases = sum(int(st.lengths.astype(np.int64).sum()) for st, _ in libs)
    log(f"generated {len(libs)} libraries, {reads} reads, {bases} bases in {time.time() - t0:.1f}s")
    eager = not args.graph
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=args.bloom_bits), timing=eager)
    for st, _ in libs:
        eng.add_library(st)
    tune = eng.autotune(reps=1) if args.autotune else eng.tune()
    log(f"tune: {tune}")

    def timed(step):
        first = step()  # an eager run sizes its bins; a graph call plans and captures
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t_start) * 1e3 / args.steps, first

    ms, first = timed(eng.run if eager else eng.graph_run)
    snap = [eng.counts(i) for i in range(len(libs))]
    if eager:
        for name, (kms, n) in sorted(eng.kernel_times().items(), key=lambda x: -x[1][0])[:12]:
            log(f"  {name:18s} {kms / max(n, 1):9.3f} ms/launch x{n}")
    other_ms, _ = timed(eng.graph_run if eager else eng.run)
    same = all(all(np.array_equal(a, b) for a, b in zip(x, eng.counts(i))) for i, x in enumerate(snap))
    log(f"{'eager' if eager else 'graph'} {ms:.3f} ms, {'graph' if eager else 'eager'} {other_ms:.3f} ms, "
        f"same tables {same}")
    n_occ = sum(eng.info(i)["n_occ"] for i in range(len(libs)))
    n_counts = sum(eng.info(i)["n_counts"] for i in range(len(libs)))
    # exactness: the oracle's digests of the libraries recorded for this scale, and the
    # device closed-form checks (kc_library_verify) of every library
    from kc.digest import diff, engine_digests
    path = os.path.join(ROOT, "tests", "golden", "full_size_digests.json")
    rec = json.load(open(path)) if os.path.exists(path) else {}
    sampled = {}
    for i in range(len(libs)):
        want = rec.get(f"64lib_{args.scale}_lib{i}_k{k}")
        if want and want["bloom_bits"] == args.bloom_bits:
            sampled[str(i)] = not diff(engine_digests(eng, eng.info(i), library=i), want)
    verify_ok = 0
    for i in range(len(libs)):
        v = eng.verify(i)
        verify_ok += (v["seen_words_differing"] == 0 and v["solid_windows"] == v["sum_counts"]
                      and v["unordered_pairs"] == 0 and v["mask_mismatches"] == 0)
    log(f"oracle digests of sampled libraries: {sampled}; device checks hold for {verify_ok} of {len(libs)}")
    eng.close()
    out = {
        "metric": "k-mers/s counted bit-exact (Bloom tiers + sorted count table + cutoff + solid set)",
        "value": round(n_occ / (ms / 1e3), 1), "unit": "k-mers/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded)",
        "exact": (all(sampled.values()) and verify_ok == len(libs)) if sampled else None,
        "exact_sampled_libraries": sampled, "device_checks_hold": f"{verify_ok}/{len(libs)}",
        "config": {"workload": f"64 libraries, genomes 5 kb-6 Mb x {args.scale}, 20-200x, reads 50-250 bp "
                               f"(lognormal, median 150), 0.3% errors", "k": k, "reads": reads, "bases": bases,
                   "n_occ": n_occ, "n_counts": n_counts,
                   "mode": "eager (binned count)" if eager else "hip graph replay (binned count)",
                   ("graph_replay_ms_per_step" if eager else "eager_ms_per_step"): round(other_ms, 3),
                   "both_modes_same_tables": same,
                   "tune": tune, "first_call_ms": round(first["ms_stage2"] + first["ms_stage3"], 3)},
    }
    print(json.dumps(out), flush=True)


def bench_meta(args):
    """Metagenome-scale counting at k = 55: 20 genomes of 1-6 Mb with lognormal(sigma 1)
    abundances, 30M 2x150 bp pairs (x --meta-scale), 0.5% errors, one library, 2^29-bit
    tiers (saturated at this size), the 64-bit event path (5.76e9 windows).  Exactness:
    the oracle's digests when recorded for the scale (1/20), and at any scale the device
    closed-form checks of kc_library_verify (seen-tier recount, Bloom-solid windows =
    sum of counts, sorted table, solid flags) plus the table checksum equal on every run."""
    import torch
    from kc import AssemblerConfig, SpectrumEngine, synth
    k = 55
    t0 = time.time()
    store, meta = synth.workload_metagenome(
        seed=4, scale=args.meta_scale,
        progress=lambda g, n: log(f"genome {g}: {n} pairs drawn ({time.time() - t0:.1f}s)"))
    log(f"generated {len(store)} reads ({int(meta['genome_sizes'].sum())} genome bases) in {time.time() - t0:.1f}s")
    lens = store.lengths.astype(np.int64)
    stream = int(((lens + 3) // 4)[lens >= k].sum())
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=args.bloom_bits), timing=True)
    eng.upload(store)
    del store
    c0 = time.perf_counter()
    info = eng.run()
    cold_ms = (time.perf_counter() - c0) * 1e3
    free, total = torch.cuda.mem_get_info()
    log(f"cold run {cold_ms:.1f} ms: {info}; device memory in use {(total - free) / 2**30:.1f} GiB")
    v_first = eng.verify()
    log(f"verify: {v_first}")
    for _ in range(args.warmup):
        eng.run()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        info = eng.run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t_start) * 1e3 / args.steps
    kt = sorted(eng.kernel_times().items(), key=lambda x: -x[1][0])
    for name, (kms, n) in kt[:14]:
        log(f"  {name:18s} {kms / max(n, 1):9.3f} ms/launch x{n}")
    v_last = eng.verify()
    props = {"seen_tier_recount_equal": v_last["seen_words_differing"] == 0,
             "solid_windows_equal_sum_of_counts": v_last["solid_windows"] == v_last["sum_counts"],
             "table_strictly_ascending": v_last["unordered_pairs"] == 0,
             "solid_flags_consistent": v_last["mask_mismatches"] == 0,
             "same_table_every_run": v_last["checksum"] == v_first["checksum"] and v_last == v_first}
    exact = None
    path = os.path.join(ROOT, "tests", "golden", "full_size_digests.json")
    want = json.load(open(path)).get(f"meta_{args.meta_scale:g}_k{k}") if os.path.exists(path) else None
    if want and want["bloom_bits"] == args.bloom_bits:
        from kc.digest import diff, engine_digests
        exact = not diff(engine_digests(eng, info), want)
    dom = kt[0]
    dom_ms = dom[1][0] / dom[1][1]
    alg = stream + info["n_occ"] * G_BYTES
    out = {
        "metric": "k-mers/s counted bit-exact (Bloom tiers + sorted count table + cutoff + solid set)",
        "value": round(info["n_occ"] / (ms / 1e3), 1), "unit": "k-mers/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded)",
        "exact": exact, "properties": props, "all_properties_hold": all(props.values()),
        "config": {"workload": f"metagenome: 20 genomes of 1-6 Mb x {args.meta_scale:g}, lognormal abundances "
                               f"(sigma 1), {meta['n_pairs']} 2x150 bp pairs, 0.5% errors, k=55",
                   "k": k, "bloom_bits": args.bloom_bits, "genome_bases": int(meta["genome_sizes"].sum()),
                   "n_occ": info["n_occ"], "n_distinct": info["n_distinct"], "n_counts": info["n_counts"],
                   "n_solid": info["n_solid"], "cutoff": info["cutoff"], "cold_first_run_ms": round(cold_ms, 3),
                   "device_memory_in_use_GiB": round((total - free) / 2**30, 1), "verify": v_last},
        "roofline": {"bound": "hbm", "kernel": dom[0], "avg_launch_ms": round(dom_ms, 4),
                     "achieved": round(alg / (dom_ms / 1e3) / 1e9, 1) if dom[0] in ("k_bin_count", "k_insert")
                     else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / (dom_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                     if dom[0] in ("k_bin_count", "k_insert") else None,
                     "traffic": None, "algorithmic_bytes_per_launch": alg,
                     "bytes_unit": "k-mer occurrence: s_k packed bytes + one 64 B count-table access"},
    }
    eng.close()
    print(json.dumps(out), flush=True)


def fasta_text(store, n_pairs, read_len, seed, width=60):
    """The same pairs and N's as fastq_text, as two FASTA files: headers '>p<8 digits>/1'
    and '/2', sequences in lines of `width` bases."""
    files = []
    for f, tag in zip(fastq_text(store, n_pairs, read_len, seed), (b"1", b"2")):
        rows = f.reshape(n_pairs, -1)
        seq = rows[:, 13:13 + read_len]
        cuts = list(range(0, read_len, width))
        out = np.empty((n_pairs, 13 + read_len + len(cuts)), dtype=np.uint8)
        out[:, :13] = rows[:, :13]
        out[:, 0] = ord(">")
        o = 13
        for c in cuts:
            w = min(width, read_len - c)
            out[:, o:o + w] = seq[:, c:c + w]
            out[:, o + w] = 10
            o += w + 1
        files.append(out.reshape(-1))
    return files


def fastq_text(store, n_pairs, read_len, seed):
    """The fixed-length pairs of `store` as two FASTQ files (uint8 arrays): headers
    '@p<8 digits>/1' and '/2', Phred+33 qualities 'I', and, seeded, an internal N in 1 of
    100 R1 reads and three leading Ns in 1 of 100 R2 reads (both N gates take work)."""
    nb = (read_len + 3) // 4
    rng = np.random.default_rng(seed)
    lut = np.frombuffer(b"ACGT", dtype=np.uint8)
    files = []
    for half, tag in ((0, b"1"), (1, b"2")):
        d = store.data[half * n_pairs * nb:(half + 1) * n_pairs * nb].reshape(n_pairs, nb)
        codes = np.stack([(d >> 6) & 3, (d >> 4) & 3, (d >> 2) & 3, d & 3], axis=-1).reshape(n_pairs, 4 * nb)
        seq = lut[codes[:, :read_len]]
        pick = rng.random(n_pairs) < 0.01
        if half == 0:
            seq[pick, read_len // 2] = ord("N")
        else:
            seq[pick, :3] = ord("N")
        L = read_len
        out = np.empty((n_pairs, 17 + 2 * L), dtype=np.uint8)
        out[:, 0] = ord("@")
        out[:, 1] = ord("p")
        idx = np.arange(n_pairs)
        for dgt in range(8):
            out[:, 2 + dgt] = ord("0") + (idx // 10 ** (7 - dgt)) % 10
        out[:, 10] = ord("/")
        out[:, 11] = tag[0]
        out[:, 12] = 10
        out[:, 13:13 + L] = seq
        out[:, 13 + L] = 10
        out[:, 14 + L] = ord("+")
        out[:, 15 + L] = 10
        out[:, 16 + L:16 + 2 * L] = ord("I")
        out[:, 16 + 2 * L] = 10
        files.append(out.reshape(-1))
    return files


def bench_ingest(args):
    """FASTQ ingestion on the device: the 5 Mb / 100x pairs (1,666,667 x 2 x 150 bp) as two
    FASTQ files -> the three-block packed layout, through one reusable kc_ingest context.
    value = input bytes / device time (files already uploaded, sizes read back between
    kernels included); the host-to-host rate of the whole call is reported beside it."""
    from kc import synth
    from kc.ingest import FastqIngest
    t0 = time.time()
    store, meta = synth.workload_5mb_100x(seed=2)
    n_pairs, L = meta["n_pairs"], meta["read_len"]
    fmt = "FASTA" if args.fasta else "FASTQ"
    r1, r2 = (fasta_text if args.fasta else fastq_text)(store, n_pairs, L, seed=5)
    del store
    n_in = int(r1.size + r2.size)
    log(f"generated {fmt} pair, {n_in / 1e9:.3f} GB, in {time.time() - t0:.1f}s")
    ctx = FastqIngest()
    for _ in range(args.warmup):
        ctx.pack(r1, r2)
    walls, devs = [], []
    info = None
    for _ in range(args.steps):
        t = time.perf_counter()
        st, info = ctx.pack(r1, r2)
        walls.append((time.perf_counter() - t) * 1e3)
        devs.append(info["ms_device"])
    ms_dev, ms_wall = float(np.median(devs)), float(np.median(walls))
    out_bytes = info["data_len"] + 12 * info["n_reads"]
    alg = n_in + out_bytes
    log(f"device {ms_dev:.3f} ms, wall {ms_wall:.3f} ms, {info}")
    res = {
        "metric": f"{fmt} bytes/s ingested into the three-block packed layout (device-resident)",
        "value": round(n_in / (ms_dev / 1e3), 1), "unit": "bytes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_dev, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded)",
        "config": {"workload": f"5 Mb genome, 2x150 bp at 100x as two {fmt} files (1% internal N in R1, "
                               "1% leading NNN in R2" + ("; 60-base lines)" if args.fasta else ")"),
                   "input_bytes": n_in, "pairs": n_pairs,
                   "info": info, "host_to_host_ms": round(ms_wall, 3),
                   "host_to_host_bytes_per_s": round(n_in / (ms_wall / 1e3), 1)},
        "roofline": {"bound": "hbm", "kernel": "whole device pipeline (12 kernels + 4 scans)",
                     "achieved": round(alg / (ms_dev / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / (ms_dev / 1e3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": alg},
    }
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        from oracle import oracle as O
        nrec = min(200_000, n_pairs)
        rec = r1.size // n_pairs  # bytes per record (fixed-size rows)
        a, b = r1[:nrec * rec].tobytes(), r2[:nrec * rec].tobytes()
        t = time.perf_counter()
        ref = O.pack_fastq(a, b)
        cpu_s = time.perf_counter() - t
        dev, _ = ctx.pack(a, b)
        d, o, ln = dev.arrays()
        same = bool(np.array_equal(d, ref[0]) and np.array_equal(o, ref[1]) and np.array_equal(ln, ref[2]))
        res["cpu_baseline"] = {"value": round((len(a) + len(b)) / cpu_s, 1), "unit": "bytes/s", "cores": 1,
                               "kind": "port", "sample": f"first {nrec} records of both files "
                                                         f"({len(a) + len(b)} bytes), the oracle's "
                                                         f"single-thread FastxReader + gates",
                               "device_matches_on_sample": same}
    ctx.close()
    print(json.dumps(res), flush=True)


def bench_unitigs(args):
    """Compacted unitigs (CompactedGraph::build, graph.rs:65-234) of the 5 Mb / 100x solid
    set at k = the first of --ks, built on the device from the context's count table.
    value = solid k-mers compacted per second of device time (the count table already
    resident); the dominant stage is the adjacency pass (8 hash lookups per k-mer)."""
    from kc import AssemblerConfig, SpectrumEngine, synth
    k = int(args.ks.split(",")[0])
    t0 = time.time()
    store, meta = synth.workload_5mb_100x(seed=2)
    log(f"generated {len(store)} reads in {time.time() - t0:.1f}s")
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=args.bloom_bits, min_coverage=5.0))
    eng.upload(store)
    del store
    info = eng.run()
    n = int(info["n_solid"])
    log(f"k={k}: {n} solid k-mers (cutoff {info['cutoff']})")
    for _ in range(args.warmup):
        eng.unitigs()
    devs, walls, stages = [], [], []
    g = None
    for _ in range(args.steps):
        t = time.perf_counter()
        g = eng.unitigs()
        walls.append((time.perf_counter() - t) * 1e3)
        devs.append(g.ms_device)
        stages.append(g.ms_stage)
    ms_dev = float(np.median(devs))
    st_med = {name: round(float(np.median([x[name] for x in stages])), 3) for name in g.ms_stage}
    ok = bool(int(g.kmers.sum()) == n and g.seq.shape[0] == n + len(g) * (k - 1))
    log(f"device {ms_dev:.3f} ms ({st_med}), {len(g)} unitigs, {g.seq.shape[0]} bases, invariants {ok}")
    kw = 1 if k <= 32 else 2 if k <= 63 else 4
    # adjacency pass, per solid k-mer: its value, 8 hash-slot reads, the keys of the found
    # neighbours (about 2 on a genome path), degree + two neighbour ids out
    adj_bytes = n * (8 * kw + 8 * 8 + 2 * 8 * kw + 9)
    adj_ms = st_med["adjacency+links"]
    res = {
        "metric": "solid k-mers compacted into unitigs per second (device-resident count table)",
        "value": round(n / (ms_dev / 1e3), 1), "unit": "k-mers/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_dev, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic (seeded)",
        "config": {"workload": f"unitigs of the 5 Mb genome, 2x150 bp at 100x solid set, k={k}", "k": k,
                   "n_solid": n, "n_unitigs": len(g), "n_bases": int(g.seq.shape[0]), "stage_ms": st_med,
                   "wall_ms": round(float(np.median(walls)), 3), "invariants_hold": ok},
        "roofline": {"bound": "hbm", "kernel": "k_adjacent + k_links (adjacency+links stage)",
                     "achieved": round(adj_bytes / (adj_ms / 1e3) / 1e9, 1) if adj_ms > 0 else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(adj_bytes / (adj_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if adj_ms > 0 else None,
                     "traffic": None, "algorithmic_bytes_per_launch": adj_bytes,
                     "random_lookups_per_s": round(8 * n / (adj_ms / 1e3), 1) if adj_ms > 0 else None},
    }
    if not args.no_cpu and k <= 32:
        sys.path.insert(0, ROOT)
        from oracle import oracle as O
        keys, cnt, mask = eng.counts()
        solid = keys[mask.astype(bool), 0]
        gen = meta["genome"][: meta["genome"].shape[0] // 8].astype(np.uint64)
        m = gen.shape[0] - k + 1
        fw = np.zeros(m, np.uint64)
        rv = np.zeros(m, np.uint64)
        for i in range(k):
            fw = (fw << np.uint64(2)) | gen[i:i + m]
            rv = (rv << np.uint64(2)) | (np.uint64(3) - gen[k - 1 - i:k - 1 - i + m])
        samp = np.intersect1d(np.minimum(fw, rv), solid)
        sk = np.zeros((samp.size, 4), np.uint64)
        sk[:, 0] = samp
        t = time.perf_counter()
        ref = O.compacted_graph(k, sk, keys, cnt)
        cpu_s = time.perf_counter() - t
        res["cpu_baseline"] = {"value": round(samp.size / cpu_s, 1), "unit": "k-mers/s", "cores": 1, "kind": "port",
                               "sample": f"the {samp.size} solid k-mers of the genome's first 1/8 "
                                         f"({gen.shape[0]} bases), {len(ref)} unitigs, oracle CompactedGraph::build "
                                         f"(single thread, ascending order) in {cpu_s:.1f} s"}
    eng.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
