#!/usr/bin/env python3
"""Writes tests/golden/full_size_digests.json: SHA-256 digests of the sequential oracle's
outputs (both Bloom tiers, the sorted count table, counts, solid mask, cutoff; corrected
reads with Hammer on) on the full-size synthetic workloads:

  5mb_100x   the 5 Mb / 100x pairs (kc.synth.workload_5mb_100x, seed 2), k = 21, 33, 55,
             2^29-bit tiers, Hammer off (assemble.rs:129-298 at -t 1)
  1mb_30x    the 1 Mb / 30x pairs (kc.synth.workload_1mb_30x, seed 1), k = 21, 33, 55,
             2^29-bit tiers, Hammer off; k = 21 also with Hammer on
  64lib_1.0  the 64 uneven libraries at full scale (kc.synth.workload_64_libraries, seed 3;
             5 kb-6 Mb genomes, 20-200x, 50-250 bp reads), k = 33, 2^29-bit tiers: every
             8th library (0, 8, ..., 56), each counted alone by the oracle
  meta_0.05  the metagenome sample at 1/20 scale (kc.synth.workload_metagenome, seed 4,
             scale 0.05: 20 genomes totalling 4.1 Mb, 1.5M pairs), k = 55, 2^29-bit tiers

The oracle runs single-threaded (the reference's sequential order); the 5 Mb cases take a
few minutes each.  Usage: python tests/golden/make_full_digests.py [--out PATH] [case ...]
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (ROOT, os.path.join(ROOT, "spades-rs_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

OUT = os.path.join(ROOT, "tests", "golden", "full_size_digests.json")
BITS = 1 << 29

CASES = {
    "5mb_100x_k21": ("5mb_100x", 21, False),
    "5mb_100x_k33": ("5mb_100x", 33, False),
    "5mb_100x_k55": ("5mb_100x", 55, False),
    "1mb_30x_k21": ("1mb_30x", 21, False),
    "1mb_30x_k21_hammer": ("1mb_30x", 21, True),
    "1mb_30x_k33": ("1mb_30x", 33, False),
    "1mb_30x_k55": ("1mb_30x", 55, False),
    "meta_0.05_k55": ("meta_0.05", 55, False),
}


def workload(name):
    from kc import synth
    if name == "5mb_100x":
        return synth.workload_5mb_100x(seed=2)[0]
    if name.startswith("meta_"):
        return synth.workload_metagenome(seed=4, scale=float(name[5:]))[0]
    return synth.workload_1mb_30x(seed=1)[0]


def many_libraries(out, path):
    """The sampled libraries of the 64-library workload, oracle runs in parallel threads
    (the oracle's ctypes calls release the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor

    from kc import synth
    from kc.digest import spectrum_digests
    from oracle import oracle as O
    libs = synth.workload_64_libraries(seed=3, scale=1.0)

    def one(i):
        d, o, l = libs[i][0].arrays()
        t0 = time.time()
        S = O.Spectrum(d, o, l, 33, BITS, False, 5.0, 1)
        seen, solid = S.bloom()
        keys, cnt, mask = S.counts()
        dig = spectrum_digests(seen, solid, keys, cnt, mask, S.cutoff, S.n_occ)
        dig.update({"workload": "64lib_1.0", "library": i, "k": 33, "bloom_bits": BITS, "error_correct": False})
        print(f"64lib_1.0 library {i}: n_occ={S.n_occ} n_counts={S.n_counts} ({time.time() - t0:.0f} s)", flush=True)
        return i, dig

    with ThreadPoolExecutor(max_workers=8) as ex:
        for i, dig in ex.map(one, range(0, 64, 8)):
            out[f"64lib_1.0_lib{i}_k33"] = dig
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main():
    from kc.digest import spectrum_digests
    from oracle import oracle as O
    args = sys.argv[1:]
    path = OUT
    if args[:1] == ["--out"]:
        path, args = args[1], args[2:]
    want = args or list(CASES) + ["64lib_1.0"]
    out = json.load(open(path)) if os.path.exists(path) else {}
    if "64lib_1.0" in want:
        want = [w for w in want if w != "64lib_1.0"]
        many_libraries(out, path)
    stores = {}
    for case in want:
        wl, k, ec = CASES[case]
        if wl not in stores:
            stores[wl] = workload(wl)
        d, o, l = stores[wl].arrays()
        t0 = time.time()
        S = O.Spectrum(d, o, l, k, BITS, ec, 5.0, 1)
        seen, solid = S.bloom()
        keys, cnt, mask = S.counts()
        corr = S.corrected_data() if ec else None
        dig = spectrum_digests(seen, solid, keys, cnt, mask, S.cutoff, S.n_occ, corr,
                               S.corrected_reads if ec else None)
        dig.update({"workload": wl, "k": k, "bloom_bits": BITS, "error_correct": ec, "top_cov": int(S.top_cov)})
        out[case] = dig
        print(f"{case}: n_occ={S.n_occ} n_counts={S.n_counts} cutoff={S.cutoff} ({time.time() - t0:.0f} s)",
              flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
