"""GPU parity: the HIP engine (libkc.so via its C ABI) against the oracle on the
same seeded inputs.  Integer work, so every comparison is bit-exact: both Bloom
tiers, the sorted (k-mer, count) list, the cutoff, the solid set and, with
Hammer on, the corrected reads."""
import sys

import numpy as np
import pytest

from kc import AssemblerConfig, ErrorCorrector, PackedReads, SpectrumEngine, TwoTierFilter, synth
from kc.packed_reads import read_fastq_sequences

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_spectrum(store, k, bits, ec=False, min_cov=5.0, load=0.0):
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec, min_coverage=min_cov),
                         table_load=load)
    try:
        eng.upload(store)
        info = eng.run()
        seen, solid = eng.bloom()
        keys, cnt, mask = eng.counts()
        reads = eng.reads()
        return info, seen, solid, keys, cnt, mask, reads
    finally:
        eng.close()


def assert_parity(oracle, store, k, bits, ec=False, min_cov=5.0, load=0.0):
    d, o, l = store.arrays()
    say(f"case k={k} bits={bits} ec={ec} reads={len(store)}")
    S = oracle.Spectrum(d, o, l, k, bits, ec, min_cov, 1)
    info, seen, solid, keys, cnt, mask, reads = gpu_spectrum(store, k, bits, ec, min_cov, load)
    oseen, osolid = S.bloom()
    okeys, ocnt, omask = S.counts()
    assert info["n_occ"] == S.n_occ
    assert np.array_equal(seen, oseen), "seen tier differs"
    assert np.array_equal(solid, osolid), "solid tier differs"
    assert keys.shape == okeys.shape, (keys.shape, okeys.shape)
    assert np.array_equal(keys, okeys), "sorted keys differ"
    assert np.array_equal(cnt, ocnt), "counts differ"
    assert info["cutoff"] == S.cutoff
    assert info["top_cov"] == min(S.top_cov, 1023)
    assert np.array_equal(mask, omask), "solid set differs"
    assert info["n_solid"] == S.n_solid
    if ec:
        assert info["corrected_reads"] == S.corrected_reads
        assert np.array_equal(reads, S.corrected_data()), "corrected reads differ"
    return info


# ---- standalone filter (bloom.rs) -----------------------------------------------

def test_filter_reference_tests():
    # bloom.rs:138-156 and tests/test_hammer_and_bloom.rs:39-70
    f = TwoTierFilter(1024 * 1024)
    a, b, c = 0x123456789ABCDEF0, 0xFEDCBA9876543210, 0xA5A5A5A5A5A5A5A5
    assert not f.is_solid(a)
    assert f.insert(a) is False and not f.is_solid(a)
    assert f.insert(a) is True and f.is_solid(a)
    assert f.insert(a) is True
    assert not f.is_solid(b) and not f.is_solid(c)
    assert f.memory_usage_bytes() == 2 * (1024 * 1024 // 8)


@pytest.mark.parametrize("bits", [64, 1000, 4096, 1 << 16, (1 << 16) + 192, 1 << 24])
def test_filter_sequential_batches_match_oracle(oracle, bits):
    rng = np.random.default_rng(bits)
    pool = np.zeros((400, 4), dtype=np.uint64)
    pool[:, 0] = rng.integers(0, 2**63, 400, dtype=np.uint64)
    pool[:80, 1] = rng.integers(0, 2**50, 80, dtype=np.uint64)
    pool[:20, 2] = rng.integers(0, 2**40, 20, dtype=np.uint64)
    occ = pool[rng.integers(0, 400, 5000)]
    f = TwoTierFilter(bits)
    g = oracle.TwoTierFilter(bits)
    for lo, hi in ((0, 1), (1, 700), (700, 701), (701, 5000)):
        prom = f.insert_kmer256_many(occ[lo:hi])
        oprom = g.insert_many(occ[lo:hi])
        assert np.array_equal(prom, oprom.astype(bool)), (lo, hi)
    s1, s2 = f.export()
    o1, o2 = g.export()
    assert np.array_equal(s1, o1) and np.array_equal(s2, o2)
    assert np.array_equal(f.is_solid_kmer256_many(pool), g.is_solid_many(pool).astype(bool))


# ---- Hammer (hammer.rs) -------------------------------------------------------------

def _insert_seq(f, seq, k, times, oracle):
    keys = [oracle.kmer_canonical(oracle.kmer_from_bytes(seq[i:i + k], k), k)[0] for i in range(len(seq) - k + 1)]
    for _ in range(times):
        f.insert_kmer256_many(np.array(keys, dtype=np.uint64))


def test_hammer_reference_tests(oracle):
    # hammer.rs:160-192, tests/test_hammer_and_bloom.rs:72-102
    f = TwoTierFilter(1024 * 1024)
    _insert_seq(f, b"ACGTACGTACGT", 5, 2, oracle)
    assert ErrorCorrector(5).correct_read_in_place(b"ACGGACGTACGT", f) == (True, b"ACGTACGTACGT")
    ref = b"GATTACATAGGATACAGGATACCAGATTAGACATAGACAG"
    f2 = TwoTierFilter(1024 * 1024)
    _insert_seq(f2, ref, 15, 2, oracle)
    bad = bytearray(ref)
    bad[20] = ord("T")
    out, n = ErrorCorrector(15).correct_reads([bytes(bad)], f2)
    assert n == 1 and out[0] == ref
    assert ErrorCorrector(15).correct_read_in_place(b"ACGT", TwoTierFilter(1024)) == (False, b"ACGT")


@pytest.mark.parametrize("k,bits", [(11, 1 << 14), (21, 1 << 22), (33, 1 << 22), (55, 1 << 12), (77, 1 << 20),
                                    (101, 1 << 16)])
def test_hammer_matches_oracle_on_noisy_reads(oracle, k, bits):
    store, _ = synth.uniform_genome_pairs(seed=100 + k, genome_size=4000, n_pairs=300, read_len=120,
                                          insert_mean=250, insert_sd=20, p_start=0.02)
    d, o, l = store.arrays()
    # filter state from the first half of the reads, corrected: the second half
    g = oracle.TwoTierFilter(bits)
    f = TwoTierFilter(bits)
    half = len(store) // 2
    keys = []
    for i in range(half):
        r = store.get_read(i)
        keys += [oracle.kmer_canonical(oracle.kmer_from_bytes(r[j:j + k], k), k)[0] for j in range(len(r) - k + 1)]
    keys = np.array(keys, dtype=np.uint64)
    g.insert_many(keys)
    f.insert_kmer256_many(keys)
    tail = PackedReads()
    for i in range(half, len(store)):
        tail.pack_clean(store.get_read(i))
    data, n = ErrorCorrector(k).correct_packed(tail, f)
    exp, en = [], 0
    for i in range(len(tail)):
        m, r = oracle.correct_read(g, k, tail.get_read(i))
        en += m
        exp.append(r)
    fixed = PackedReads.from_arrays(data, tail.offsets, tail.lengths)
    assert n == en
    assert [fixed.get_read(i) for i in range(len(fixed))] == exp


# ---- spectrum (assemble.rs Stage 2-3) ---------------------------------------------

@pytest.mark.parametrize("k,bits,ec", [
    (21, 1 << 20, False), (21, 1 << 16, False), (21, 1 << 20, True), (31, (1 << 20) + 64 * 5, False),
    (32, 1 << 18, False), (33, 1 << 20, False), (33, 1 << 20, True), (55, 1 << 20, False), (55, 4096, True),
    (63, 1 << 19, False), (11, 1 << 12, False), (5, 64 * 3, False), (1, 64, False),
    (64, 1 << 18, False), (65, 1 << 20, False), (77, 1 << 20, True), (99, 1 << 16, False), (127, 1 << 20, True),
    (127, 4096, False), (96, (1 << 18) + 64 * 3, True),
])
def test_spectrum_parity_synthetic(oracle, k, bits, ec):
    store, _ = synth.uniform_genome_pairs(seed=7 * k + bits % 97 + ec, genome_size=20000, n_pairs=2000,
                                          read_len=150, p_start=0.005)
    assert_parity(oracle, store, k, bits, ec)


def test_spectrum_parity_high_load_and_min_cov(oracle):
    store, _ = synth.uniform_genome_pairs(seed=11, genome_size=30000, n_pairs=3000, read_len=150, p_start=0.01)
    assert_parity(oracle, store, 21, 1 << 20, load=0.9)
    assert_parity(oracle, store, 21, 1 << 20, min_cov=12.0)


def test_spectrum_parity_valley_cutoff(oracle):
    # 60x coverage so the count spectrum has a valley above 2
    store, _ = synth.uniform_genome_pairs(seed=12, genome_size=20000, n_pairs=4000, read_len=150, p_start=0.01)
    info = assert_parity(oracle, store, 21, 1 << 22)
    say("cutoff", info["cutoff"], "top", info["top_cov"])


def test_spectrum_ragged_and_edge_reads(oracle):
    from test_packed_reads import lcg_workload, modulo4_reads
    st = PackedReads()
    for r in modulo4_reads() + lcg_workload()[:500] + [b"A" * 21, b"T" * 40, b"ACGT" * 10]:
        st.add_read(r)
    for k in (21, 33, 63, 64, 65, 127):
        assert_parity(oracle, st, k, 1 << 16, ec=(k in (21, 65)))


@pytest.mark.parametrize("k,run", [(33, 4), (55, 12), (55, 300), (63, 80), (55, 1500)])
def test_count_table_order_with_shared_prefixes(oracle, k, run):
    """Count-table entries that share their top 64 bits: k-mers A^32 + a varied tail (forward
    strand canonical), each read twice so that it is Bloom-solid.  After the sort by the top
    64 bits, a run of equal tops up to 16 entries is ordered by one thread (4, 12), up to 1024
    by a workgroup in LDS (300 at k = 55, 80 at k = 63), and a longer one (1500) makes the
    two-word sort redo the table.  Keys, counts and mask match the oracle."""
    rng = np.random.default_rng(k * 1000 + run)
    st = PackedReads()
    tail = k - 32
    bases = np.frombuffer(b"ACGT", dtype=np.uint8)
    seen = set()
    while len(seen) < run:
        # the forward strand is canonical: its 32 leading A's beat the reverse complement,
        # which has at most `tail` < 32 of them before the complement of the A run
        t = bytes(bases[rng.integers(0, 4, tail)])
        if t in seen:
            continue
        seen.add(t)
        for _ in range(2):
            st.add_read(b"A" * 32 + t)
    for r in range(300):  # and unrelated reads around them
        st.add_read(bytes(bases[rng.integers(0, 4, 150)]))
    assert_parity(oracle, st, k, 1 << 20)


def test_spectrum_empty_and_all_short(oracle):
    empty = PackedReads()
    info, seen, solid, keys, cnt, mask, _ = gpu_spectrum(empty, 21, 4096)
    assert info["n_occ"] == 0 and keys.shape[0] == 0 and not seen.any() and info["cutoff"] == 2
    short = PackedReads()
    for r in (b"ACGT", b"", b"A" * 20):
        short.add_read(r)
    assert_parity(oracle, short, 21, 4096, ec=True)


def test_spectrum_unaligned_offsets(oracle):
    # reads placed at arbitrary byte offsets with gaps between them
    base, _ = synth.uniform_genome_pairs(seed=3, genome_size=5000, n_pairs=200, read_len=101)
    d, o, l = base.arrays()
    gap = np.arange(len(base), dtype=np.uint64) * np.uint64(3)
    nb = (l.astype(np.uint64) + np.uint64(3)) // np.uint64(4)
    new_off = o + gap + np.uint64(1)
    data = np.full(int(new_off[-1] + nb[-1]) + 5, 0xA5, dtype=np.uint8)
    for i in range(len(base)):
        data[int(new_off[i]):int(new_off[i] + nb[i])] = d[int(o[i]):int(o[i] + nb[i])]
    st = PackedReads.from_arrays(data, new_off, l)
    assert_parity(oracle, st, 21, 1 << 18)


def test_ecoli_1k_fixture(oracle):
    # tests/test_pipeline.rs:33-91 configuration: k=31, 2^24 bits, Hammer on.  The reads
    # carry no sequencing errors (the oracle corrects none either); every 31-mer of the
    # 1 kb reference must be in the solid set.
    import gzip
    import os
    from conftest import GOLDEN
    from oracle import hash_np as H
    r1 = read_fastq_sequences(os.path.join(GOLDEN, "ecoli_1K_1.fq.gz"))
    r2 = read_fastq_sequences(os.path.join(GOLDEN, "ecoli_1K_2.fq.gz"))
    store = PackedReads.from_paired_slices(r1, r2)
    assert assert_parity(oracle, store, 31, 16 * 1024 * 1024, ec=True)["corrected_reads"] == 0
    info, _, _, keys, cnt, mask, _ = gpu_spectrum(store, 31, 16 * 1024 * 1024)
    with gzip.open(os.path.join(GOLDEN, "reference_1K.fa.gz"), "rb") as fh:
        ref = b"".join(ln.strip() for ln in fh if not ln.startswith(b">"))
    solid = {tuple(int(x) for x in w) for w in keys[mask.astype(bool)]}
    genome = {H.canonical_words_of_string(ref[i:i + 31]) for i in range(len(ref) - 30)}
    assert len(ref) == 1000 and genome <= solid


def test_spectrum_1mb_30x_full_size(oracle):
    """The 1 Mb / 30x / k=21 workload at full size with the default 2^29-bit tiers."""
    store, _ = synth.workload_1mb_30x(seed=1)
    assert_parity(oracle, store, 21, 512 * 1024 * 1024)


def test_repeated_runs_are_identical():
    store, _ = synth.uniform_genome_pairs(seed=9, genome_size=50000, n_pairs=5000, p_start=0.005)
    eng = SpectrumEngine(AssemblerConfig(k=33, bloom_bits=1 << 20))
    eng.upload(store)
    outs = []
    for _ in range(3):
        eng.run()
        outs.append((eng.bloom(), eng.counts()))
    for o in outs[1:]:
        for a, b in zip(o[0] + o[1], outs[0][0] + outs[0][1]):
            assert np.array_equal(a, b)
    eng.close()


# ---- several libraries in one context ----------------------------------------------

def _uneven_library(seed, genome, pairs, min_len=50, max_len=250):
    rng = np.random.default_rng(seed)
    st, _ = synth.uniform_genome_pairs(seed=seed, genome_size=genome, n_pairs=pairs, read_len=max_len,
                                       insert_mean=max_len * 2, insert_sd=20, p_start=0.01)
    out = PackedReads()
    for i in range(len(st)):
        r = st.get_read(i)
        out.pack_clean(r[: int(rng.integers(min_len, max_len + 1))])
    return out


@pytest.mark.parametrize("k,ec", [(21, False), (33, True), (55, False), (31, True), (63, False), (77, True),
                                  (127, False)])
def test_many_libraries_match_each_alone(oracle, k, ec):
    libs = [_uneven_library(1000 + i, genome=3000 + 1500 * i, pairs=150 + 40 * i) for i in range(5)]
    libs.insert(2, PackedReads())  # an empty library
    short = PackedReads()
    short.pack_clean(b"ACGT" * 4)
    libs.append(short)  # reads shorter than k only
    bits = [1 << 16, (1 << 18) + 64, 4096, 1 << 20, 1 << 17, 1 << 15, 1 << 12]
    n_libs = len(libs)  # at k = 63, (library, V) takes more than two words: the record layout
    eng = SpectrumEngine(AssemblerConfig(k=k, error_correct=ec))
    for i in range(n_libs):
        assert eng.add_library(libs[i], bloom_bits=bits[i], min_coverage=5.0 if i != 3 else 12.0) == i
    eng.run()
    for i in range(n_libs):
        say(f"library {i} k={k} ec={ec}")
        d, o, l = libs[i].arrays()
        S = oracle.Spectrum(d, o, l, k, bits[i], ec, 5.0 if i != 3 else 12.0, 1)
        info = eng.info(i)
        seen, solid = eng.bloom(i)
        keys, cnt, mask = eng.counts(i)
        oseen, osolid = S.bloom()
        okeys, ocnt, omask = S.counts()
        assert info["n_occ"] == S.n_occ
        assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid)
        assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt)
        assert np.array_equal(mask, omask) and info["cutoff"] == S.cutoff and info["n_solid"] == S.n_solid
        if ec:
            assert info["corrected_reads"] == S.corrected_reads
            assert np.array_equal(eng.reads(i), S.corrected_data())
    eng.close()
