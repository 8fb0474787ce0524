"""GPU parity of Hammer (ErrorCorrector::correct_read_in_place, hammer.rs:34-120): the
device kernel (one read per wave, the read staged in LDS a chunk at a time) against the
hand-derived known answers of tests/hammer_kats.py, and against the oracle bit for bit on
reads of any length (past the LDS chunk of 2048 bases), through the standalone filter and
through the spectrum engine (error_correct), including multi-line FASTA records of several
kilobases ingested on the device."""
import sys

import numpy as np
import pytest

from hammer_kats import all_cases, preconditions
from kc import AssemblerConfig, ErrorCorrector, PackedReads, SpectrumEngine, TwoTierFilter
from kc.ingest import fastq_pack

from test_oracle_fastq import fasta

pytestmark = pytest.mark.gpu

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def canon_limbs(oracle, seqs, k):
    return np.array([oracle.kmer_canonical(oracle.kmer_from_bytes(s, k), k)[0] for s in seqs], dtype=np.uint64)


@pytest.mark.parametrize("case", all_cases(), ids=lambda c: c[0])
def test_known_answers_on_device(oracle, case):
    name, k, kmers, read, expected, states = case
    say(f"case hammer kat {name} k={k}")
    f = TwoTierFilter(1 << 24)
    keys = canon_limbs(oracle, [x.encode() for x in kmers], k)
    f.insert_kmer256_many(np.concatenate([keys, keys]))

    def is_solid(w):
        return bool(f.is_solid_kmer256_many(canon_limbs(oracle, [w.encode()], k))[0])

    preconditions(k, kmers, read, expected, states, is_solid)
    mod, out = ErrorCorrector(k).correct_read_in_place(read.encode(), f)
    f.close()
    assert (mod, out.decode() if isinstance(out, bytes) else out) == expected


def random_genome(rng, n):
    return ACGT[rng.integers(0, 4, n)].tobytes()


def noisy_reads(rng, genome, lengths, err):
    reads = []
    for L in lengths:
        L = min(int(L), len(genome))
        s = int(rng.integers(0, len(genome) - L + 1))
        r = bytearray(genome[s:s + L])
        for p in np.flatnonzero(rng.random(L) < err):
            r[p] = ACGT[(int(np.flatnonzero(ACGT == r[p])[0]) + int(rng.integers(1, 4))) & 3]
        if rng.random() < 0.5:  # reverse strand
            r = bytearray(r[::-1].translate(bytes.maketrans(b"ACGT", b"TGCA")))
        reads.append(bytes(r))
    return reads


@pytest.mark.parametrize("k,bits", [(11, 1 << 22), (21, 1 << 20), (33, (1 << 20) + 192), (55, 1 << 16),
                                    (77, 1 << 21), (127, 1 << 21)])
def test_long_reads_match_oracle(oracle, k, bits):
    """Reads of k to 9000 bases (several LDS chunks), 0.4% errors, both strands, against
    the oracle's correct_read on the same filter state (built by the same inserts)."""
    rng = np.random.default_rng(k)
    genome = random_genome(rng, 60000)
    store = PackedReads()
    store.pack_clean(genome)
    lens = np.concatenate([[k, k + 1, 2047, 2048, 2049, 4100, 9000], rng.integers(k, 400, 200),
                           rng.integers(2000, 6000, 12)])
    reads = noisy_reads(rng, genome, lens, 0.004)
    say(f"case hammer long reads k={k} bits={bits} reads={len(reads)} max={max(len(r) for r in reads)}")
    # the solid set: every window of the genome inserted twice, on the device and in the oracle
    keys = np.concatenate([canon_limbs(oracle, [genome[i:i + k] for i in range(len(genome) - k + 1)], k)] * 2)
    f = TwoTierFilter(bits)
    f.insert_kmer256_many(keys)
    of = oracle.TwoTierFilter(bits)
    of.insert_many(keys)
    assert all(np.array_equal(a, b) for a, b in zip(f.export(), of.export()))
    out, n = ErrorCorrector(k).correct_reads(reads, f)
    f.close()
    want = [oracle.correct_read(of, k, r) for r in reads]
    assert n == sum(m for m, _ in want)
    bad = [i for i, (o, (_, w)) in enumerate(zip(out, want)) if (o if isinstance(o, bytes) else o.encode()) != w]
    assert not bad, (bad[:5], [len(reads[i]) for i in bad[:5]])


@pytest.mark.parametrize("k", [21, 33, 55])
def test_engine_long_reads_with_hammer_match_oracle(oracle, k):
    """Stage 2 + 3 with error_correct over reads of 2-5 kb mixed with 150 bp reads: Bloom
    tiers, corrected reads, counts, cutoff and solid set equal the oracle's."""
    rng = np.random.default_rng(100 + k)
    genome = random_genome(rng, 40000)
    lens = np.concatenate([np.full(1500, 150), rng.integers(2000, 5000, 40)])
    rng.shuffle(lens)
    reads = noisy_reads(rng, genome, lens, 0.005)
    store = PackedReads()
    for r in reads:
        store.pack_clean(r)
    d, o, l = store.arrays()
    bits = 1 << 22
    say(f"case engine hammer long reads k={k} reads={len(store)}")
    S = oracle.Spectrum(d, o, l, k, bits, True, 5.0, 1)
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=True))
    eng.upload(store)
    info = eng.run()
    seen, solid = eng.bloom()
    keys, cnt, mask = eng.counts()
    corr = eng.reads()
    eng.close()
    oseen, osolid = S.bloom()
    okeys, ocnt, omask = S.counts()
    assert info["corrected_reads"] == S.corrected_reads > 0
    assert np.array_equal(corr, S.corrected_data())
    assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid)
    assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt) and np.array_equal(mask, omask)
    assert info["cutoff"] == S.cutoff


@pytest.mark.parametrize("width", [60, 0])
def test_fasta_records_of_kilobases_with_hammer(oracle, width):
    """Multi-line (60-base lines) and unwrapped FASTA records of 2-5 kb ingested on the
    device, then counted with Hammer on: equal to the oracle's ingestion + Spectrum."""
    rng = np.random.default_rng(7 + width)
    genome = random_genome(rng, 30000)
    reads = noisy_reads(rng, genome, rng.integers(2000, 5000, 30), 0.004) + noisy_reads(
        rng, genome, np.full(800, 150), 0.004)
    text = fasta(reads, width)
    dev, _ = fastq_pack(text)
    d, o, l = dev.arrays()
    ref = oracle.pack_fastq(text)
    assert np.array_equal(d, ref[0]) and np.array_equal(o, ref[1]) and np.array_equal(l, ref[2])
    k, bits = 31, 1 << 22
    say(f"case fasta hammer width={width} reads={len(dev)}")
    S = oracle.Spectrum(d, o, l, k, bits, True, 5.0, 1)
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=True))
    eng.upload(dev)
    info = eng.run()
    corr = eng.reads()
    keys, cnt, mask = eng.counts()
    eng.close()
    okeys, ocnt, omask = S.counts()
    assert info["corrected_reads"] == S.corrected_reads > 0
    assert np.array_equal(corr, S.corrected_data())
    assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt) and np.array_equal(mask, omask)


def test_hamming_distance_2bit_on_device(oracle):
    """hamming_distance_2bit (hammer.rs:13-18): the reference's known answers
    (hammer.rs:150-158), and random words against the oracle and a per-base count."""
    from kc import hamming_distance_2bit
    assert list(hamming_distance_2bit([0b00011011, 0b00011011, 0b00000000],
                                      [0b00011011, 0b00011000, 0b01010101])) == [0, 1, 4]
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2**63, 100000, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, 100000, dtype=np.uint64)
    b = a.copy()
    flip = rng.random((100000, 32)) < 0.1
    for j in range(32):
        b ^= np.where(flip[:, j], rng.integers(1, 4, 100000, dtype=np.uint64) << np.uint64(2 * j), np.uint64(0))
    got = hamming_distance_2bit(a, b)
    assert np.array_equal(got, flip.sum(axis=1).astype(np.uint32))
    assert all(int(got[i]) == oracle.hamming_2bit(int(a[i]), int(b[i])) for i in range(0, 100000, 997))
