"""Pins the oracle's hammer.rs restatement to the reference's tests
(hammer.rs:160-192, tests/test_hammer_and_bloom.rs:72-102) and to a planted-error
property: a single substitution inside a read drawn from a clean, solid genome
is restored."""
import random

import pytest


def insert_seq(oracle, f, seq, k, times):
    for _ in range(times):
        for i in range(len(seq) - k + 1):
            w = oracle.kmer_from_bytes(seq[i:i + k], k)
            f.insert_kmer256(oracle.kmer_canonical(w, k)[0])


def test_correct_read_single_substitution(oracle):
    # hammer.rs:160-183
    f = oracle.TwoTierFilter(1024 * 1024)
    insert_seq(oracle, f, b"ACGTACGTACGT", 5, 2)
    mod, out = oracle.correct_read(f, 5, b"ACGGACGTACGT")
    assert mod and out == b"ACGTACGTACGT"


def test_error_corrector_single_base_correction(oracle):
    # tests/test_hammer_and_bloom.rs:72-102
    ref = b"GATTACATAGGATACAGGATACCAGATTAGACATAGACAG"
    f = oracle.TwoTierFilter(1024 * 1024)
    insert_seq(oracle, f, ref, 15, 2)
    bad = bytearray(ref)
    assert bad[20] == ord("A")
    bad[20] = ord("T")
    mod, out = oracle.correct_read(f, 15, bytes(bad))
    assert mod and out == ref


def test_short_read_bypass(oracle):
    # hammer.rs:185-192
    f = oracle.TwoTierFilter(1024)
    assert oracle.correct_read(f, 15, b"ACGT") == (False, b"ACGT")


@pytest.mark.parametrize("k", [11, 21, 31])
def test_planted_substitution_is_restored(oracle, k):
    rng = random.Random(k)
    genome = bytes(rng.choice(b"ACGT") for _ in range(3000))
    f = oracle.TwoTierFilter(1 << 22)
    insert_seq(oracle, f, genome, k, 2)
    restored = 0
    for _ in range(40):
        s = rng.randrange(0, len(genome) - 100)
        read = bytearray(genome[s:s + 100])
        pos = rng.randrange(k, 100 - k)
        read[pos] = rng.choice([c for c in b"ACGT" if c != read[pos]])
        mod, out = oracle.correct_read(f, k, bytes(read))
        assert mod
        restored += out == genome[s:s + 100]
    assert restored == 40


# ---- hand-derived known answers (tests/hammer_kats.py) --------------------------

from hammer_kats import all_cases, preconditions  # noqa: E402


@pytest.mark.parametrize("case", all_cases(), ids=lambda c: c[0])
def test_known_answers(oracle, case):
    name, k, kmers, read, expected, states = case
    f = oracle.TwoTierFilter(1 << 24)
    for _ in range(2):
        for x in kmers:
            f.insert_kmer256(oracle.kmer_canonical(oracle.kmer_from_bytes(x.encode(), k), k)[0])

    def is_solid(w):
        return f.is_solid_kmer256(oracle.kmer_canonical(oracle.kmer_from_bytes(w.encode(), k), k)[0])

    preconditions(k, kmers, read, expected, states, is_solid)
    mod, out = oracle.correct_read(f, k, read.encode())
    assert (mod, out.decode()) == expected
