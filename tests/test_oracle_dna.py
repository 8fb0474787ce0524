"""Pins the oracle's dna.rs restatement to the reference's own known answers
(dna.rs:329-432, tests/test_dna_kmer.rs) and to string-level forms computed
independently in Python (oracle/hash_np.py)."""
import random

import pytest

from oracle import hash_np as H

TEMPLATE = (b"ATGCGATCGATCGATAGCTAGCTAGCTAGCTAAGCTAGCTAGCTAGCTAATGCGATCGATCGATAGCTAGCTAGCTAGCTAAGCTAGCTA"
            b"GCTAGCTAATGCGATCGATCGATAGCTAGCTAGCTAGCTAAGCTAGCTAGCTAGCTAATGCGATCGATCGATAGCTAGCTAGCTAGCTAA"
            b"GCTAGCTAGCTAGCTA")


def enc(s, k):
    """string_to_kmer (dna.rs:104-114) written out for the test."""
    v = 0
    for ch in s[:k]:
        v = (v << 2) | b"ACGT".index(bytes([ch]).upper())
    return v


def test_base_encoding_table(oracle):
    # tests/test_dna_kmer.rs:9-24
    for ch, v in zip(b"ACGTacgt", (0, 1, 2, 3, 0, 1, 2, 3)):
        assert oracle.lib().or_ascii_to_2bit(ch) == v
    for v, ch in enumerate(b"ACGT"):
        assert oracle.lib().or_bit2_to_base(v) == ch
    for bad in b"N1 -n*":
        assert oracle.lib().or_ascii_to_2bit(bad) == 0xFF


def test_invalid_windows_rejected(oracle):
    # tests/test_dna_kmer.rs:26-31
    for s in (b"ACGTN", b"ACGT1", b"AC GT", b"AC-GT"):
        assert oracle.kmer_from_bytes(s, 5) is None
    assert oracle.kmer_from_bytes(b"ACGT", 5) is None
    assert oracle.kmer_from_bytes(b"", 0) is None


def test_revcomp_u64_known_answers(oracle):
    # dna.rs:329-342: ACGT is its own reverse complement; AAAA <-> TTTT
    acgt = enc(b"ACGT", 4)
    assert acgt == 0x1B
    assert oracle.revcomp_u64(acgt, 4) == acgt
    a4, t4 = enc(b"AAAA", 4), enc(b"TTTT", 4)
    assert oracle.revcomp_u64(a4, 4) == t4
    assert oracle.revcomp_u64(t4, 4) == a4


def test_revcomp_u64_involution_and_canonical(oracle):
    # tests/test_dna_kmer.rs:47-106
    for k in range(1, 33):
        seq = bytes(b"ACGT"[i % 4] for i in range(k))
        km = enc(seq, k)
        assert oracle.revcomp_u64(oracle.revcomp_u64(km, k), k) == km
    for k in (1, 7, 15, 21, 31, 32):
        seq = bytes(b"ACGT"[(i * 7) % 4] for i in range(k))
        km = enc(seq, k)
        c1, rc1 = oracle.canonical_u64(km, k)
        assert oracle.canonical_u64(c1, k)[0] == c1
        rc = oracle.revcomp_u64(km, k)
        c2, rc2 = oracle.canonical_u64(rc, k)
        assert c1 == c2
        if km != rc:
            assert rc1 != rc2


@pytest.mark.parametrize("k", [21, 33, 55, 64, 77, 99, 127])
def test_kmer256_roundtrip_and_revcomp(oracle, k):
    # dna.rs:354-367
    seq = TEMPLATE[:k]
    km = oracle.kmer_from_bytes(seq, k)
    assert oracle.kmer_to_string(km, k) == seq
    assert oracle.kmer_revcomp(oracle.kmer_revcomp(km, k), k) == km


def test_kmer256_revcomp_known_answers(oracle):
    # dna.rs:370-414
    base = b"ACGT" * 16 + b"A"
    km = oracle.kmer_from_bytes(base, 65)
    assert oracle.kmer_to_string(oracle.kmer_revcomp(km, 65), 65) == b"T" + b"ACGT" * 16
    b96 = b"ACGT" * 24
    k96 = oracle.kmer_from_bytes(b96, 96)
    assert oracle.kmer_to_string(oracle.kmer_revcomp(k96, 96), 96) == b96
    s127 = b"ATCG" * 31 + b"ATG"
    k127 = oracle.kmer_from_bytes(s127, 127)
    assert oracle.kmer_to_string(oracle.kmer_revcomp(k127, 127), 127) == H.revcomp_string(s127)


@pytest.mark.parametrize("k", [21, 33, 55, 64, 77, 99, 127])
def test_extend_and_prepend(oracle, k):
    # dna.rs:416-432
    seq = TEMPLATE[:k]
    km = oracle.kmer_from_bytes(seq, k)
    assert oracle.kmer_to_string(oracle.kmer_extend_right(km, 1, k), k) == seq[1:] + b"C"
    assert oracle.kmer_to_string(oracle.kmer_prepend_left(km, 2, k), k) == b"G" + seq[:-1]


@pytest.mark.parametrize("k", [33, 55, 64, 65, 77, 99, 111, 127, 128])
def test_kmer256_large_k_invariance(oracle, k):
    # tests/test_dna_kmer.rs:108-146
    seq = bytes(b"ACGT"[(i * 13 + 5) % 4] for i in range(k))
    km = oracle.kmer_from_bytes(seq, k)
    rc = oracle.kmer_revcomp(km, k)
    assert oracle.kmer_revcomp(rc, k) == km
    assert oracle.kmer_canonical(km, k)[0] == oracle.kmer_canonical(rc, k)[0]
    assert oracle.kmer_to_string(km, k) == seq


def test_extend_prepend_k63(oracle):
    # tests/test_dna_kmer.rs:148-175
    seq = bytes(b"ACGT"[i % 4] for i in range(63))
    km = oracle.kmer_from_bytes(seq, 63)
    assert oracle.kmer_to_string(oracle.kmer_extend_right(km, 3, 63), 63) == seq[1:] + b"T"
    assert oracle.kmer_to_string(oracle.kmer_prepend_left(km, 2, 63), 63) == b"G" + seq[:-1]


@pytest.mark.parametrize("k", [1, 5, 11, 21, 31, 32, 33, 55, 63, 64, 65, 100, 127])
def test_canonical_matches_string_level(oracle, k):
    # canonical over Kmer256 (lo, hi) order == min of the string-level encodings
    rng = random.Random(k)
    for _ in range(200):
        s = bytes(rng.choice(b"ACGT") for _ in range(k))
        km = oracle.kmer_from_bytes(s, k)
        assert km == H.kmer_int_to_words(H.encode_kmer(s), k)
        assert oracle.kmer_to_string(oracle.kmer_revcomp(km, k), k) == H.revcomp_string(s)
        assert oracle.kmer_canonical(km, k)[0] == H.canonical_words_of_string(s)


def test_hamming_distance(oracle):
    # hammer.rs:150-158, tests/test_hammer_and_bloom.rs:7-37
    assert oracle.hamming_2bit(0b00011011, 0b00011011) == 0
    assert oracle.hamming_2bit(0b00011011, 0b00011000) == 1
    assert oracle.hamming_2bit(0b00000000, 0b01010101) == 4
    for i in range(4):
        for j in range(4):
            assert oracle.hamming_2bit(i, j) == (0 if i == j else 1)
    assert oracle.hamming_2bit(enc(b"ACGTACGTACGTACGT", 16), enc(b"ACCTACATACGAACGT", 16)) == 3
