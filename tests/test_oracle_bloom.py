"""Pins the oracle's bloom.rs restatement: the reference's filter tests
(bloom.rs:138-156, tests/test_hammer_and_bloom.rs:39-70), the golden hash
vectors (tests/golden/hash_kat.json, written by the independent NumPy
transcription), agreement of the two transcriptions on random keys, the
closed form of the seen tier, and a pure-Python sequential replay of
insert_kmer256 for the solid tier."""
import json
import os

import numpy as np
import pytest

from oracle import hash_np as H

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_two_tier_filter_reference_test(oracle):
    # bloom.rs:138-156
    f = oracle.TwoTierFilter(1024 * 1024)
    assert not f.is_solid(123456789)
    f.insert(123456789)
    assert not f.is_solid(123456789)
    f.insert(123456789)
    assert f.is_solid(123456789)
    assert not f.is_solid(987654321)


def test_two_tier_bloom_transitions(oracle):
    # tests/test_hammer_and_bloom.rs:39-70
    f = oracle.TwoTierFilter(1024 * 1024)
    a, b, c = 0x123456789ABCDEF0, 0xFEDCBA9876543210, 0xA5A5A5A5A5A5A5A5
    assert not any(f.is_solid(x) for x in (a, b, c))
    assert f.insert(a) is False
    assert not f.is_solid(a)
    assert f.insert(a) is True
    assert f.is_solid(a)
    assert f.insert(a) is True
    assert not f.is_solid(b) and not f.is_solid(c)


def test_sizing_rounds_up_to_words(oracle):
    # bloom.rs:26-28
    for bits, words in ((0, 1), (1, 1), (64, 1), (65, 2), (1 << 20, 1 << 14), ((1 << 20) + 100, (1 << 14) + 2)):
        assert oracle.TwoTierFilter(bits).num_words == words
        assert H.size_bits_for(bits) == (64 * words, words)


def test_hash_golden_vectors(oracle):
    with open(os.path.join(GOLDEN, "hash_kat.json")) as fh:
        kat = json.load(fh)
    seeds = [int(s) for s in kat["seeds"]]
    assert seeds == [int(s) for s in H.SEEDS] == list(oracle.SEEDS)
    for c in kat["cases"]:
        limbs = [int(x) for x in c["limbs"]]
        if c["k"]:
            assert oracle.kmer_canonical(oracle.kmer_from_bytes(c["seq"].encode(), c["k"]), c["k"])[0] == tuple(limbs)
        for j, s in enumerate(seeds):
            assert oracle.hash_kmer256(limbs, s) == int(c[f"h{j + 1}"])


@pytest.mark.parametrize("k", [21, 31, 33, 55, 63, 65, 127])
def test_two_transcriptions_agree(oracle, k):
    rng = np.random.default_rng(k)
    n = 20000
    keys = np.zeros((n, 4), dtype=np.uint64)
    bits = 2 * k
    keys[:, 0] = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    if bits < 64:
        keys[:, 0] &= np.uint64((1 << bits) - 1)
    if bits > 64:
        keys[:, 1] = rng.integers(0, 2**63, n, dtype=np.uint64)
        if bits < 128:
            keys[:, 1] &= np.uint64((1 << (bits - 64)) - 1)
    if bits > 128:
        keys[:, 2] = rng.integers(0, 2**63, n, dtype=np.uint64) & np.uint64((1 << min(bits - 128, 63)) - 1)
    for s in H.SEEDS:
        assert np.array_equal(oracle.hash_many(keys, int(s)), H.hash_kmer256(keys[:, 0], keys[:, 1], keys[:, 2],
                                                                             keys[:, 3], s))


@pytest.mark.parametrize("bits", [64, 1000, 4096, 1 << 16, (1 << 16) + 192])
def test_sequential_insert_matches_python_replay(oracle, bits):
    # keys drawn from a small pool so repeats, promotions and bit collisions all occur
    rng = np.random.default_rng(bits)
    pool = np.zeros((300, 4), dtype=np.uint64)
    pool[:, 0] = rng.integers(0, 2**62, 300, dtype=np.uint64)
    pool[:50, 1] = rng.integers(0, 2**40, 50, dtype=np.uint64)
    occ = pool[rng.integers(0, 300, 2000)]
    f = oracle.TwoTierFilter(bits)
    prom = f.insert_many(occ)
    seen, solid = f.export()
    seen_py, solid_py, prom_py = H.sequential_two_tier(occ, bits)
    assert np.array_equal(seen, seen_py)
    assert np.array_equal(solid, solid_py)
    assert np.array_equal(prom.astype(bool), prom_py)
    assert np.array_equal(seen, H.seen_tier(occ, bits))
    assert np.array_equal(f.is_solid_many(pool).astype(bool), H.is_solid(solid, pool, bits))
