"""Pins the oracle's Stage 2-3 restatement (assemble.rs:129-298, 692-735):
  * seen tier == OR over all occurrences (closed form, NumPy transcription);
  * solid tier == pure-Python sequential replay of insert_kmer256;
  * count table == brute-force multiplicities restricted to Bloom-solid k-mers;
  * determine_solid_kmer_cutoff on hand-built count distributions;
  * the multi-thread baseline agrees on everything order-free."""
from collections import Counter

import numpy as np
import pytest

from kc import synth
from oracle import hash_np as H


def windows_canonical(store, k):
    keys = []
    for i in range(len(store)):
        r = store.get_read(i)
        for j in range(len(r) - k + 1):
            keys.append(H.canonical_words_of_string(r[j:j + k]))
    return np.array(keys, dtype=np.uint64).reshape(-1, 4)


@pytest.mark.parametrize("k,bits", [(21, 1 << 16), (15, 5000), (33, 1 << 18), (55, 3000 * 64), (5, 64 * 7)])
def test_spectrum_against_brute_force(oracle, k, bits):
    store, _ = synth.uniform_genome_pairs(seed=k, genome_size=3000, n_pairs=120, read_len=100,
                                          insert_mean=200, insert_sd=20, p_start=0.01)
    d, o, l = store.arrays()
    S = oracle.Spectrum(d, o, l, k, bits, False, 5.0, 1)
    occ = windows_canonical(store, k)
    assert S.n_occ == occ.shape[0]
    seen, solid = S.bloom()
    assert np.array_equal(seen, H.seen_tier(occ, bits))
    seen_py, solid_py, _ = H.sequential_two_tier(occ, bits)
    assert np.array_equal(solid, solid_py)
    # count table
    cnt = Counter(map(tuple, occ.tolist()))
    distinct = np.array(sorted(cnt.keys(), key=lambda w: ((w[1] << 64) | w[0], (w[3] << 64) | w[2])),
                        dtype=np.uint64).reshape(-1, 4)
    mask = H.is_solid(solid, distinct, bits)
    exp_keys = distinct[mask]
    exp_cnt = np.array([cnt[tuple(w)] for w in exp_keys.tolist()], dtype=np.uint32)
    keys, counts, smask = S.counts()
    assert np.array_equal(keys, exp_keys)
    assert np.array_equal(counts, exp_cnt)
    # every k-mer seen at least twice is Bloom-solid
    multi = np.array([cnt[tuple(w)] >= 2 for w in distinct.tolist()])
    assert mask[multi].all()
    assert np.array_equal(smask.astype(bool), counts >= S.cutoff)


def test_cutoff_known_answers(oracle):
    # min_coverage > 5 takes the f64 path: max(min_cov * 0.35, 2.0) as u32
    assert oracle.cutoff([1, 2, 3], 10.0)[0] == 3
    assert oracle.cutoff([1, 2, 3], 5.5)[0] == 2
    assert oracle.cutoff([], 20.0)[0] == 7
    # `as u32` saturates: 1e20 * 0.35 -> u32::MAX; +inf likewise
    assert oracle.cutoff([1, 2], 1e20)[0] == 0xFFFFFFFF
    assert oracle.cutoff([1, 2], float("inf"))[0] == 0xFFFFFFFF
    assert oracle.cutoff([1, 2], 12345678.9)[0] == 4320987  # trunc(4320987.615)
    # <= 100 entries -> 2
    assert oracle.cutoff([50] * 100, 5.0)[0] == 2
    # top (95th percentile) < 30 -> 2
    assert oracle.cutoff([29] * 101, 5.0) == (2, 29)
    # a trough at 5 then a rebound: top = 40 -> search window clamp(40/4, 10, 35) = 10.
    # c=2..5 keep lowering the minimum (valley 5, min 10); c=6 (30) and c=7 (60) exceed
    # 2*min but are not > valley+2; c=8 (80) is -> cutoff 5
    covs = [2] * 400 + [3] * 100 + [4] * 20 + [5] * 10 + [6] * 30 + [7] * 60 + [8] * 80 + [40] * 2000
    assert oracle.cutoff(covs, 5.0) == (5, 40)
    # same trough, rebound only at c=7 (not > valley+2), then zeros to the end -> 2
    covs2 = [2] * 400 + [3] * 100 + [4] * 20 + [5] * 10 + [6] * 30 + [7] * 80 + [40] * 2000
    assert oracle.cutoff(covs2, 5.0) == (2, 40)
    # ties (h <= min) move the valley right: 2,3 tie at 5, 4,5 tie at 3 -> valley 5; 8 rebounds
    covs3 = [2] * 5 + [3] * 5 + [4] * 3 + [5] * 3 + [6] * 4 + [7] * 4 + [8] * 50 + [200] * 3000
    assert oracle.cutoff(covs3, 5.0) == (5, 200)
    # an empty bin at c=2 is the minimum; a rebound at 5 > 2+2 -> cutoff 2 through `found`
    covs4 = [3] * 10 + [4] * 10 + [5] * 10 + [9] * 50 + [200] * 3000
    assert oracle.cutoff(covs4, 5.0) == (2, 200)


def test_cutoff_brute_force_grid(oracle):
    rng = np.random.default_rng(7)
    for _ in range(200):
        n = int(rng.integers(90, 400))
        covs = rng.integers(1, int(rng.integers(3, 300)), n).astype(np.uint32)
        got, top = oracle.cutoff(covs, 5.0)
        # independent statement of assemble.rs:692-735
        v = sorted(int(c) for c in covs)
        if len(v) <= 100:
            exp = 2
        else:
            t = v[len(v) * 95 // 100]
            assert top == t
            if t < 30:
                exp = 2
            else:
                mvs = min(max(t // 4, 10), 35)
                hist = [sum(1 for c in v if c == x) for x in range(mvs + 1)]
                valley, mn, found = 2, float("inf"), False
                for c in range(2, mvs + 1):
                    if hist[c] <= mn:
                        mn, valley = hist[c], c
                    elif hist[c] > mn * 2 and c > valley + 2:
                        found = True
                        break
                exp = valley if found else 2
        assert got == exp


def test_multithread_baseline_order_free_parts(oracle):
    store, _ = synth.workload_1mb_30x(seed=5, scale=0.02)
    d, o, l = store.arrays()
    s1 = oracle.Spectrum(d, o, l, 21, 1 << 20, False, 5.0, 1)
    s8 = oracle.Spectrum(d, o, l, 21, 1 << 20, False, 5.0, 8)
    assert s1.n_occ == s8.n_occ
    assert np.array_equal(s1.bloom()[0], s8.bloom()[0])  # seen tier is an OR: order-free


def test_empty_and_short_inputs(oracle):
    S = oracle.Spectrum(np.zeros(0, np.uint8), np.zeros(0, np.uint64), np.zeros(0, np.uint32), 21, 4096, False, 5.0, 1)
    assert S.n_occ == 0 and S.n_counts == 0 and S.cutoff == 2
    from kc.packed_reads import PackedReads
    st = PackedReads()
    for r in (b"ACGT", b"", b"A" * 20):
        st.add_read(r)
    d, o, l = st.arrays()
    S = oracle.Spectrum(d, o, l, 21, 4096, True, 5.0, 1)
    assert S.n_occ == 0 and S.n_counts == 0
