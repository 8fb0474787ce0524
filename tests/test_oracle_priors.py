"""Pins the oracle's prior / trusted-contig boost (assemble.rs:244-284):
  * prior_boost = (min_coverage as u32 * 4).max(50), with Rust's saturating float cast
    and wrapping u32 multiply (known answers);
  * counts with priors == counts without priors + boost x (prior windows per canonical
    k-mer), with absent k-mers inserted; the boost is computed here by an independent
    Python enumeration of the prior windows (N windows skipped, lowercase accepted,
    sequences shorter than k skipped);
  * the cutoff and solid set are recomputed over the boosted counts."""
from collections import Counter

import numpy as np
import pytest

from kc import synth
from oracle import hash_np as H

_ACGT = set(b"ACGTacgt")


def prior_windows(priors, k):
    """Canonical limbs of every valid window of every prior (from_bytes: all k bases ACGT/acgt)."""
    out = Counter()
    for p in priors:
        b = p.encode() if isinstance(p, str) else bytes(p)
        for i in range(len(b) - k + 1):
            w = b[i:i + k]
            if all(c in _ACGT for c in w):
                out[tuple(H.canonical_words_of_string(w.upper()))] += 1
    return out


def test_prior_boost_known_answers(oracle):
    assert oracle.prior_boost(5.0) == 50          # 20 -> max(., 50)
    assert oracle.prior_boost(12.7) == 50         # as u32 truncates: 12 * 4 = 48
    assert oracle.prior_boost(13.0) == 52
    assert oracle.prior_boost(100.0) == 400
    assert oracle.prior_boost(float("nan")) == 50  # NaN as u32 == 0
    assert oracle.prior_boost(-7.0) == 50          # negative saturates to 0
    assert oracle.prior_boost(2.0 ** 30) == 50     # 2^32 wraps to 0
    assert oracle.prior_boost(1e20) == 0xFFFFFFFC  # saturates to u32::MAX, * 4 wraps


def _key_int(w):
    return ((int(w[1]) << 64) | int(w[0]), (int(w[3]) << 64) | int(w[2]))


@pytest.mark.parametrize("k,ec,min_cov", [(21, False, 5.0), (33, True, 5.0), (55, False, 13.0), (21, False, 1e20)])
def test_priors_add_boost_per_window(oracle, k, ec, min_cov):
    store, meta = synth.uniform_genome_pairs(seed=100 + k, genome_size=4000, n_pairs=300, read_len=100,
                                             insert_mean=200, insert_sd=20, p_start=0.004, p_end=0.004)
    genome = "".join("ACGT"[int(x)] for x in meta["genome"])
    rng = np.random.default_rng(k)
    rand = "".join("ACGT"[x] for x in rng.integers(0, 4, 300))
    priors = [
        genome[100:900],                                  # k-mers already counted
        rand,                                             # k-mers absent from the reads
        genome[1000:1200].lower(),                        # lowercase is valid (dna.rs:11-20)
        genome[2000:2100] + "N" + genome[2100:2200],      # windows over the N are skipped
        genome[:k - 1],                                   # shorter than k: skipped
        "",
        genome[100:300],                                  # repeats an earlier prior
    ]
    d, o, l = store.arrays()
    bits = 1 << 18
    base = oracle.Spectrum(d, o, l, k, bits, ec, min_cov, 1)
    S = oracle.Spectrum(d, o, l, k, bits, ec, min_cov, 1, priors=priors)
    # the Bloom tiers do not see the priors
    b0, b1 = base.bloom(), S.bloom()
    assert np.array_equal(b0[0], b1[0]) and np.array_equal(b0[1], b1[1])
    boost = oracle.prior_boost(min_cov)
    add = prior_windows(priors, k)
    assert sum(add.values()) > 0
    exp = {tuple(w): int(c) for w, c in zip(base.counts()[0].tolist(), base.counts()[1].tolist())}
    for w, n in add.items():
        exp[w] = (exp.get(w, 0) + n * boost) & 0xFFFFFFFF
    keys, counts, mask = S.counts()
    order = sorted(exp, key=_key_int)
    assert np.array_equal(keys, np.array(order, dtype=np.uint64).reshape(-1, 4))
    assert np.array_equal(counts, np.array([exp[w] for w in order], dtype=np.uint32))
    cut, _ = oracle.cutoff(counts, min_cov)
    assert S.cutoff == cut
    assert np.array_equal(mask.astype(bool), counts >= cut)
    # k-mers absent from the reads enter with exactly n * boost
    absent = [w for w in add if w not in {tuple(x) for x in base.counts()[0].tolist()}]
    assert absent
