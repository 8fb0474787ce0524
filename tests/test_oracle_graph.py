"""The oracle's CompactedGraph::build (graph.rs:65-234), pinned without the reference.

The reference's own tests build Unitig values by hand and never call build(), so the
restatement is pinned here by (a) constructions whose unitigs are known in closed form
(one linear genome, one circular genome, a hairpin, a self loop, a repeat shared by two
genomes) and (b) the defining properties of maximal non-branching paths, checked with
plain string operations on every result: each solid k-mer in exactly one unitig, every
step inside a unitig a link (unique successor / unique predecessor, another k-mer), no
unitig end that a link would extend except where a cycle closes, a cycle starting at its
smallest k-mer on the forward strand (the reference's pass 2 in ascending order), and the
mean coverage equal to the mean count.
"""
import random
from collections import Counter

import numpy as np
import pytest

COMP = str.maketrans("ACGT", "TGCA")


def rc(s):
    return s.translate(COMP)[::-1]


def canon(s):
    return min(s, rc(s))


def code(s):
    v = 0
    for ch in s:
        v = (v << 2) | "ACGT".index(ch)
    return v


def limbs(s):
    """Kmer256 limbs (dna.rs:134-138): k <= 64 -> (V, 0); else (first 64 bases, rest)."""
    k = len(s)
    if k <= 64:
        v = code(s)
        return [v & (2**64 - 1), v >> 64, 0, 0]
    a, b = code(s[:64]), code(s[64:])
    return [a & (2**64 - 1), a >> 64, b & (2**64 - 1), b >> 64]


def kmers_of(seq, k, circular=False):
    s = seq + (seq[:k - 1] if circular else "")
    return {canon(s[i:i + k]) for i in range(len(s) - k + 1)}


def build(oracle, k, solid, counts=None):
    solid = sorted(solid)
    keys = np.array([limbs(s) for s in solid], dtype=np.uint64).reshape(-1, 4)
    if counts is None:
        counts = {s: 7 for s in solid}
    ck = sorted(counts)
    ckeys = np.array([limbs(s) for s in ck], dtype=np.uint64).reshape(-1, 4)
    cv = np.array([counts[s] for s in ck], dtype=np.uint32)
    return [(s.decode(), m, c) for s, m, c in oracle.compacted_graph(k, keys, ckeys, cv)]


def succs(s, solid):
    return [s[1:] + b for b in "ACGT" if canon(s[1:] + b) in solid]


def preds(s, solid):
    return [a + s[:-1] for a in "ACGT" if canon(a + s[:-1]) in solid]


def link(a, b, solid):
    """a -> b in the reference's walk: b is a's only successor, a is b's only predecessor,
    and b is another canonical k-mer."""
    sa = succs(a, solid)
    return len(sa) == 1 and sa[0] == b and len(preds(b, solid)) == 1 and canon(a) != canon(b)


def check_unitigs(k, solid, counts, unitigs):
    solid = set(solid)
    seen = Counter()
    for seq, m, cov in unitigs:
        assert len(seq) == m + k - 1
        kms = [seq[i:i + k] for i in range(m)]
        for x in kms:
            seen[canon(x)] += 1
        want = sum(counts.get(canon(x), 2) for x in kms) / m
        assert cov == want
        for a, b in zip(kms, kms[1:]):
            assert link(a, b, solid), (a, b)
        last, first = kms[-1], kms[0]
        cyc = [s for s in succs(last, solid) if link(last, s, solid)]
        if cyc:  # only a closing cycle may continue past the end
            assert cyc == [first]
            assert first == canon(first) and canon(first) == min(canon(x) for x in kms)
        for p in preds(first, solid):
            if link(p, first, solid):
                assert p == last
    assert set(seen) == solid and all(v == 1 for v in seen.values())


def random_seq(rng, n):
    return "".join(rng.choice("ACGT") for _ in range(n))


def test_linear_genome_is_one_unitig(oracle):
    rng = random.Random(1)
    for k in (21, 33, 55, 77):
        g = random_seq(rng, 600)
        solid = kmers_of(g, k)
        assert len(solid) == len(g) - k + 1
        counts = {s: rng.randrange(1, 90) for s in solid}
        u = build(oracle, k, solid, counts)
        assert len(u) == 1 and canon(u[0][0]) == canon(g) and u[0][1] == len(g) - k + 1
        check_unitigs(k, solid, counts, u)


def test_circular_genome_is_one_cycle(oracle):
    rng = random.Random(2)
    for k in (21, 55):
        g = random_seq(rng, 500)
        solid = kmers_of(g, k, circular=True)
        u = build(oracle, k, solid)
        assert len(u) == 1
        seq, m, _ = u[0]
        assert m == len(g) and seq[len(g):] == seq[:k - 1]
        body = seq[:len(g)]
        assert body in g + g or body in rc(g) + rc(g)
        assert seq[:k] == min(solid)  # the smallest k-mer, forward strand
        check_unitigs(k, solid, {s: 7 for s in solid}, u)


def test_hairpin_and_self_loop_are_single_kmers(oracle):
    # AACGT's successor ACGTT is its own reverse complement: a link onto itself, never followed
    assert rc("AACGT") == "ACGTT"
    for solid in ({"AACGT"}, {"AAAAA"}, {"AACGT", "AAAAA"}):
        u = build(oracle, 5, solid)
        assert sorted(canon(s) for s, _, _ in u) == sorted(solid)
        assert all(m == 1 for _, m, _ in u)
        check_unitigs(5, solid, {s: 7 for s in solid}, u)


def test_shared_repeat_splits_into_five(oracle):
    rng = random.Random(3)
    k = 21
    r = random_seq(rng, 200)
    x, y, z, w = (random_seq(rng, 150) for _ in range(4))
    solid = kmers_of(x + r + y, k) | kmers_of(z + r + w, k)
    u = build(oracle, k, solid)
    assert len(u) == 5
    assert sum(m for _, m, _ in u) == len(solid)
    assert canon(r[1:-1]) in {canon(s) for s, _, _ in u} or any(r[1:-1] in s or r[1:-1] in rc(s) for s, _, _ in u)
    check_unitigs(k, solid, {s: 7 for s in solid}, u)


def test_missing_coverage_defaults_to_two(oracle):
    rng = random.Random(4)
    g = random_seq(rng, 100)
    solid = kmers_of(g, 21)
    u = build(oracle, 21, solid, counts={})
    assert len(u) == 1 and u[0][2] == 2.0


def test_empty_and_k0(oracle):
    assert build(oracle, 21, set()) == []
    keys = np.zeros((1, 4), np.uint64)
    assert oracle.compacted_graph(0, keys, keys, np.ones(1, np.uint32)) == []


@pytest.mark.parametrize("seed", range(6))
def test_random_graphs_satisfy_the_definition(oracle, seed):
    """Dense random k-mer sets at small k (hairpins, self loops, branches everywhere) and
    repeat-rich genomes, linear and circular."""
    rng = random.Random(100 + seed)
    for trial in range(40):
        mode = trial % 3
        if mode == 0:
            k = rng.choice([1, 3, 5, 7])
            solid = {canon(random_seq(rng, k)) for _ in range(rng.randrange(1, min(4**k, 300)))}
        else:
            k = rng.choice([9, 11, 21, 33, 77])
            g = random_seq(rng, rng.randrange(k, 400))
            solid = kmers_of(g, k, circular=(mode == 2))
            for _ in range(rng.randrange(0, 4)):
                a = rng.randrange(len(g))
                alt = random_seq(rng, rng.randrange(0, 30))
                s = alt + g[a:a + rng.randrange(k, 3 * k)] + alt[::-1]
                if len(s) >= k:
                    solid |= kmers_of(s, k)
        counts = {s: rng.randrange(1, 1000) for s in solid}
        check_unitigs(k, solid, counts, build(oracle, k, solid, counts))


def test_plausible_mistakes_are_caught(oracle):
    """The property check rejects a unitig with a wrong last base, a reversed strand
    inside a unitig, a dropped k-mer and a wrong mean."""
    rng = random.Random(5)
    k = 21
    g = random_seq(rng, 300)
    solid = kmers_of(g, k)
    counts = {s: rng.randrange(1, 90) for s in solid}
    (seq, m, cov), = build(oracle, k, solid, counts)
    bad = [
        [(seq[:-1] + ("A" if seq[-1] != "A" else "C"), m, cov)],
        [(seq[:100] + rc(seq[100:]), m, cov)],
        [(seq[:-1], m - 1, cov)],
        [(seq, m, cov + 0.5)],
    ]
    for u in bad:
        with pytest.raises(AssertionError):
            check_unitigs(k, solid, counts, u)
