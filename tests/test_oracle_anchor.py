"""Pins the oracle's unique-anchor k-mer indexes (polisher.rs:33-60, paired_info.rs:34-70,
splitter.rs:57-70) against an order-free characterisation written independently here: a
k-mer is kept when it occurs once in all (polisher at k >= 11, paired_info), or when all
its occurrences lie in one sequence (splitter, polisher at k < 11); its value is its first
occurrence.  Windows holding a non-ACGT base are skipped, lower case counts as upper."""
import numpy as np
import pytest

COMP = {ord("A"): "T", ord("C"): "G", ord("G"): "C", ord("T"): "A"}


def canonical_str(w):
    rc = "".join(COMP[c] for c in reversed(w.encode()))
    return (w, False) if w <= rc else (rc, True)  # 'A' < 'C' < 'G' < 'T' as bytes


def expected(kind, k, seqs):
    occ = {}
    for si, s in enumerate(seqs):
        u = s.upper()
        for i in range(len(u) - k + 1):
            w = u[i:i + k]
            if any(c not in "ACGT" for c in w):
                continue
            can, rc = canonical_str(w)
            occ.setdefault(can, []).append((si, i, rc))
    unique = kind == "paired_info" or (kind == "polisher" and k >= 11)
    out = {}
    for can, lst in occ.items():
        if unique and len(lst) != 1:
            continue
        if not unique and len({x[0] for x in lst}) != 1:
            continue
        out[can] = lst[0]
    return out


def random_contigs(rng, n, genome_len=3000):
    g = "".join(rng.choice(list("ACGT"), genome_len))
    seqs = []
    for i in range(n):
        a = int(rng.integers(0, genome_len - 200))
        s = list(g[a:a + int(rng.integers(5, 200))])
        if i % 3 == 0 and len(s) > 60:  # an internal repeat
            s[30:45] = s[0:15]
        if i % 4 == 1 and s:
            s[int(rng.integers(0, len(s)))] = "N"
        if i % 5 == 2:
            s = [c.lower() for c in s]
        seqs.append("".join(s))
    seqs.append("ACGTACGTACGTACGTACGTACGTACGTACGT")  # tandem repeat of a palindrome
    seqs.append("")
    return seqs


@pytest.mark.parametrize("kind", ["polisher", "paired_info", "splitter"])
@pytest.mark.parametrize("k", [4, 5, 10, 11, 21, 31, 33, 64, 65, 99])
def test_anchor_index_matches_characterisation(oracle, kind, k):
    rng = np.random.default_rng(k * 7 + len(kind))
    seqs = random_contigs(rng, 40)
    keys, seq, pos, rc = oracle.anchor_index(kind, k, seqs)
    exp = expected(kind, k, seqs)
    assert len(keys) == len(exp)
    got = {}
    for j in range(len(keys)):
        s = oracle.kmer_to_string(tuple(int(x) for x in keys[j]), k).decode()
        got[s] = (int(seq[j]), int(pos[j]), bool(rc[j]))
    assert got == exp
    # the export is sorted by k-mer, which is the string order of the canonical k-mers
    assert list(got) == sorted(got)


def test_polisher_modes_differ_below_k11(oracle):
    # a k-mer repeated inside one contig: kept at k < 11 (first occurrence), dropped at k >= 11
    seqs = ["AAAACCCCGGTTAAAACCCCGGTT" + "ACGGT", "TTTGGGCA"]
    keys10, seq10, pos10, _ = oracle.anchor_index("polisher", 10, seqs)
    keys11, *_ = oracle.anchor_index("polisher", 11, seqs)
    k10 = {oracle.kmer_to_string(tuple(int(x) for x in w), 10).decode(): (int(s), int(p))
           for w, s, p in zip(keys10, seq10, pos10)}
    assert k10[canonical_str("AAAACCCCGG")[0]] == (0, 0)
    k11 = {oracle.kmer_to_string(tuple(int(x) for x in w), 11).decode() for w in keys11}
    assert canonical_str("AAAACCCCGGT")[0] not in k11
