"""GPU: the unique-anchor k-mer indexes (polisher.rs:33-60, paired_info.rs:34-70,
splitter.rs:57-70) against the oracle's restatement of the reference loops, bit-exact:
the sorted anchors, their sequence, position and strand; lookups of every anchor, of
dropped repeats and of absent k-mers."""
import sys

import numpy as np
import pytest

from kc import KcError, _abi
from kc.anchor import AnchorIndex
from test_oracle_anchor import random_contigs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", ["polisher", "paired_info", "splitter"])
@pytest.mark.parametrize("k", [5, 10, 11, 21, 33, 55, 64, 77, 127])
def test_anchor_index_matches_oracle(oracle, kind, k):
    rng = np.random.default_rng(1000 + k)
    seqs = random_contigs(rng, 60, genome_len=4000)
    seqs += ["".join(rng.choice(list("ACGT"), 400)) for _ in range(3)]
    say(f"case anchor {kind} k={k}")
    ok, os_, op, orc = oracle.anchor_index(kind, k, seqs)
    ix = AnchorIndex(kind, k, seqs)
    keys, seq, pos, rc = ix.export()
    assert keys.shape == ok.shape
    assert np.array_equal(keys, ok) and np.array_equal(seq, os_) and np.array_equal(pos, op)
    assert np.array_equal(rc, orc)
    # every anchor is found at its own index; dropped and absent k-mers are not
    if len(keys):
        assert np.array_equal(ix.query(keys), np.arange(len(keys)))
    probes = []
    for s in seqs[:20]:
        u = s.upper()
        for i in range(0, max(len(u) - k + 1, 0), 7):
            w = oracle.kmer_from_bytes(u[i:i + k].encode(), k)
            if w is not None:
                probes.append(oracle.kmer_canonical(w, k)[0])
    if probes:
        probes = np.array(probes, dtype=np.uint64)
        got = ix.query(probes)
        have = {tuple(int(x) for x in w): j for j, w in enumerate(keys)}
        exp = np.array([have.get(tuple(int(x) for x in w), -1) for w in probes])
        assert np.array_equal(got, exp)
    ix.close()


def test_anchor_index_larger_input(oracle):
    rng = np.random.default_rng(5)
    g = "".join(rng.choice(list("ACGT"), 200_000))
    seqs = [g[a:a + int(rng.integers(500, 5000))] for a in rng.integers(0, 190_000, 300)]
    for kind in ("paired_info", "splitter"):
        ok, os_, op, orc = oracle.anchor_index(kind, 31, seqs)
        keys, seq, pos, rc = AnchorIndex(kind, 31, seqs).export()
        assert np.array_equal(keys, ok) and np.array_equal(seq, os_) and np.array_equal(pos, op)
        assert np.array_equal(rc, orc)


def test_anchor_edge_cases(oracle):
    assert AnchorIndex("paired_info", 21, []).n == 0
    assert AnchorIndex("polisher", 21, ["ACGT", "", "NNNNNNNNNNNNNNNNNNNNNNNNN"]).n == 0
    ix = AnchorIndex("splitter", 3, ["ACGTT"])
    assert ix.n == oracle.anchor_index("splitter", 3, ["ACGTT"])[0].shape[0]
    with pytest.raises(KcError) as e:
        AnchorIndex("polisher", 128, ["A" * 200])
    assert e.value.status == _abi.KC_EUNSUPPORTED
    with pytest.raises(KcError) as e:
        AnchorIndex("polisher", 0, ["ACGT"])
    assert e.value.status == _abi.KC_EINVAL
