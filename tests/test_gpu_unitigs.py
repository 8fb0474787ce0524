"""GPU unitig compaction (kc_unitigs_build) against the oracle's CompactedGraph::build
(graph.rs:65-234) over the same solid set and counts.

Unitigs are compared as a multiset of (canonical sequence, k-mer count, mean coverage):
the reference's walk order and a linear unitig's strand follow its hash order.  Cycles
are compared as emitted (both start at the smallest k-mer, forward strand), and the mean
coverage must be equal as a float.
"""
import sys

import numpy as np
import pytest

from kc import AssemblerConfig, KcError, PackedReads, SpectrumEngine, synth
from kc import _abi

pytestmark = pytest.mark.gpu

COMP = bytes.maketrans(b"ACGT", b"TGCA")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def canon(s):
    return min(s, s.translate(COMP)[::-1])


def oracle_unitigs(oracle, eng, k, library=0):
    keys, cnt, mask = eng.counts(library)
    ref = oracle.compacted_graph(k, keys[mask.astype(bool)], keys, cnt)
    return sorted((canon(s), m, c) for s, m, c in ref)


def gpu_unitigs(eng, library=0):
    g = eng.unitigs(library)
    out = sorted((canon(g.sequence(i)), int(g.kmers[i]), float(g.mean_cov[i])) for i in range(len(g)))
    return g, out


def run_case(oracle, store, k, min_cov=5.0, bits=1 << 22):
    say(f"case k={k} reads={len(store)}")
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, min_coverage=min_cov))
    try:
        eng.upload(store)
        info = eng.run()
        g, got = gpu_unitigs(eng)
        want = oracle_unitigs(oracle, eng, k)
        assert len(got) == len(want), (len(got), len(want))
        assert got == want
        assert int(g.kmers.sum()) == info["n_solid"]
        assert g.seq.shape[0] == info["n_solid"] + len(g) * (k - 1)
        return g, info
    finally:
        eng.close()


@pytest.mark.parametrize("k", [21, 33, 55, 77, 127])
def test_unitigs_match_oracle(oracle, k):
    store, _ = synth.uniform_genome_pairs(11, 20_000, 3_000, read_len=150, p_start=0.004)
    g, info = run_case(oracle, store, k)
    assert len(g) >= 1 and info["n_solid"] > 15_000


def test_circular_genome_gives_one_cycle(oracle):
    """Error-free reads of a circular genome (reads drawn from the genome with its first
    1000 bases appended): the solid graph is one cycle."""
    rng = np.random.Generator(np.random.PCG64(5))
    g = synth.random_genome(rng, 8_000)
    ext = np.concatenate([g, g[:1000]])
    store, _ = synth.uniform_genome_pairs(12, 0, 2_000, read_len=150, p_start=0.0, genome=ext)
    k = 21
    gr, info = run_case(oracle, store, k, min_cov=2.0)
    if info["n_solid"] == 8_000:
        assert len(gr) == 1
        s = gr.sequence(0)
        assert s[8_000:] == s[:k - 1]


def test_many_libraries_each_alone(oracle):
    k = 33
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=1 << 22, min_coverage=5.0))
    try:
        for seed in (21, 22, 23):
            store, _ = synth.uniform_genome_pairs(seed, 6_000 + 1000 * seed, 1_500, p_start=0.004)
            eng.add_library(store)
        eng.run()
        for lib in range(3):
            say(f"library {lib}")
            _, got = gpu_unitigs(eng, lib)
            assert got == oracle_unitigs(oracle, eng, k, lib)
    finally:
        eng.close()


def test_1mb_30x_full_size(oracle):
    store, _ = synth.workload_1mb_30x()
    run_case(oracle, store, 21, bits=1 << 26)


def test_errors_and_empty():
    eng = SpectrumEngine(AssemblerConfig(k=22, bloom_bits=1 << 20))
    try:
        store, _ = synth.uniform_genome_pairs(1, 3_000, 300)
        eng.upload(store)
        with pytest.raises(KcError) as e:  # before kc_count
            eng.unitigs()
        assert e.value.status == _abi.KC_ESTATE
        eng.run()
        with pytest.raises(KcError) as e:  # even k
            eng.unitigs()
        assert e.value.status == _abi.KC_EUNSUPPORTED
        with pytest.raises(KcError):
            eng.unitigs(library=5)
    finally:
        eng.close()
    eng = SpectrumEngine(AssemblerConfig(k=21, bloom_bits=1 << 20))
    try:  # reads shorter than k: no k-mers at all
        data = np.zeros(4, np.uint8)
        eng.upload(PackedReads.from_arrays(data, np.array([0], np.uint64), np.array([10], np.uint32)))
        eng.run()
        g = eng.unitigs()
        assert len(g) == 0 and g.seq.shape[0] == 0
    finally:
        eng.close()
