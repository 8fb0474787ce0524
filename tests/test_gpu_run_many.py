"""GPU parity of kc_run_many: several contexts (one per k, and a Hammer-on one) run
together from their own host threads give, context by context, the oracle's outputs on
the same reads -- the same checks as one context run alone."""
import sys

import numpy as np
import pytest

from kc import AssemblerConfig, KcError, SpectrumEngine, run_many, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_run_many_matches_oracle_per_context(oracle):
    store, _ = synth.uniform_genome_pairs(seed=31, genome_size=60000, n_pairs=12000, p_start=0.002, p_end=0.01)
    d, o, l = store.arrays()
    cases = [(21, 1 << 20, False), (33, 1 << 18, False), (55, 1 << 20, False), (31, 1 << 16, True)]
    engs = []
    try:
        for k, bits, ec in cases:
            e = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec))
            e.upload(store)
            engs.append(e)
        for rep in range(2):  # a second call reuses every context's buffers
            say(f"run_many rep {rep}: {[c[0] for c in cases]}")
            infos = run_many(engs)
            for (k, bits, ec), e, info in zip(cases, engs, infos):
                S = oracle.Spectrum(d, o, l, k, bits, ec, 5.0, 1)
                seen, solid = e.bloom()
                keys, cnt, mask = e.counts()
                oseen, osolid = S.bloom()
                okeys, ocnt, omask = S.counts()
                assert info["n_occ"] == S.n_occ and info["cutoff"] == S.cutoff, (k, ec)
                assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid), (k, ec)
                assert np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt), (k, ec)
                assert np.array_equal(mask, omask), (k, ec)
                if ec:
                    assert info["corrected_reads"] == S.corrected_reads
                    assert np.array_equal(e.reads(), S.corrected_data())
    finally:
        for e in engs:
            e.close()


def test_run_many_rejects_a_repeated_context():
    store, _ = synth.uniform_genome_pairs(seed=32, genome_size=5000, n_pairs=200)
    e = SpectrumEngine(AssemblerConfig(k=21, bloom_bits=1 << 16))
    try:
        e.upload(store)
        with pytest.raises(KcError, match="listed twice"):
            run_many([e, e])
        assert run_many([]) == []
    finally:
        e.close()
