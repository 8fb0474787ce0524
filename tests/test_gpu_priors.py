"""GPU parity of the prior / trusted-contig boost (assemble.rs:244-284) against the
oracle: the boosted count table, cutoff and solid set are bit-exact, with Bloom tiers
untouched, for single and several libraries, with and without Hammer, with no reads at
all, and through AssemblerConfig.prior_contigs / trusted_contigs."""
import sys

import numpy as np
import pytest

from kc import AssemblerConfig, PackedReads, SpectrumEngine, count_kmers, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _genome_str(meta):
    return "".join("ACGT"[int(x)] for x in meta["genome"])


def _priors(genome, k, seed):
    rng = np.random.default_rng(seed)
    rand = "".join("ACGT"[x] for x in rng.integers(0, 4, 400))
    return [genome[100:1500], rand, genome[1600:1900].lower(), genome[2000:2150] + "N" + genome[2150:2300],
            genome[:max(k - 1, 0)], "", genome[100:400], "acgtNNNNacgt" * 8]


def _compare(eng, lib, S):
    info = eng.info(lib)
    seen, solid = eng.bloom(lib)
    keys, cnt, mask = eng.counts(lib)
    oseen, osolid = S.bloom()
    okeys, ocnt, omask = S.counts()
    assert np.array_equal(seen, oseen) and np.array_equal(solid, osolid), "Bloom tiers differ"
    assert keys.shape == okeys.shape, (keys.shape, okeys.shape)
    assert np.array_equal(keys, okeys), "sorted keys differ"
    assert np.array_equal(cnt, ocnt), "counts differ"
    assert info["cutoff"] == S.cutoff and info["n_solid"] == S.n_solid
    assert np.array_equal(mask, omask)


@pytest.mark.parametrize("k,ec,min_cov", [(21, False, 5.0), (33, True, 5.0), (55, False, 13.0), (21, False, 1e20),
                                          (63, True, 5.0), (11, False, 5.0), (32, False, 40.0), (77, True, 5.0),
                                          (127, False, 5.0)])
def test_priors_match_oracle(oracle, k, ec, min_cov):
    store, meta = synth.uniform_genome_pairs(seed=300 + k, genome_size=5000, n_pairs=400, read_len=120,
                                             insert_mean=250, insert_sd=20, p_start=0.003, p_end=0.008)
    priors = _priors(_genome_str(meta), k, k)
    bits = 1 << 18
    say(f"case priors k={k} ec={ec} min_cov={min_cov}")
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits, error_correct=ec, min_coverage=min_cov))
    eng.upload(store)
    eng.set_priors(priors)
    eng.run()
    d, o, l = store.arrays()
    S = oracle.Spectrum(d, o, l, k, bits, ec, min_cov, 1, priors=priors)
    _compare(eng, 0, S)
    # Stage 3 again over the same Bloom build gives the same table
    first = eng.counts()
    eng.count()
    again = eng.counts()
    assert all(np.array_equal(a, b) for a, b in zip(first, again))
    # clearing the priors restores the plain result
    eng.set_priors([])
    eng.run()
    _compare(eng, 0, oracle.Spectrum(d, o, l, k, bits, ec, min_cov, 1))
    eng.close()


@pytest.mark.parametrize("k", [21, 33, 99])
def test_priors_without_reads(oracle, k):
    store, meta = synth.uniform_genome_pairs(seed=7, genome_size=3000, n_pairs=10, read_len=100,
                                             insert_mean=200, insert_sd=20)
    priors = _priors(_genome_str(meta), k, 1)
    empty = PackedReads()
    say(f"case priors without reads k={k}")
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=1 << 16))
    eng.upload(empty)
    eng.set_priors(priors)
    info = eng.run()
    S = oracle.Spectrum(np.zeros(1, np.uint8), np.zeros(0, np.uint64), np.zeros(0, np.uint32), k, 1 << 16, False,
                        5.0, 1, priors=priors)
    assert info["n_counts"] == S.n_counts > 0
    _compare(eng, 0, S)
    eng.close()


def test_priors_per_library(oracle):
    k = 33
    libs, metas = [], []
    for i in range(3):
        st, meta = synth.uniform_genome_pairs(seed=900 + i, genome_size=3000 + 1000 * i, n_pairs=200 + 50 * i,
                                              read_len=110, insert_mean=230, insert_sd=20, p_start=0.005)
        libs.append(st)
        metas.append(meta)
    pri = {0: _priors(_genome_str(metas[0]), k, 0), 2: _priors(_genome_str(metas[2]), k, 2)[:3]}
    mcs = [5.0, 5.0, 20.0]
    say("case priors per library")
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=1 << 17, error_correct=True))
    for i, st in enumerate(libs):
        eng.add_library(st, bloom_bits=1 << 17, min_coverage=mcs[i])
    for i, p in pri.items():
        eng.set_priors(p, library=i)
    eng.run()
    for i, st in enumerate(libs):
        d, o, l = st.arrays()
        S = oracle.Spectrum(d, o, l, k, 1 << 17, True, mcs[i], 1, priors=pri.get(i))
        _compare(eng, i, S)
    eng.close()


def test_config_prior_and_trusted_contigs(oracle, tmp_path):
    k = 21
    store, meta = synth.uniform_genome_pairs(seed=55, genome_size=4000, n_pairs=300, read_len=100,
                                             insert_mean=200, insert_sd=20, p_start=0.004)
    g = _genome_str(meta)
    fa = tmp_path / "trusted.fa"
    fa.write_bytes(b">c1 first\n" + g[0:60].encode() + b"\n" + g[60:200].encode() + b"\r\n>c2\n\n>c3\n"
                   + g[3000:3400].encode() + b"\n")
    missing = tmp_path / "absent.fa"
    cfg = AssemblerConfig(k=k, bloom_bits=1 << 18, prior_contigs=[g[500:900]], trusted_contigs=[str(fa), str(missing)])
    priors = cfg.all_priors()
    assert priors == [g[500:900].encode(), g[0:200].encode(), g[3000:3400].encode()]
    say("case config priors")
    spec = count_kmers(store, cfg)
    d, o, l = store.arrays()
    S = oracle.Spectrum(d, o, l, k, 1 << 18, False, 5.0, 1, priors=priors)
    okeys, ocnt, omask = S.counts()
    assert np.array_equal(spec.keys, okeys) and np.array_equal(spec.counts, ocnt)
    assert np.array_equal(spec.solid_mask, omask) and spec.cutoff == S.cutoff
