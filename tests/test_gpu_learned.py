"""GPU: the learned first pass's features against their Python restatement
(oracle/features_np.py), and the selection's contract: inside the margin the label is the
exact rule count >= cutoff, outside it the network's decision."""
import sys

import numpy as np
import pytest

from kc import AssemblerConfig, SpectrumEngine, synth
from kc.learn import Mlp
from oracle import features_np as F

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine(store, quals, qoff, k, bits=1 << 20):
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=bits))
    eng.upload(store)
    eng.set_quals(quals, qoff)
    eng.run()
    return eng


@pytest.mark.parametrize("k", [21, 33])
def test_features_match_oracle(k):
    store, quals, meta = synth.repeat_pairs_with_quals(31 + k, genome_size=12_000, coverage=25.0)
    say(f"case features k={k} reads={len(store)}")
    eng = _engine(store, quals, meta["qoff"], k)
    info = eng.info()
    keys, cnt, _ = eng.counts()
    f = eng.features()
    cw, qs = F.window_stats(store, k, quals, meta["qoff"])
    ref = F.features(keys, cnt, k, info["cutoff"], info["top_cov"], cw, qs)
    assert f.shape == ref.shape and f.shape[0] > 100
    err = np.abs(f - ref)
    assert err.max() < 1e-4 * (1 + np.abs(ref).max()), (err.max(), np.unravel_index(err.argmax(), err.shape))
    # without qualities the quality features are zero and the rest unchanged
    eng.set_quals(np.zeros(0, np.uint8), np.zeros(0, np.uint64))
    f0 = eng.features()
    assert np.all(f0[:, 1] == 0) and np.all(f0[:, 10] == 0)
    assert np.array_equal(np.delete(f0, [1, 10], axis=1), np.delete(f, [1, 10], axis=1))
    eng.close()


def test_learned_select_contract():
    k = 21
    train_store, tq, tm = synth.repeat_pairs_with_quals(71, genome_size=40_000, coverage=30.0)
    test_store, sq, sm = synth.repeat_pairs_with_quals(72, genome_size=40_000, coverage=30.0)
    eng = _engine(train_store, tq, tm["qoff"], k)
    Xtr = eng.features()
    _, ctr, _ = eng.counts()
    ytr = (ctr >= eng.info()["cutoff"]).astype(np.uint8)
    eng.close()
    m = Mlp(seed=3)
    m.train(Xtr, ytr, epochs=4, batch=1024, lr=3e-3, seed=1)
    eng = _engine(test_store, sq, sm["qoff"], k)
    X = eng.features()
    _, cnt, mask = eng.counts()
    cutoff = eng.info()["cutoff"]
    p = m.predict(X)
    for margin in (0.0, 0.1, 0.49):
        say(f"case learned select margin={margin}")
        lab, fb, ms = eng.learned_select(m, margin)
        m32 = np.float32(margin)
        inside = (p > m32) & (p < np.float32(1) - m32)
        assert np.array_equal(fb, inside)
        assert np.array_equal(lab[fb], cnt[fb] >= cutoff)
        assert np.array_equal(lab[~fb], p[~fb] >= 0.5)
        assert ms > 0
    # a margin of 0.5 is rejected; the exact rule alone is the solid mask
    with pytest.raises(Exception):
        eng.learned_select(m, 0.5)
    assert np.array_equal(mask.astype(bool), cnt >= cutoff)
    eng.close()
    m.close()


def _labelled(seed, genome_size, k=21):
    """Uncorrected table features and the exact Hammer-on labels of one library."""
    store, quals, meta = synth.repeat_pairs_with_quals(seed, genome_size=genome_size, coverage=60.0)
    ex = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=1 << 29, error_correct=True))
    ex.upload(store)
    ex.run()
    kex, _, mex = ex.counts()
    judge = {tuple(int(x) for x in w) for w in kex[mex.astype(bool)]}
    ex.close()
    eng = SpectrumEngine(AssemblerConfig(k=k, bloom_bits=1 << 29))
    eng.upload(store)
    info = eng.run()
    eng.set_quals(quals, meta["qoff"])
    keys, cnt, _ = eng.counts()
    X = eng.features()
    y = np.array([tuple(int(x) for x in w) in judge for w in keys], dtype=np.uint8)
    return eng, X, y, cnt, info["cutoff"]


def test_network_gain_over_the_count_rule():
    """What the trained network gains over the method without it: on held-out libraries
    whose cutoff correction moves (no valley before Hammer, one after), the plain rule
    count >= uncorrected cutoff keeps every low-count k-mer and misses about 10% of the
    exact labels; the network, trained on other seeds' exact Hammer-on labels, reaches
    the exact labels.  Both libraries are ones the library-level margin would send to
    the exact selection (kc_library_cutoff_fragile on the device's histogram, and
    kc_cutoff_is_fragile on the exported counts), which the test also checks."""
    from kc.learn import cutoff_is_fragile
    g = 240_000
    Xs, ys = [], []
    for s in range(8):
        eng, X, y, _, _ = _labelled(1000 + s, g)
        eng.close()
        Xs.append(X)
        ys.append(y)
    m = Mlp(seed=1)
    m.train(np.concatenate(Xs), np.concatenate(ys), epochs=8, batch=4096, lr=2e-3, seed=7)
    for s in (5001, 5003):
        say(f"case learned gain seed={s}")
        eng, X, y, cnt, cutoff = _labelled(s, g)
        lab, fb, _ = eng.learned_select(m, 0.05)
        net = float(np.mean(lab == y))
        plain = float(np.mean((cnt >= cutoff) == y))
        say(f"  cutoff {cutoff}: network {net:.5f}, count rule {plain:.5f}, fallback {fb.mean():.2e}")
        assert cutoff == 2 and cutoff_is_fragile(cnt) and eng.cutoff_fragile()
        assert net >= 0.999 and net >= plain + 0.05
        eng.close()
    m.close()


@pytest.mark.parametrize("seed,cov,min_cov", [(11, 15.0, 5.0), (12, 60.0, 5.0), (13, 200.0, 5.0), (14, 60.0, 12.0)])
def test_library_margin_matches_host_rule(seed, cov, min_cov):
    """kc_library_cutoff_fragile (the device count histogram of a counted library) gives the
    decision kc_cutoff_is_fragile gives for the library's exported counts, for any seed."""
    from kc.learn import cutoff_is_fragile
    g = 200_000
    store, _ = synth.uniform_genome_pairs(seed=seed, genome_size=g, n_pairs=int(cov * g / 300), p_start=0.005)
    say(f"case library margin seed={seed} cov={cov} min_cov={min_cov}")
    eng = SpectrumEngine(AssemblerConfig(k=21, bloom_bits=1 << 24, min_coverage=min_cov))
    eng.upload(store)
    eng.run()
    _, cnt, _ = eng.counts()
    for rs in (0, 3):
        assert eng.cutoff_fragile(seed=rs) == cutoff_is_fragile(cnt, min_coverage=min_cov, seed=rs)
    if min_cov > 5.0:
        assert not eng.cutoff_fragile()
    eng.close()
