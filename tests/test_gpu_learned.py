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
