"""Pins the NumPy restatement of the learned first pass's network (oracle/mlp_np.py):
  * bf16 rounding known answers (ties to even, carries into the exponent);
  * the analytic gradient equals central finite differences of the loss (with the bf16
    roundings off, where the loss is differentiable), for every parameter block;
  * p is the logistic of the logit, and the loss is the weighted BCE computed directly."""
import numpy as np

from oracle import mlp_np as M


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 65504.0, 1.0 + 2 ** -8 + 2 ** -20,
                  (2 - 2 ** -8)], dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 65536.0, 1.0 + 2 ** -7, 2.0], dtype=np.float32)
    assert np.array_equal(M.bf16(x), want)
    # every bf16 value is a fixed point
    v = np.arange(0, 1 << 16, 7, dtype=np.uint32) << 16
    f = v.view(np.float32)
    f = f[np.isfinite(f)]
    assert np.array_equal(M.bf16(f).view(np.uint32), f.view(np.uint32))


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = M.init_params(1).astype(np.float64)
    X = rng.normal(0, 1, (40, M.N_IN))
    y = rng.integers(0, 2, 40)
    sw = rng.uniform(0.5, 2.0, 40)

    def L(wv):
        return M.loss_and_grad(wv, X, y, sw, pos_weight=1.7, rounding=False)[0]

    _, g = M.loss_and_grad(w, X, y, sw, pos_weight=1.7, rounding=False)
    idx = list(range(0, M.OFF_B1, 97)) + list(range(M.OFF_B1, M.OFF_W2, 13)) + \
        list(range(M.OFF_W2, M.OFF_B2, 311)) + list(range(M.OFF_B2, M.OFF_W3, 17)) + \
        list(range(M.OFF_W3, M.OFF_B3, 9)) + [M.OFF_B3]
    for i in idx:
        e = np.zeros_like(w)
        e[i] = 1e-5
        fd = (L(w + e) - L(w - e)) / 2e-5
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_forward_is_logistic_of_logit_and_loss_is_bce():
    rng = np.random.default_rng(5)
    w = M.init_params(2)
    X = rng.normal(0, 1, (50, M.N_IN)).astype(np.float32)
    y = rng.integers(0, 2, 50)
    p, inter = M.forward(w, X)
    z3 = inter[-1]
    assert np.allclose(p, 1 / (1 + np.exp(-z3)))
    loss, _ = M.loss_and_grad(w, X, y)
    direct = -np.mean(y * np.log(p) + (1 - y) * np.log(1 - p))
    assert abs(loss - direct) < 1e-9


def _valley(hist):
    """The valley search of determine_solid_kmer_cutoff (assemble.rs:692-735) restated
    over hist[0..mvs]: (cutoff, found)."""
    valley, low = 2, None
    for c in range(2, len(hist)):
        h = int(hist[c])
        if low is None or h <= low:
            low, valley = h, c
        elif h > low * 2 and c > valley + 2:
            return valley, True
    return 2, False


def test_cutoff_fragility_rule(oracle):
    """The library-level margin (kc_cutoff_is_fragile, host code of libkc.so, no GPU):
    every count vector whose valley search finds nothing (the oracle's cutoff, restated
    above, gives 2 without a valley) is fragile, and so is any with <= 100 entries or a
    95th percentile under 30; a fixed cutoff (min_coverage > 5) never is; a deep valley
    between large bins survives every resampling and is not; a valley that one Poisson
    step removes is; and the decision is reproducible."""
    from kc.learn import count_histogram, cutoff_is_fragile
    rng = np.random.default_rng(4)
    for t in range(300):
        n = int(rng.integers(101, 3000))
        counts = np.concatenate([rng.geometric(0.5, n) + 1, rng.poisson(int(rng.integers(30, 90)), n)]).astype(np.uint32)
        covs = np.sort(counts)
        top = int(covs[len(covs) * 95 // 100])
        want = oracle.cutoff(counts, 5.0)[0]
        if top < 30:
            assert want == 2 and cutoff_is_fragile(counts)
            continue
        mvs = min(max(top // 4, 10), 35)
        hist = np.bincount(np.minimum(covs.astype(np.int64), mvs + 1), minlength=mvs + 2)[:mvs + 1]
        got, found = _valley(hist)
        assert got == want, t
        if not found:
            assert cutoff_is_fragile(counts)
        assert cutoff_is_fragile(counts) == cutoff_is_fragile(counts)
        assert cutoff_is_fragile(counts, seed=5) == cutoff_is_fragile(counts, seed=5)
    assert not cutoff_is_fragile(np.arange(500), min_coverage=12.0)
    assert cutoff_is_fragile(np.full(50, 40))
    assert cutoff_is_fragile(np.full(5000, 20))
    def of(bins):
        return np.concatenate([np.full(n, c) for c, n in bins.items()])

    # error k-mers at 2..4, a valley of 10 at 7, genomic k-mers around 60: robust
    deep = of({2: 200000, 3: 50000, 4: 20000, 5: 100, 6: 40, 7: 10, 8: 30, 9: 100, 10: 1000, 11: 2000, 12: 3000,
               13: 4000, 14: 5000, 15: 6000, 60: 100000})
    assert _valley(count_histogram(deep)[:16]) == (7, True) and not cutoff_is_fragile(deep)
    # the same shape with single-digit bins: Poisson resamplings move the valley past it
    thin = of({2: 40, 3: 9, 4: 5, 5: 4, 6: 3, 7: 5, 8: 6, 9: 7, 60: 600})
    assert _valley(count_histogram(thin)[:16]) == (6, True) and cutoff_is_fragile(thin)
