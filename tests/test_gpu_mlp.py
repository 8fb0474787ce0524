"""GPU: the learned first pass's network against its NumPy restatement (oracle/mlp_np.py):
the MFMA forward pass, the training kernels' gradient (the gradient check), reproducible
training, and a learnable rule actually learned."""
import numpy as np
import pytest

from kc.learn import Mlp
from oracle import mlp_np as M

pytestmark = pytest.mark.gpu


def _data(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1.5, (n, M.N_IN)).astype(np.float32)
    y = (X[:, 0] + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3] > 0.2).astype(np.uint8)
    return X, y


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 40001])
def test_forward_matches_oracle(n):
    X, _ = _data(n, n)
    m = Mlp(seed=7)
    w = M.init_params(3)
    m.params = w
    assert np.array_equal(m.params, w)
    p = m.predict(X)
    q, _ = M.forward(w, X)
    err = np.abs(p - q)
    # fp32 accumulation order differs; a bf16 rounding of h1 can land one step apart
    assert err.max() < 5e-3 and np.mean(err) < 1e-4, (err.max(), np.mean(err))
    m.close()


@pytest.mark.parametrize("pos_weight,weighted", [(1.0, False), (3.0, True)])
def test_gradient_check_against_oracle(pos_weight, weighted):
    X, y = _data(11, 3000)
    sw = np.random.default_rng(2).uniform(0.2, 2.0, 3000).astype(np.float32) if weighted else None
    m = Mlp(seed=5)
    w = M.init_params(4)
    m.params = w
    loss, g = m.gradient(X, y, sw, pos_weight)
    lref, gref = M.loss_and_grad(w, X, y, sw, pos_weight)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    assert rel < 2e-3, rel
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3 + 1)):
        blk = np.linalg.norm(g[a:b] - gref[a:b]) / max(np.linalg.norm(gref[a:b]), 1e-12)
        assert blk < 1e-2, (a, blk)
    m.close()


def test_training_is_reproducible_and_learns():
    X, y = _data(21, 60000)
    Xt, yt = _data(22, 20000)
    runs = []
    for _ in range(2):
        m = Mlp(seed=9)
        before = np.mean((m.predict(Xt) > 0.5) == yt)
        m.train(X, y, epochs=10, batch=2048, lr=3e-3, seed=4)
        runs.append(m.params)
        acc = np.mean((m.predict(Xt) > 0.5) == yt)
        m.close()
    assert np.array_equal(runs[0], runs[1])
    assert acc > 0.95 and acc > before + 0.1, (before, acc)
