"""The CPU restatement of the list-length predictor's network
(oracle/mlp_oracle.py) pinned by the mathematics: its analytic gradient equals
central finite differences of its own loss, and its forward pass and Adam step
reduce to closed forms in special cases."""
import numpy as np

import mlp_oracle as M


def data(seed, n=300):
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.normal(0, 1 / np.sqrt(M.F), M.H * M.F), rng.normal(0, 0.1, M.H),
                        rng.normal(0, 1 / np.sqrt(M.H), M.H * M.H), rng.normal(0, 0.1, M.H),
                        rng.normal(0, 1 / np.sqrt(M.H), M.O * M.H), rng.normal(0, 0.1, M.O)])
    return p, rng.normal(size=(n, M.F)), rng.normal(size=(n, M.O))


def test_layout_size():
    assert M.NPARAMS == 4866
    p, _, _ = data(0)
    assert len(p) == M.NPARAMS and np.array_equal(M.pack(*M.unpack(p)), p)


def test_gradient_equals_central_differences():
    p, x, t = data(1)
    g = M.gradient(p, x, t, batch=len(x))
    rng = np.random.default_rng(2)
    # every bias and a spread of weights of each layer
    picks = np.concatenate([np.arange(M.OFF_B1, M.OFF_W2), np.arange(M.OFF_B2, M.OFF_W3),
                            np.arange(M.OFF_B3, M.NPARAMS),
                            rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 40, replace=False),
                            rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 80, replace=False),
                            rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 40, replace=False)])
    fd = np.array([M.finite_difference(p, x, t, len(x), i) for i in picks])
    scale = np.abs(g).max()
    assert np.all(np.abs(g[picks] - fd) <= 1e-6 * scale + 1e-5 * np.abs(fd))


def test_gradient_of_a_wrong_network_fails_the_check():
    # a plausible slip (ReLU derivative dropped in the first layer) must show
    p, x, t = data(3)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    y, h1, h2 = M.forward(p, x)
    d3 = 2.0 * (y - t) / (len(x) * M.O)
    d2 = (d3 @ W3) * (h2 > 0)
    d1_wrong = d2 @ W2
    wrong = d1_wrong.sum(0)
    fd = np.array([M.finite_difference(p, x, t, len(x), M.OFF_B1 + i) for i in range(M.H)])
    assert np.abs(wrong - fd).max() > 1e-3 * np.abs(fd).max()


def test_forward_closed_forms():
    p, x, _ = data(4, 50)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    # no hidden weights: constant output b3 + W3 relu(b2)
    q = M.pack(np.zeros_like(W1), np.zeros_like(b1), np.zeros_like(W2), b2, W3, b3)
    y = M.forward(q, x)[0]
    assert np.allclose(y, b3 + W3 @ np.maximum(b2, 0.0), rtol=0, atol=1e-14)
    # all first-layer units dead: output independent of x
    q = M.pack(W1, np.full(M.H, -1e6), W2, b2, W3, b3)
    y = M.forward(q, x)[0]
    assert np.allclose(y, y[0], rtol=0, atol=1e-12)


def test_adam_first_step_moves_by_the_learning_rate():
    g = np.array([3.0, -0.5, 1e-3])
    p, m, v = M.adam(np.zeros(3), np.zeros(3), np.zeros(3), g, 0, 0.01)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-4)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_minibatch_indices():
    idx = M.minibatch(1000, 3, 7)
    assert idx.shape == (3, 64)
    k = 7 * 192 + 64 * 2 + 5
    assert idx[2, 5] == (k * 2654435761) % 1000
