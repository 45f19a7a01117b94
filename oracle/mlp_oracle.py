"""CPU restatement (numpy, FP64) of the list-length predictor's network and
its training step, as the library's kernels compute them
(galactic-tides_amd/csrc/gt_predict.hip: mlp_forward, mlp_gradients,
adam_step).  The reference has no such network; the restatement is pinned by
the mathematics instead: its analytic gradient is checked against central
finite differences of its own loss (tests/test_mlp_oracle.py).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the library.
"""
import numpy as np

F, H, O = 8, 64, 2
C = F + O  # columns of a sample: raw features, log(1 + cells), log(1 + bodies)
# parameter layout: W1 (H x F, row-major), b1, W2 (H x H), b2, W3 (O x H), b3
OFF_W1, OFF_B1 = 0, H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + O * H
NPARAMS = OFF_B3 + O
BLOCK = 64  # samples per workgroup
PICK = 2654435761  # the minibatch hash multiplier


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, F), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(O, H), p[OFF_B3:OFF_B3 + O])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def forward(p, x):
    """Standardised outputs (n x O) at normalised inputs x (n x F); also the
    hidden activations."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3, h1, h2


def minibatch(n, blocks, iteration):
    """Sample indices (blocks x BLOCK) of one training iteration: sample
    (iteration * batch + b * BLOCK + t) * PICK mod n."""
    batch = blocks * BLOCK
    k = (np.uint64(iteration) * np.uint64(batch)
         + np.arange(blocks, dtype=np.uint64)[:, None] * np.uint64(BLOCK)
         + np.arange(BLOCK, dtype=np.uint64)[None, :])
    return ((k * np.uint64(PICK)) % np.uint64(n)).astype(np.int64)


def standardise(samples, stats):
    samples = np.asarray(samples, np.float64)
    stats = np.asarray(stats, np.float64)
    mean, dev = stats[:C], stats[C:2 * C]
    return (samples[:, :F] - mean[:F]) / dev[:F], (samples[:, F:] - mean[F:]) / dev[F:]


def loss(p, x, t, batch):
    """Sum over the given samples of their squared errors / (batch * O): the
    share of the minibatch's mean squared error these samples carry."""
    y = forward(p, x)[0]
    return float(np.sum((y - t) ** 2) / (batch * O))


def gradient(p, x, t, batch):
    """Analytic gradient of loss(p, x, t, batch) by backpropagation."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, h1, h2 = forward(p, x)
    d3 = 2.0 * (y - t) / (batch * O)          # d loss / d y
    d2 = (d3 @ W3) * (h2 > 0)                 # through layer 3 and its ReLU
    d1 = (d2 @ W2) * (h1 > 0)                 # through layer 2 and its ReLU
    return pack(d1.T @ x, d1.sum(0), d2.T @ h1, d2.sum(0), d3.T @ h2, d3.sum(0))


def finite_difference(p, x, t, batch, index, h=1e-6):
    """Central difference of loss along parameter `index`."""
    q = np.array(p, np.float64)
    q[index] += h
    up = loss(q, x, t, batch)
    q[index] -= 2 * h
    return (up - loss(q, x, t, batch)) / (2 * h)


def adam(p, m, v, g, iteration, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step with bias correction for step iteration + 1."""
    m = b1 * np.asarray(m, np.float64) + (1 - b1) * g
    v = b2 * np.asarray(v, np.float64) + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** (iteration + 1), 1 - b2 ** (iteration + 1)
    return np.asarray(p, np.float64) - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
