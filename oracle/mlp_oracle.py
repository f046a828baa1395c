"""Double-precision CPU restatement of the bucket-width MLP (8 -> 32 -> 32 -> 2,
ReLU, masked squared loss) for checking the GPU forward and backward kernels.

TEST INFRASTRUCTURE ONLY (used by tests/ as the checker).  Pinned
independently of the GPU code by central finite differences of the loss
(tests/test_gpu_mlp.py checks the analytic gradient below against them).
"""
import numpy as np

NIN, NH, NOUT = 8, 32, 2


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + NH * NIN].reshape(NH, NIN); o += NH * NIN
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    W3 = p[o:o + NOUT * NH].reshape(NOUT, NH); o += NOUT * NH
    b3 = p[o:o + NOUT]; o += NOUT
    mu = p[o:o + NIN]; o += NIN
    sd = p[o:o + NIN]
    return W1, b1, W2, b2, W3, b3, mu, sd


def forward(p, x):
    W1, b1, W2, b2, W3, b3, mu, sd = unpack(p)
    xn = (np.asarray(x, np.float64) - mu) / sd
    a1 = xn @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    return h2 @ W3.T + b3


def loss(p, x, t, m):
    y = forward(p, x)
    return float(np.sum(0.5 * m * (y - t) ** 2) / x.shape[0])


def gradient(p, x, t, m):
    """Mean loss and gradient w.r.t. the 1,410 trained parameters."""
    W1, b1, W2, b2, W3, b3, mu, sd = unpack(p)
    n = x.shape[0]
    xn = (np.asarray(x, np.float64) - mu) / sd
    a1 = xn @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ W3.T + b3
    dy = m * (y - t)
    dW3 = dy.T @ h2
    db3 = dy.sum(0)
    d2 = (dy @ W3) * (a2 > 0)
    dW2 = d2.T @ h1
    db2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    dW1 = d1.T @ xn
    db1 = d1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3]) / n
    return float(np.sum(0.5 * m * (y - t) ** 2) / n), g
