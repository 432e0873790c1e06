"""FP64 numpy restatement of the surrogate MLP 3 -> 64 -> 64 -> 1 (tanh) and its
gradient (test infrastructure only). Parameter layout: W1[64][3], b1[64],
W2[64][64], b2[64], W3[64], b3 (4481 values)."""
import numpy as np

H = 64
NP = 3 * H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + 3 * H].reshape(H, 3); o += 3 * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3, h1, h2


def loss_grad(p, X, t):
    """loss = (1/n) sum (y - t)^2 and its gradient, standardised units."""
    n = X.shape[0]
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, h1, h2 = forward(p, X)
    e = y - t
    dy = 2.0 * e / n
    d2 = dy[:, None] * W3[None, :] * (1 - h2 ** 2)
    d1 = (d2 @ W2) * (1 - h1 ** 2)
    g = np.concatenate([(d1.T @ X).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0), h2.T @ dy, [dy.sum()]])
    return float(np.mean(e ** 2)), g


def bf16_to_float(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)
