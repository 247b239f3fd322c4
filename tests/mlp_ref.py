"""A float64 numpy restatement of the traversal's cost model (treecorr_amd/csrc/tc_cost.hip):
forward pass, mean-squared-error loss, backward pass, Adam, and the mini-batch index
hash, for checking the GPU kernels.  Its gradient is itself pinned by central finite
differences (tests/test_gpu_cost.py)."""
import numpy as np

NF, H = 11, 64
O_W1 = 0
O_B1 = O_W1 + H * NF
O_W2 = O_B1 + H
O_B2 = O_W2 + H * H
O_W3 = O_B2 + H
O_B3 = O_W3 + H
NP = O_B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[O_W1:O_B1].reshape(H, NF), p[O_B1:O_W2], p[O_W2:O_B2].reshape(H, H),
            p[O_B2:O_W3], p[O_W3:O_B3], p[O_B3])


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.)
    return h2 @ W3 + b3, (X, h1, h2)


def loss_and_grad(p, X, y):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    yhat, (X, h1, h2) = forward(p, X)
    n = len(X)
    r = yhat - np.asarray(y, np.float64)
    loss = float(np.mean(r * r))
    dy = 2. * r / n
    gW3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, W3) * (h2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (h1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** t, 1 - b2 ** t
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def batch_idx(seed, step, n, batch):
    """The mini-batch of training step `step` (k_batch_idx's counter-based hash)."""
    M = (1 << 64) - 1
    out = np.empty(batch, np.int64)
    for s in range(batch):
        z = (seed + 0x9e3779b97f4a7c15 * (step * batch + s + 1)) & M
        z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M
        z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M
        z ^= z >> 31
        out[s] = z % n
    return out
