"""FP64 restatement of the learned projection used to prune join candidates.

TEST INFRASTRUCTURE ONLY (it checks the library's MLP kernels).  The
reference has no learned component; this is the closed-form mathematics of
the model and loss the library trains:
    f(x) = W2 relu(W1 x + b1) + b2
    for pairs (a_i, b_i) with target t_i and scale s_i:
    p_i = |f(x_a) - f(x_b)|^2,  r_i = (p_i - t_i) / (t_i + s_i),  L = mean r_i^2
Gradients are derived by hand below and pinned by central finite
differences in the tests.
"""
import numpy as np


def forward(W1, b1, W2, b2, X):
    pre = X @ W1.T + b1
    h = np.maximum(pre, 0.0)
    z = h @ W2.T + b2
    return pre, h, z


def loss_grad(W1, b1, W2, b2, xa, xb, t, s):
    W1, b1, W2, b2, xa, xb, t, s = [np.asarray(a, np.float64) for a in (W1, b1, W2, b2, xa, xb, t, s)]
    B = t.size
    X = np.empty((2 * B, xa.shape[1]))
    X[0::2] = xa
    X[1::2] = xb
    pre, h, z = forward(W1, b1, W2, b2, X)
    e = z[0::2] - z[1::2]
    p = (e * e).sum(1)
    den = t + s
    r = (p - t) / den
    L = float(np.mean(r * r))
    dLdp = 2.0 * r / den / B
    dz = np.empty_like(z)
    dz[0::2] = 2.0 * dLdp[:, None] * e
    dz[1::2] = -dz[0::2]
    gW2 = dz.T @ h
    gb2 = dz.sum(0)
    dh = dz @ W2
    dh[pre <= 0] = 0.0
    gW1 = dh.T @ X
    gb1 = dh.sum(0)
    return L, gW1, gb1, gW2, gb2


def loss(W1, b1, W2, b2, xa, xb, t, s):
    return loss_grad(W1, b1, W2, b2, xa, xb, t, s)[0]
