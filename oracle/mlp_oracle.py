"""CPU reference of the learned part's per-pair MLP (TEST INFRASTRUCTURE ONLY).

Checks the GPU forward/backward kernels: 3 hidden SiLU layers, linear output,
weighted mean-squared loss  L = sum_o w_o sum_r (Y_ro - T_ro)^2 / rows.
The gradients are pinned in tests/test_oracle.py against central finite
differences.  Parameters are one flat vector in the order W1 b1 W2 b2 W3 b3 W4 b4.
"""
import numpy as np


def _unpack(params, shapes):
    out, off = [], 0
    for s in shapes:
        k = int(np.prod(s))
        out.append(params[off:off + k].reshape(s))
        off += k
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, shapes, X, keep=False):
    W1, b1, W2, b2, W3, b3, W4, b4 = _unpack(params, shapes)
    acts, zs = [X], []
    H = X
    for W, b, act in ((W1, b1, 1), (W2, b2, 1), (W3, b3, 1), (W4, b4, 0)):
        Z = H @ W + b
        zs.append(Z)
        H = _silu(Z) if act else Z
        acts.append(H)
    return (H, acts, zs) if keep else H


def loss_and_grad(params, shapes, X, T, w):
    Y, acts, zs = forward(params, shapes, X, keep=True)
    rows = X.shape[0]
    d = Y - T
    loss = float(np.sum(w[None, :] * d * d) / rows)
    G = 2.0 * w[None, :] * d / rows
    Ws = _unpack(params, shapes)[0::2]
    grads = [None] * 8
    for layer in (3, 2, 1, 0):
        grads[2 * layer] = acts[layer].T @ G
        grads[2 * layer + 1] = G.sum(0)
        if layer > 0:
            G = (G @ Ws[layer].T) * _dsilu(zs[layer - 1])
    return loss, np.concatenate([g.ravel() for g in grads])
