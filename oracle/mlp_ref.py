"""CPU restatement of the cost-model MLP (TEST INFRASTRUCTURE).

float64 numpy forward and backward of the network libhsv trains on the GPU
(hisvsim_amd/csrc/mlp.hip): in -> H -> H -> 1, ReLU, mean squared error.
Parameter layout W1[in][H] | b1[H] | W2[H][H] | b2[H] | W3[H] | b3.
Pinned by central finite differences (tests/test_costmodel.py).
"""
import numpy as np


def unpack(p: np.ndarray, n_in: int, h: int):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + n_in * h].reshape(n_in, h); o += n_in * h
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    W3 = p[o:o + h]; o += h
    b3 = p[o]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, X, n_in, h):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, h)
    X = np.asarray(X, np.float64)
    a1 = np.maximum(X @ W1 + b1, 0.0)
    a2 = np.maximum(a1 @ W2 + b2, 0.0)
    return a2 @ W3 + b3, (X, a1, a2)


def loss_and_grad(p, X, y, n_in, h):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, h)
    out, (X, a1, a2) = forward(p, X, n_in, h)
    y = np.asarray(y, np.float64)
    N = len(X)
    e = out - y
    loss = float(np.mean(e * e))
    d3 = 2.0 * e / N
    gW3 = a2.T @ d3
    gb3 = d3.sum()
    d2 = np.outer(d3, W3) * (a2 > 0)
    gW2 = a1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (a1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g
