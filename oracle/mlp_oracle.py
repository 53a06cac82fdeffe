"""FP64 numpy restatement of the learned scheduler's network (test
infrastructure only): 12 -> 64 -> 64 -> 3, ReLU, standardised inputs, loss
sum(mask * (out - y)^2) / n.  Parameter layout as in subway_amd/csrc/mlp.hpp.
Pinned by central finite differences in the tests."""
import numpy as np

IN, H, OUT = 12, 64, 3
NORM = 2 * IN
W1 = NORM
B1 = W1 + IN * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H * OUT
PARAMS = B3 + OUT


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[:IN], p[IN:NORM], p[W1:B1].reshape(IN, H), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3],
            p[W3:B3].reshape(H, OUT), p[B3:PARAMS])


def forward(p, x):
    mu, sc, w1, b1, w2, b2, w3, b3 = unpack(p)
    xn = (np.asarray(x, np.float64) - mu) * sc
    z1 = xn @ w1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2 + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3, (xn, z1, h1, z2, h2)


def loss(p, x, y, m):
    out, _ = forward(p, x)
    return float(np.sum(m * (out - y) ** 2) / len(x))


def grad(p, x, y, m):
    mu, sc, w1, b1, w2, b2, w3, b3 = unpack(p)
    out, (xn, z1, h1, z2, h2) = forward(p, x)
    n = len(x)
    dout = 2.0 * m * (out - y) / n
    g = np.zeros(PARAMS)
    g[W3:B3] = (h2.T @ dout).ravel()
    g[B3:PARAMS] = dout.sum(0)
    dz2 = (dout @ w3.T) * (z2 > 0)
    g[W2:B2] = (h1.T @ dz2).ravel()
    g[B2:W3] = dz2.sum(0)
    dz1 = (dz2 @ w2.T) * (z1 > 0)
    g[W1:B1] = (xn.T @ dz1).ravel()
    g[B1:W2] = dz1.sum(0)
    return g


def finite_difference(p, x, y, m, idx, eps=1e-6):
    p = np.asarray(p, np.float64)
    out = []
    for i in idx:
        a = p.copy()
        b = p.copy()
        a[i] += eps
        b[i] -= eps
        out.append((loss(a, x, y, m) - loss(b, x, y, m)) / (2 * eps))
    return np.array(out)
