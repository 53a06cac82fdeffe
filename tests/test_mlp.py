"""The learned scheduler's network: the FP64 numpy oracle is pinned by
finite differences (CPU); the HIP forward, gradient and training kernels are
checked against it (GPU)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402


def sample_problem(seed, n=200):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, M.IN))
    y = np.stack([np.tanh(x[:, 0] + 0.5 * x[:, 1]), x[:, 2] ** 2 * 0.1, 0.3 * x[:, 3] - x[:, 4]], 1)
    m = (rng.random((n, M.OUT)) < 0.8).astype(np.float64)
    p = np.zeros(M.PARAMS)
    p[:M.IN] = rng.normal(0, 0.1, M.IN)
    p[M.IN:M.NORM] = rng.uniform(0.5, 2.0, M.IN)
    p[M.NORM:] = rng.normal(0, 0.2, M.PARAMS - M.NORM)
    return p, x, y, m


@pytest.mark.parametrize("seed", range(3))
def test_oracle_gradient_matches_finite_differences(seed):
    p, x, y, m = sample_problem(seed, n=40)
    g = M.grad(p, x, y, m)
    rng = np.random.default_rng(seed)
    idx = np.concatenate([rng.choice(np.arange(M.NORM, M.PARAMS), 60, replace=False),
                          np.arange(M.B3, M.PARAMS)])
    fd = M.finite_difference(p, x, y, m, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-8)
    assert np.all(g[:M.NORM] == 0)


def test_oracle_planted_mistake_caught():
    """Dropping the ReLU derivative in a copy of the backward disagrees with
    the finite differences."""
    p, x, y, m = sample_problem(7, n=30)
    mu, sc, w1, b1, w2, b2, w3, b3 = M.unpack(p)
    out, (xn, z1, h1, z2, h2) = M.forward(p, x)
    dout = 2.0 * m * (out - y) / len(x)
    dz2_bad = dout @ w3.T  # missing * (z2 > 0)
    g_bad = (h1.T @ dz2_bad).ravel()
    fd = M.finite_difference(p, x, y, m, range(M.W2, M.W2 + 200))
    assert np.max(np.abs(g_bad[:200] - fd)) > 1e-4


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(sw, gpu):
    from subway_amd import policy
    p, x, y, m = sample_problem(1, n=500)
    out = policy.forward(p.astype(np.float32), x.astype(np.float32))
    exp, _ = M.forward(p.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64))
    np.testing.assert_allclose(out, exp, rtol=1e-4, atol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 33, 1000])
def test_gpu_gradient_matches_oracle(sw, gpu, n):
    from subway_amd import policy
    p, x, y, m = sample_problem(2, n=n)
    p32, x32, y32, m32 = (a.astype(np.float32) for a in (p, x, y, m))
    g, loss = policy.grad(p32, x32, y32, m32)
    eg = M.grad(*(a.astype(np.float64) for a in (p32, x32, y32, m32)))
    el = M.loss(*(a.astype(np.float64) for a in (p32, x32, y32, m32)))
    assert abs(loss - el) <= 1e-4 * abs(el) + 1e-6
    scale = np.max(np.abs(eg))
    np.testing.assert_allclose(g, eg, rtol=1e-3, atol=1e-4 * scale)


@pytest.mark.gpu
def test_gpu_training_fits(sw, gpu):
    from subway_amd import policy
    _, x, y, m = sample_problem(3, n=2000)
    p0 = policy.init_params(0, x.mean(0), x.std(0))
    p1, losses = policy.train(p0, x, y, np.ones_like(m), steps=800, lr=3e-3)
    assert losses[-1] < 0.2 * losses[0]
    # the trained network predicts held-out samples from the same function
    _, xt, yt, _ = sample_problem(4, n=500)
    pred = policy.forward(p1, xt)
    assert np.mean((pred - yt) ** 2) < 0.5 * np.mean((yt - yt.mean(0)) ** 2)
