"""Pins of the FP64 guiding-network oracle: central finite differences of the
loss through the whole network, the closed-form first Adam step, and the
normalisation of the vMF mixture density."""
import numpy as np
import pytest


def _params(rng, scale=0.3):
    from oracle.oracle import MLP_NB, MLP_NW
    p = rng.normal(size=MLP_NW + MLP_NB) * scale
    # head padding columns 41..47 are unused
    W4 = p[96 * 64 + 2 * 64 * 64:MLP_NW].reshape(64, 48)
    W4[:, 41:] = 0
    p[MLP_NW + 3 * 64 + 41:] = 0
    return p


def _batch(rng, n):
    pos = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    w = rng.normal(size=(n, 3))
    w = (w / np.linalg.norm(w, axis=1, keepdims=True)).astype(np.float32)
    pb = rng.uniform(0.05, 1.0, n).astype(np.float32)
    wt = rng.uniform(0.1, 2.0, n).astype(np.float32)
    return pos, d, w, pb, wt


def test_mlp_gradient_finite_differences(oracle_mod):
    rng = np.random.default_rng(0)
    p = _params(rng)
    pos, d, w, pb, wt = _batch(rng, 6)
    x = oracle_mod.mlp_encode(pos, d, [0, 0, 0], [1, 1, 1])

    def loss(pp):
        z4, _ = oracle_mod.mlp_forward(x, pp)
        return oracle_mod.mlp_loss(z4, w, pb, wt, want_grad=False)[0]

    z4, acts = oracle_mod.mlp_forward(x, p)
    _, dz4, _, _ = oracle_mod.mlp_loss(z4, w, pb, wt)
    g = oracle_mod.mlp_backward(x, p, acts, dz4)
    idx = np.concatenate([rng.choice(len(p) - 300, 60, replace=False), len(p) - 240 + rng.choice(240, 20, replace=False)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (loss(p + e) - loss(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_adam_first_step_closed_form(oracle_mod):
    rng = np.random.default_rng(1)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    lr, eps = 1e-2, 1e-8
    p1, m1, v1 = oracle_mod.adam(p, np.zeros(50), np.zeros(50), g, lr, 0.9, 0.999, eps, 1)
    # bias-corrected first step: m_hat = g, v_hat = g^2
    assert np.allclose(p1, p - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-12)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)


def test_vmf_mixture_normalised(oracle_mod):
    """The mixture density integrates to one over the sphere (Monte Carlo)."""
    rng = np.random.default_rng(2)
    p = _params(rng, 0.5)
    x = oracle_mod.mlp_encode(rng.uniform(0, 1, (3, 3)).astype(np.float32),
                              np.tile(np.array([[0, 0, 1]], np.float32), (3, 1)), [0, 0, 0], [1, 1, 1])
    z4, _ = oracle_mod.mlp_forward(x, p)
    n = 400000
    w = rng.normal(size=(n, 3))
    w = (w / np.linalg.norm(w, axis=1, keepdims=True)).astype(np.float32)
    for r in range(3):
        _, _, q, _ = oracle_mod.mlp_loss(np.repeat(z4[r:r + 1], n, 0), w, np.ones(n, np.float32),
                                         np.ones(n, np.float32), want_grad=False)
        est = 4 * np.pi * q.mean()
        assert abs(est - 1.0) < 5 * 4 * np.pi * q.std() / np.sqrt(n)
