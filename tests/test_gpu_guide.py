"""GPU guiding network (BF16 MFMA, FP32 accumulation) against the FP64 oracle:
forward outputs, loss gradients (mean over the batch), one Adam step, and a
training run that fits a known directional target."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _params(rng, scale=0.25):
    from oracle.oracle import MLP_NB, MLP_NW
    p = (rng.normal(size=MLP_NW + MLP_NB) * scale).astype(np.float32)
    W4 = p[96 * 64 + 2 * 64 * 64:MLP_NW].reshape(64, 48)
    W4[:, 41:] = 0
    p[MLP_NW + 3 * 64 + 41:] = 0
    return p


class Guide:
    def __init__(self, rx, ctx):
        from renderx_amd import _lib
        self.L = _lib.lib()
        self.ctx = ctx
        p = np.zeros(1, rx.GUIDE_PARAMS)
        p["train_passes"], p["batch_size"], p["steps_per_pass"], p["learning_rate"] = 8, 4096, 16, 1e-2
        p["beta1"], p["beta2"], p["eps"], p["seed"] = 0.9, 0.999, 1e-8, 7
        ctx.check(self.L.rx_guide_init(ctx.h, p.ctypes.data_as(C.c_void_p)), "rx_guide_init")
        lo = np.zeros(3, np.float32)
        hi = np.ones(3, np.float32)
        ctx.check(self.L.rx_guide_set_bounds(ctx.h, lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p)),
                  "bounds")

    def set(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.ctx.check(self.L.rx_guide_set_weights(self.ctx.h, p.ctypes.data_as(C.c_void_p)), "set")

    def get(self):
        p = np.empty(17648, np.float32)
        self.ctx.check(self.L.rx_guide_get_weights(self.ctx.h, p.ctypes.data_as(C.c_void_p)), "get")
        return p

    def eval(self, pos, d):
        out = np.empty((len(pos), 48), np.float32)
        P = lambda a: np.ascontiguousarray(a, np.float32).ctypes.data_as(C.c_void_p)
        pos = np.ascontiguousarray(pos, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        self.ctx.check(self.L.rx_guide_eval(self.ctx.h, len(pos), P(pos), P(d), out.ctypes.data_as(C.c_void_p)),
                       "eval")
        return out

    def eval_mixture(self, pos, d):
        out = np.empty((len(pos), 52), np.float32)
        pos = np.ascontiguousarray(pos, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        self.ctx.check(self.L.rx_guide_eval_mixture(self.ctx.h, len(pos), pos.ctypes.data_as(C.c_void_p),
                                                    d.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)),
                       "eval_mixture")
        return out

    def train(self, pos, d, w, pb, wt, want_grad=True):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (pos, d, w, pb, wt)]
        loss = C.c_float()
        g = np.empty(17648, np.float32) if want_grad else None
        self.ctx.check(self.L.rx_guide_train_batch(self.ctx.h, len(pos), *[a.ctypes.data_as(C.c_void_p) for a in arrs],
                                                   C.byref(loss), None if g is None else g.ctypes.data_as(C.c_void_p)),
                       "train")
        return loss.value, g


def _batch(rng, n):
    pos = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    w = rng.normal(size=(n, 3))
    w = (w / np.linalg.norm(w, axis=1, keepdims=True)).astype(np.float32)
    pb = rng.uniform(0.05, 1.0, n).astype(np.float32)
    wt = rng.uniform(0.1, 2.0, n).astype(np.float32)
    return pos, d, w, pb, wt


def _oracle_inputs(oracle_mod, p, pos, d):
    # the GPU computes on bf16 weights and bf16 encoded inputs; feed the oracle the same values
    from oracle.oracle import MLP_NW
    q = p.astype(np.float64).copy()
    q[:MLP_NW] = _bf16(p[:MLP_NW])
    x = _bf16(oracle_mod.mlp_encode(pos, d, [0, 0, 0], [1, 1, 1]).astype(np.float32)).astype(np.float64)
    return q, x


def test_forward_matches_oracle(rx, oracle_mod, gpu_ctx):
    print("[guide] forward parity", flush=True)
    rng = np.random.default_rng(0)
    g = Guide(rx, gpu_ctx)
    p = _params(rng)
    g.set(p)
    pos, d, _, _, _ = _batch(rng, 1000)
    z = g.eval(pos, d)
    q, x = _oracle_inputs(oracle_mod, p, pos, d)
    zr, _ = oracle_mod.mlp_forward(x, q)
    err = np.abs(z - zr)
    tol = 0.03 * (np.abs(zr) + np.sqrt((zr ** 2).mean()))
    print(f"  max |err| {err.max():.4f}, rms(z) {np.sqrt((zr ** 2).mean()):.4f}")
    assert (err <= tol).mean() > 0.999 and err.max() < 0.2 * np.abs(zr).max()
    assert np.abs(z[:, 41:]).max() == 0.0


def _decode(z):
    """The renderer's mixture from raw outputs, in FP64: softmax lobe weights
    as a cumulative table, unit means, kappa = exp(clamp(z, -3, 8)), the vMF
    normaliser times the weight, alpha = 0.05 + 0.45 sigmoid(z40)."""
    z = z.astype(np.float64)
    lo = z[:, 0:40].reshape(-1, 8, 5)
    e = np.exp(lo[:, :, 0] - lo[:, :, 0].max(axis=1, keepdims=True))
    w = e / e.sum(axis=1, keepdims=True)
    cdf = np.cumsum(w, axis=1)
    mu = lo[:, :, 1:4] / np.linalg.norm(lo[:, :, 1:4], axis=2, keepdims=True)
    k = np.exp(np.clip(lo[:, :, 4], -3.0, 8.0))
    cw = w * k / (2 * np.pi * (1 - np.exp(-2 * k)))
    alpha = 0.05 + 0.45 / (1 + np.exp(-z[:, 40]))
    return mu, k, cw, cdf, alpha


def test_decoded_mixture_matches_outputs(rx, gpu_ctx):
    """The decode the renderer samples and evaluates (one 208-byte record per
    query, written by the inference kernel) against an FP64 decode of the same
    network's raw outputs, on a row count that is not a multiple of the
    16-row tile."""
    print("[guide] decoded mixture", flush=True)
    rng = np.random.default_rng(5)
    g = Guide(rx, gpu_ctx)
    p = _params(rng, scale=0.5)
    g.set(p)
    n = 1000
    pos, d, _, _, _ = _batch(rng, n)
    z = g.eval(pos, d)
    m = g.eval_mixture(pos, d)
    mu, k, cw, cdf, alpha = _decode(z)
    lobes = m[:, :32].reshape(n, 8, 4)
    assert np.abs(lobes[:, :, :3] - mu).max() < 2e-6
    assert np.abs(lobes[:, :, 3] / k - 1).max() < 2e-6
    assert np.abs(m[:, 40:48] - cdf).max() < 2e-6
    # cw is the stored table's weights (what the sampler realises) times the vMF normaliser
    wg = np.diff(m[:, 40:48].astype(np.float64), axis=1, prepend=0.0)
    expect = wg * k / (2 * np.pi * (1 - np.exp(-2 * k)))
    big = wg > 1e-6
    assert np.abs(m[:, 32:40][big] / expect[big] - 1).max() < 1e-5
    assert np.abs(m[:, 32:40] - expect).max() < 1e-5 * cw.max()
    assert (m[:, 47] == 1.0).all()  # the sampler's last bin always closes
    assert np.abs(m[:, 48] - alpha).max() < 1e-6
    # an index slip in the decode (weight logits read one output late) fails the table check
    zl = z[:, 1:40:5].astype(np.float64)
    wrong = np.exp(zl - zl.max(axis=1, keepdims=True))
    wrong /= wrong.sum(axis=1, keepdims=True)
    assert np.abs(np.cumsum(wrong, axis=1) - m[:, 40:48]).max() > 1e-2


def test_gradient_matches_oracle(rx, oracle_mod, gpu_ctx):
    print("[guide] gradient parity", flush=True)
    rng = np.random.default_rng(1)
    g = Guide(rx, gpu_ctx)
    p = _params(rng)
    g.set(p)
    n = 4096
    pos, d, w, pb, wt = _batch(rng, n)
    loss, grad = g.train(pos, d, w, pb, wt)
    q, x = _oracle_inputs(oracle_mod, p, pos, d)
    z4, acts = oracle_mod.mlp_forward(x, q)
    tot, dz4, _, _ = oracle_mod.mlp_loss(z4, w, pb, wt)
    gr = oracle_mod.mlp_backward(x, q, acts, dz4) / n
    print(f"  loss gpu {loss:.5f} oracle {tot / n:.5f}")
    assert abs(loss - tot / n) < 0.02 * abs(tot / n) + 1e-4
    # per parameter block: relative error of the gradient vector
    offs = [0, 96 * 64, 96 * 64 + 4096, 96 * 64 + 8192, 96 * 64 + 8192 + 3072, 17648]
    for a, b in zip(offs[:-1], offs[1:]):
        rel = np.linalg.norm(grad[a:b] - gr[a:b]) / (np.linalg.norm(gr[a:b]) + 1e-12)
        cos = np.dot(grad[a:b], gr[a:b]) / (np.linalg.norm(grad[a:b]) * np.linalg.norm(gr[a:b]) + 1e-30)
        print(f"  block {a}:{b} rel err {rel:.4f} cos {cos:.5f}")
        assert rel < 0.05 and cos > 0.998


def test_adam_step_matches_oracle(rx, oracle_mod, gpu_ctx):
    print("[guide] adam step", flush=True)
    rng = np.random.default_rng(2)
    g = Guide(rx, gpu_ctx)
    p = _params(rng)
    g.set(p)
    pos, d, w, pb, wt = _batch(rng, 2048)
    _, grad = g.train(pos, d, w, pb, wt)
    p1 = g.get()
    ref, _, _ = oracle_mod.adam(p.astype(np.float64), np.zeros(len(p)), np.zeros(len(p)), grad.astype(np.float64),
                                1e-2, 0.9, 0.999, 1e-8, 1)
    live = np.abs(grad) > 1e-6
    assert np.allclose(p1[live], ref[live], rtol=0, atol=2e-6)


def _vmf_dirs(rng, mu, kappa, n):
    u1, u2 = rng.uniform(size=n), rng.uniform(size=n)
    w = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
    r = np.sqrt(np.maximum(0, 1 - w * w))
    t1 = np.cross(mu, [1.0, 0, 0]) if abs(mu[0]) < 0.9 else np.cross(mu, [0, 1.0, 0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(mu, t1)
    ph = 2 * np.pi * u2
    return (w[:, None] * mu + r[:, None] * (np.cos(ph)[:, None] * t1 + np.sin(ph)[:, None] * t2)).astype(np.float32)


def test_training_fits_a_directional_target(rx, oracle_mod, gpu_ctx):
    """Samples drawn uniformly (p_sample = 1/4pi) with weight = target density
    * 4pi: minimising the loss pulls alpha*q + (1-alpha)/4pi toward the target."""
    print("[guide] training fit", flush=True)
    rng = np.random.default_rng(3)
    g = Guide(rx, gpu_ctx)
    mu = np.array([0.3, 0.8, -0.52])
    mu /= np.linalg.norm(mu)
    kappa = 20.0
    n = 1 << 16

    def batch():
        pos = rng.uniform(0, 1, (n, 3)).astype(np.float32)
        d = np.tile(np.array([[0, 0, 1]], np.float32), (n, 1))
        w = rng.normal(size=(n, 3))
        w = (w / np.linalg.norm(w, axis=1, keepdims=True)).astype(np.float32)
        tgt = kappa / (2 * np.pi * (1 - np.exp(-2 * kappa))) * np.exp(kappa * (w @ mu - 1))
        pb = np.full(n, 1 / (4 * np.pi), np.float32)
        return pos, d, w, pb, (tgt * 4 * np.pi).astype(np.float32)

    losses = []
    for it in range(150):
        l, _ = g.train(*batch(), want_grad=False)
        losses.append(l)
    print(f"  loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < losses[0] - 0.5
    # the trained mixture concentrates around mu
    z = g.eval(np.full((4, 3), 0.5, np.float32), np.tile(np.array([[0, 0, 1]], np.float32), (4, 1)))
    q_mu = oracle_mod.mlp_loss(z.astype(np.float64), np.tile(mu.astype(np.float32), (4, 1)),
                               np.ones(4, np.float32), np.ones(4, np.float32), want_grad=False)[2]
    q_anti = oracle_mod.mlp_loss(z.astype(np.float64), np.tile(-mu.astype(np.float32), (4, 1)),
                                 np.ones(4, np.float32), np.ones(4, np.float32), want_grad=False)[2]
    print(f"  q(mu) {q_mu.mean():.3f} q(-mu) {q_anti.mean():.4f}")
    assert q_mu.min() > 20 * q_anti.max()
