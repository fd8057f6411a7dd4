"""The proposal MLP's hand-written HIP kernels against the float64 checker, the GPU
parameter sampler against the host one, and the gain of the trained proposals.

Bars: loss and gradients of one batch within 5e-3 (relative L2 per tensor) of the
checker evaluated with the same bf16 operand rounding (only fp32-vs-fp64 accumulation
differs), within 5e-2 of exact float64; forward outputs within 2e-3; sampler features
bit-identical; one Adam step within 1e-5 of the checker's update; the trained proposals'
survival rate at least 1.5x that of uniform sampling.
"""
import numpy as np
import pytest

from lenia_amd import mlp as M
from lenia_amd import proposal as P
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def params_with_bias(seed):
    rng = np.random.default_rng(seed)
    w = MO.unpack(M.init_params(seed).astype(np.float64))
    for l in (1, 2, 3):
        w[f"b{l}"][:] = rng.normal(0, 0.1, w[f"b{l}"].shape)
    w["b4"][:3] = rng.normal(0, 0.1, 3)
    return MO.pack(w).astype(np.float32)


def batch(seed, n):
    rng = np.random.default_rng(seed)
    x = P.features(P.sample(seed, np.arange(n)))
    y = np.stack([rng.integers(0, 2, n), rng.uniform(0, 1, n), rng.uniform(0, 0.5, n)], axis=1).astype(np.float32)
    return x, y


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gradients_match_checker():
    p = params_with_bias(3)
    x, y = batch(4, 200)
    net = M.MLP(max_batch=256)
    net.set_params(p)
    loss, g = net.gradients(x, y)
    net.close()
    lref, gref = MO.loss_and_grad(p.astype(np.float64), x.astype(np.float64), y.astype(np.float64), bf16=True)
    l64, g64 = MO.loss_and_grad(p.astype(np.float64), x.astype(np.float64), y.astype(np.float64), bf16=False)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in MO.SHAPES])
    for t, (name, _) in enumerate(MO.SHAPES):
        a, b, c = g[offs[t]:offs[t + 1]], gref[offs[t]:offs[t + 1]], g64[offs[t]:offs[t + 1]]
        assert rel(a, b) <= 5e-3, (name, rel(a, b))
        assert rel(a, c) <= 5e-2, (name, rel(a, c))


def test_forward_matches_checker():
    p = params_with_bias(5)
    x, _ = batch(6, 1000)
    net = M.MLP(max_batch=512)
    net.set_params(p)
    out = net.forward(x)
    net.close()
    ref, _ = MO.forward(p.astype(np.float64), x.astype(np.float64), bf16=True)
    assert np.abs(out - ref).max() <= 2e-3


def test_gpu_sampler_matches_host():
    net = M.MLP(max_batch=4096)
    got = net.sample_features(11, 1000, 4096)
    net.close()
    ref = P.features(P.sample(11, np.arange(1000, 5096)))
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_score_sampled_equals_forward():
    p = params_with_bias(7)
    net = M.MLP(max_batch=2048)
    net.set_params(p)
    s = net.score_sampled(12, 5000)
    out = net.forward(P.features(P.sample(12, np.arange(5000))))
    net.close()
    assert np.array_equal(s.view(np.uint32), out[:, 0].copy().view(np.uint32))


def test_adam_step_matches_checker():
    p = params_with_bias(8)
    x, y = batch(9, 256)
    net = M.MLP(max_batch=256)
    net.set_params(p)
    net.set_data(x, y)
    net.train(1, 256, 1e-3, seed=0)  # one full-batch step (the order is a permutation)
    p1 = net.params()
    net.close()
    _, g = MO.loss_and_grad(p.astype(np.float64), x.astype(np.float64), y.astype(np.float64), bf16=True)
    m = 0.1 * g
    v = 0.001 * g * g
    ref = p - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    # where |g| is tiny Adam's step depends on its sign only; compare the rest
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    assert np.abs(p1[big] - ref[big]).max() <= 1e-5
    assert np.array_equal(p1[g == 0], p[g == 0])


def test_training_learns_a_rule():
    """A synthetic rule of the features is learnt to > 90 % accuracy."""
    x = P.features(P.sample(13, np.arange(16384)))
    surv = (x[:, 54] + x[:, 57] > 1.0).astype(np.float32)
    y = np.stack([surv, 0.5 * x[:, 55], 0.25 * x[:, 58]], axis=1).astype(np.float32)
    net = M.MLP(max_batch=256)
    net.set_params(M.init_params(1))
    net.set_data(x[:12288], y[:12288])
    hist = net.train(15, 256, 2e-3, seed=2)
    out = net.forward(x[12288:])
    net.close()
    assert hist[-1] < 0.5 * hist[0], hist
    acc = np.mean((out[:, 0] > 0) == (y[12288:, 0] > 0.5))
    assert acc > 0.9, acc


def test_proposals_beat_uniform_sampling():
    """Rollouts of the batched engine -> training -> scoring fresh sets on the GPU -> the
    top sets re-simulated by the plain engine survive more often than uniform draws."""
    steps = 100
    data = P.rollout(21, np.arange(8192), steps)
    net = M.MLP(max_batch=4096)
    net.set_params(M.init_params(4))
    net.set_data(data["features"], P.targets(data))
    net.train(30, 256, 1e-3, seed=3)
    logits = net.score_sampled(22, 65536)
    net.close()
    top = np.argsort(-logits)[:512]
    hit = P.rollout(22, top, steps)["survival"].mean()
    base = P.rollout(22, np.arange(512), steps)["survival"].mean()
    print(f"base rate {base:.3f}, proposal hit rate {hit:.3f}")
    assert hit >= 1.5 * base and hit >= 0.1, (hit, base)
