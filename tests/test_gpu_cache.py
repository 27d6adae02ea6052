"""The radiance cache's GPU kernels (st_cache_*) against the numpy oracle with
the same fp16 rounding points: inference, the gradients of one training step
(weights and hash table, which exercise every MFMA layout: forward,
backward through transposed weights and ReLU masks, and the weight-gradient
product), the Adam update, and that training fits a smooth radiance field.
Sample counts are not multiples of 64, so every kernel's row tail runs.

Tolerances: the GPU accumulates in fp32 in another order, so an fp16
rounding can land one unit away; sums are compared to 2% of each value plus
1% of the largest (with a relative norm of the difference under 1%)."""
import numpy as np
import pytest

from oracle.cache_oracle import CacheOracle

pytestmark = pytest.mark.gpu

CFG = dict(levels=8, log2_table=14, base_resolution=4, max_resolution=64)
BOX = ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))


def _queries(rng, n):
    q = np.zeros((n, 12), np.float32)
    q[:, :3] = rng.uniform(-1, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    q[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    q[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    q[:, 9:12] = rng.uniform(0, 1, (n, 3))
    return q


def _close(got, want, rel=2e-2, frac=1e-2):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    tol = rel * np.abs(want) + frac * np.abs(want).max()
    bad = np.abs(got - want) > tol
    assert bad.mean() < 1e-3, (int(bad.sum()), got[bad][:5], want[bad][:5])
    assert np.linalg.norm(got - want) <= 1e-2 * np.linalg.norm(want) + 1e-12


@pytest.fixture(scope="module")
def cache(gpu_ctx):
    from stratum2_amd.cache import RadianceCache
    c = RadianceCache(gpu_ctx, *BOX, seed=3, **CFG)
    yield c
    c.close()


def test_inference_matches_oracle(cache):
    import torch
    rng = np.random.default_rng(0)
    table, W = cache.params()
    # larger table values than the initial 1e-4, so the hash features matter
    table = rng.normal(scale=0.5, size=table.shape).astype(np.float32)
    cache.set_params(table=table)
    q = _queries(rng, 3001)
    got = cache.infer(torch.from_numpy(q).cuda()).cpu().numpy()
    o = CacheOracle(CFG["levels"], CFG["log2_table"], CFG["base_resolution"], CFG["max_resolution"], *BOX)
    Y = o.forward(W, o.encode(table, q))[-1]
    _close(got, np.maximum(Y[:, :3], 0))


def test_training_step_gradients_and_adam_match_oracle(cache):
    import torch
    rng = np.random.default_rng(1)
    table0 = rng.normal(scale=0.5, size=cache.params()[0].shape).astype(np.float32)
    W0 = cache.params()[1]
    cache.set_params(table=table0)
    q = _queries(rng, 5003)
    t = rng.uniform(0, 2, (5003, 3)).astype(np.float32)
    loss = cache.train(torch.from_numpy(q).cuda(), torch.from_numpy(t).cuda())
    table1, W1, tg, wg = cache.params(grads=True)
    o = CacheOracle(CFG["levels"], CFG["log2_table"], CFG["base_resolution"], CFG["max_resolution"], *BOX)
    acts = o.forward(W0, o.encode(table0, q))
    L, dY, _ = o.loss(acts[-1], t)
    assert abs(loss - L / len(q)) <= 1e-3 * L / len(q)
    dW, dX = o.backward(W0, acts, dY)
    _close(wg, dW)
    dT = o.table_grad(q, dX)
    _close(tg, dT)
    # Adam (first step) from the GPU's own gradients: dense for the weights,
    # sparse for the table
    c = cache.config
    p, _, _ = o.adam(W0.astype(np.float64), np.zeros(W0.shape), np.zeros(W0.shape), wg / len(q), 1, c.learning_rate, c.beta1,
                     c.beta2, c.epsilon, sparse=False)
    assert np.allclose(W1, p, rtol=1e-5, atol=1e-6)
    p, _, _ = o.adam(table0.astype(np.float64), np.zeros(table0.shape), np.zeros(table0.shape), tg / len(q), 1, c.learning_rate,
                     c.beta1, c.beta2, c.epsilon, sparse=True)
    assert np.allclose(table1, p, rtol=1e-5, atol=1e-6)
    assert np.array_equal(table1[tg == 0], table0[tg == 0])


def test_training_fits_a_smooth_field(gpu_ctx):
    import torch
    from stratum2_amd.cache import RadianceCache
    c = RadianceCache(gpu_ctx, *BOX, seed=5, **CFG)
    rng = np.random.default_rng(2)

    def field(q):
        p = q[:, :3]
        return np.stack([1 + np.sin(2 * p[:, 0]), 1 + 0.5 * p[:, 1] * p[:, 2], np.exp(-p[:, 0] ** 2 - p[:, 1] ** 2)], 1)

    losses = []
    for it in range(300):
        q = _queries(rng, 4096)
        t = field(q).astype(np.float32)
        losses.append(c.train(torch.from_numpy(q).cuda(), torch.from_numpy(t).cuda()))
    qe = _queries(rng, 2000)
    pred = c.infer(torch.from_numpy(qe).cuda()).cpu().numpy()
    err = np.abs(pred - field(qe)).mean() / np.abs(field(qe)).mean()
    print("loss", losses[0], "->", np.mean(losses[-10:]), "mean relative error", err)
    assert np.mean(losses[-10:]) < 0.05 * losses[0] and err < 0.1
    c.close()


def test_inference_reads_the_moving_average(gpu_ctx):
    """With ema_decay d, inference after two steps reads
    d (d p0 + (1 - d) p1) + (1 - d) p2, the moving average restarted at the
    parameters set (p0), while training reads the parameters themselves."""
    import torch
    from stratum2_amd.cache import RadianceCache
    d = 0.75
    c = RadianceCache(gpu_ctx, *BOX, seed=6, ema_decay=d, **CFG)
    rng = np.random.default_rng(4)
    table0 = rng.normal(scale=0.5, size=c.params()[0].shape).astype(np.float32)
    c.set_params(table=table0)
    W0 = c.params()[1]
    q = _queries(rng, 3001)
    got0 = c.infer(torch.from_numpy(q).cuda()).cpu().numpy()
    ps = [(table0, W0)]
    for _ in range(2):
        t = rng.uniform(0, 2, (3001, 3)).astype(np.float32)
        c.train(torch.from_numpy(q).cuda(), torch.from_numpy(t).cuda())
        ps.append(c.params())
    (t0, w0), (t1, w1), (t2, w2) = ps
    te = d * (d * t0 + (1 - d) * t1) + (1 - d) * t2
    we = d * (d * w0 + (1 - d) * w1) + (1 - d) * w2
    got = c.infer(torch.from_numpy(q).cuda()).cpu().numpy()
    o = CacheOracle(CFG["levels"], CFG["log2_table"], CFG["base_resolution"], CFG["max_resolution"], *BOX)
    _close(got0, np.maximum(o.forward(W0, o.encode(table0, q))[-1][:, :3], 0))
    _close(got, np.maximum(o.forward(we, o.encode(te, q))[-1][:, :3], 0))
    # not the raw parameters
    raw = np.maximum(o.forward(w2, o.encode(t2, q))[-1][:, :3], 0)
    assert np.abs(got - raw).mean() > 5 * np.abs(got - np.maximum(o.forward(we, o.encode(te, q))[-1][:, :3], 0)).mean()
    c.close()


def test_cache_argument_errors(gpu_ctx):
    from stratum2_amd.cache import RadianceCache
    with pytest.raises(RuntimeError, match="levels"):
        RadianceCache(gpu_ctx, *BOX, levels=17)
    with pytest.raises(RuntimeError, match="bounding box"):
        RadianceCache(gpu_ctx, (0, 0, 0), (1, 0, 1))
