"""Pins the radiance-cache oracle (oracle/cache_oracle.py), the checker of the
GPU cache kernels:

- its corner index is the direct index x + n (y + n z) where a level's
  corners fit the table and the published spatial hash elsewhere, checked
  against hand-computed values;
- in float64 its backward pass (weight gradients and the hash-table
  gradient) equals central differences of its forward loss with the loss's
  stop-gradient denominators frozen;
- its encoding is the trilinear interpolation of the level's cell corners:
  weights sum to 1 and a table that is linear in the corner coordinates
  encodes to the same linear function of the position.

A transposed weight, a dropped ReLU mask, a wrong loss factor or a table
gradient credited to the wrong corner fails one of these."""
import numpy as np

from oracle.cache_oracle import CacheOracle, cell_hash


def _queries(rng, n):
    q = np.zeros((n, 12), np.float32)
    q[:, :3] = rng.uniform(-1, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    q[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    q[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    q[:, 9:12] = rng.uniform(0, 1, (n, 3))
    return q


def test_corner_index():
    # the spatial hash in uint32 arithmetic, against values worked by hand
    assert cell_hash(1, 0, 0) == 1
    assert cell_hash(0, 1, 0) == 2654435761
    assert cell_hash(0, 0, 1) == 805459861
    assert cell_hash(3, 2, 0) == (3 ^ ((2 * 2654435761) % 2 ** 32))
    assert cell_hash(-1, 0, 0) == 2 ** 32 - 1  # negatives wrap as two's complement
    o = CacheOracle(3, 12, 4, 30, (0, 0, 0), (1, 1, 1))  # res 4, 10, 30: (res + 2)^3 = 216, 1728 fit 4096, 32768 does not
    assert o.dense == [True, True, False]
    assert o.corner_index(1, 2, 3, 0) == 1 + 6 * (2 + 6 * 3)
    assert o.corner_index(5, 0, 11, 1) == 5 + 12 * 12 * 11
    assert o.corner_index(7, 9, 4, 2) == (7 ^ ((9 * 2654435761) % 2 ** 32) ^ ((4 * 805459861) % 2 ** 32)) % 4096
    # the x and x + 1 corners of a hashed level sit in the same 16-entry (64-byte) block when x is even
    for x in range(0, 40, 2):
        assert o.corner_index(x, 9, 4, 2) // 2 == o.corner_index(x + 1, 9, 4, 2) // 2


def test_backward_equals_central_differences_float64():
    rng = np.random.default_rng(1)
    o = CacheOracle(4, 8, 2, 16, (-1, -1, -1), (1, 1, 1))
    q = _queries(rng, 40)
    table = rng.normal(scale=0.5, size=(4 * 256, 2))
    W = rng.normal(scale=0.3, size=(4, 64, 64))
    W[3, 3:] = 0
    t = rng.uniform(0, 2, (40, 3))

    def loss(table, W, den=None):
        acts = o.forward(W, o.encode(table, q, half=False), half=False)
        return o.loss(acts[-1], t, den)

    L0, dY, den = loss(table, W)
    acts = o.forward(W, o.encode(table, q, half=False), half=False)
    dW, dX = o.backward(W, acts, dY, half=False)
    dT = o.table_grad(q, dX)
    eps = 1e-6
    for (l, a, b) in [(0, 5, 7), (1, 10, 3), (2, 63, 0), (3, 1, 40), (0, 0, 2 * 4 + 9)]:
        Wp, Wm = W.copy(), W.copy()
        Wp[l, a, b] += eps
        Wm[l, a, b] -= eps
        fd = (loss(table, Wp, den)[0] - loss(table, Wm, den)[0]) / (2 * eps)
        assert abs(fd - dW[l, a, b]) <= 1e-6 * max(1, abs(fd)), (l, a, b, fd, dW[l, a, b])
    touched = np.argwhere(np.abs(dT) > 1e-3)
    for e, f in touched[rng.choice(len(touched), 8, replace=False)]:
        Tp, Tm = table.copy(), table.copy()
        Tp[e, f] += eps
        Tm[e, f] -= eps
        fd = (loss(Tp, W, den)[0] - loss(Tm, W, den)[0]) / (2 * eps)
        assert abs(fd - dT[e, f]) <= 1e-6 * max(1, abs(fd)), (e, f, fd, dT[e, f])


def test_encoding_is_trilinear():
    rng = np.random.default_rng(2)
    o = CacheOracle(3, 20, 4, 16, (0, 0, 0), (1, 1, 1))  # 2^20 entries: no collisions among these cells
    q = _queries(rng, 200)
    q[:, :3] = rng.uniform(0, 1, (200, 3))
    for l in range(3):
        idx, wt = o.corners(q, l)
        assert np.allclose(wt.sum(1), 1, atol=1e-6)
    # a table holding a linear function of each corner's integer coordinates
    # (queries touching an entry that two cells share are left out)
    table = np.zeros((3 * (1 << 20), 2))
    a = np.array([0.3, -0.2, 0.5])
    shared = np.zeros(3 * (1 << 20), bool)
    for l in range(3):
        r = int(o.res[l])
        g = np.stack(np.meshgrid(*[np.arange(r + 2)] * 3, indexing="ij"), -1).reshape(-1, 3)
        e = l * (1 << 20) + o.corner_index(g[:, 0], g[:, 1], g[:, 2], l)
        u, cnt = np.unique(e, return_counts=True)
        shared[u[cnt > 1]] = True
        table[e, 0] = g @ a
        table[e, 1] = 1.0
    keep = np.ones(len(q), bool)
    for l in range(3):
        keep &= ~shared[o.corners(q, l)[0]].any(1)
    assert keep.sum() > 150
    X = o.encode(table, q[keep], half=False)
    for l in range(3):
        assert np.allclose(X[:, 2 * l], (q[keep, :3] * o.res[l]) @ a, atol=1e-5)
        assert np.allclose(X[:, 2 * l + 1], 1.0, atol=1e-6)
