"""CPU ORACLE (test infrastructure only) of the radiance cache
(stratum2_amd/csrc/st_cache.h): the hash-grid encoding, the 64-wide ReLU MLP,
the relative L2 loss, the backward pass and Adam, in numpy.

With half=True it rounds to fp16 where the GPU stores fp16 (the table, the
encoded features, the weights, every activation and every gradient between
layers) and accumulates in float64, so it differs from the GPU only by the
order of fp32 accumulation.  With half=False everything is float64, which the
tests use to pin the backward pass against central differences.

A level whose (res + 2)^3 corner points fit its table indexes corner
(x, y, z) directly as x + n (y + n z), n = res + 2; a finer level hashes it
with the spatial hash of Mueller et al., "Instant Neural Graphics
Primitives", 2022: (x xor 2654435761 y xor 805459861 z) mod table size, in
uint32 arithmetic."""
import math

import numpy as np

WIDTH, LAYERS = 64, 4
M32 = np.uint64(0xFFFFFFFF)


def pcg(v):
    v = np.asarray(v, np.uint64) & M32
    s = (v * np.uint64(747796405) + np.uint64(2891336453)) & M32
    w = (((s >> ((s >> np.uint64(28)) + np.uint64(4))) ^ s) * np.uint64(277803737)) & M32
    return ((w >> np.uint64(22)) ^ w) & M32


def cell_hash(x, y, z):
    """The spatial hash of a corner (uint32 arithmetic; negatives wrap as two's complement)."""
    u = lambda a: np.asarray(a, np.int64).astype(np.uint64) & M32  # noqa: E731
    return (u(x) ^ ((u(y) * np.uint64(2654435761)) & M32) ^ ((u(z) * np.uint64(805459861)) & M32)) & M32


class CacheOracle:
    def __init__(self, levels, log2_table, base_resolution, max_resolution, bbox_min, bbox_max):
        self.levels, self.entries = levels, 1 << log2_table
        self.lo = np.asarray(bbox_min, np.float32)
        self.inv_ext = (np.float32(1) / (np.asarray(bbox_max, np.float32) - self.lo)).astype(np.float32)
        b = math.exp((math.log(max_resolution) - math.log(base_resolution)) / (levels - 1)) if levels > 1 else 1.0
        self.res = np.array([math.floor(base_resolution * b ** l) for l in range(levels)], np.float32)
        self.dense = [(float(r) + 2) ** 3 <= self.entries for r in self.res]

    def corner_index(self, x, y, z, l):
        """Table entry (within level l) of corner (x, y, z)."""
        if self.dense[l]:
            n = np.int64(self.res[l]) + 2
            return (np.asarray(x, np.int64) + n * (np.asarray(y, np.int64) + n * np.asarray(z, np.int64))).astype(np.int64)
        return (cell_hash(x, y, z) & np.uint64(self.entries - 1)).astype(np.int64)

    def corners(self, q, l):
        """Table indices (n, 8) and trilinear weights (n, 8) of level l, in float32 as the GPU."""
        u = np.clip((q[:, :3].astype(np.float32) - self.lo) * self.inv_ext, 0, 1).astype(np.float32)
        p = (u * self.res[l]).astype(np.float32)
        c = np.floor(p)
        f = (p - c).astype(np.float32)
        c = c.astype(np.int64)
        idx = np.zeros((len(q), 8), np.int64)
        wt = np.zeros((len(q), 8), np.float32)
        one = np.float32(1)
        for k in range(8):
            dx, dy, dz = k & 1, (k >> 1) & 1, k >> 2
            idx[:, k] = l * self.entries + self.corner_index(c[:, 0] + dx, c[:, 1] + dy, c[:, 2] + dz, l)
            wx = f[:, 0] if dx else one - f[:, 0]
            wy = f[:, 1] if dy else one - f[:, 1]
            wz = f[:, 2] if dz else one - f[:, 2]
            wt[:, k] = ((wx * wy).astype(np.float32) * wz).astype(np.float32)
        return idx, wt

    def encode(self, table, q, half=True):
        """X (n, 64): table is (levels * entries, 2)."""
        t = table.astype(np.float16).astype(np.float64) if half else table.astype(np.float64)
        n = len(q)
        X = np.zeros((n, WIDTH))
        for l in range(self.levels):
            idx, wt = self.corners(q, l)
            X[:, 2 * l:2 * l + 2] = (wt[:, :, None].astype(np.float64) * t[idx]).sum(1)
        o = 2 * self.levels
        X[:, o:o + 3] = (q[:, 3:6] + 1) * 0.5
        X[:, o + 3:o + 6] = (q[:, 6:9] + 1) * 0.5
        X[:, o + 6:o + 9] = q[:, 9:12]
        X[:, o + 9] = 1.0
        return self._r(X, half)

    @staticmethod
    def _r(a, half):
        return a.astype(np.float16).astype(np.float64) if half else a

    def forward(self, weights, X, half=True):
        """Activations [X, H1, H2, H3, Y]; weights (4, 64, 64) [layer][out][in]."""
        W = self._r(weights.astype(np.float64), half)
        acts = [X]
        for l in range(LAYERS):
            a = acts[-1] @ W[l].T
            if l < LAYERS - 1:
                a = np.maximum(a, 0)
            acts.append(self._r(a, half))
        return acts

    @staticmethod
    def loss(Y, targets, den=None):
        """Summed relative L2 loss and the per-sample gradient dY (n, 64);
        den (the stop-gradient denominators) defaults to Y's own."""
        p = Y[:, :3]
        den = p * p + 0.01 if den is None else den
        L = ((p - targets) ** 2 / den / 3).sum()
        dY = np.zeros_like(Y)
        dY[:, :3] = np.clip(2 * (p - targets) / den / 3, -30000, 30000)
        return L, dY, den

    def backward(self, weights, acts, dY, half=True):
        """Weight gradients (4, 64, 64) and dL/dX (n, 64)."""
        W = self._r(weights.astype(np.float64), half)
        d = self._r(dY, half)
        dW = np.zeros((LAYERS, WIDTH, WIDTH))
        for l in range(LAYERS - 1, -1, -1):
            dW[l] = d.T @ acts[l]
            g = d @ W[l]
            if l > 0:
                g = g * (acts[l] > 0)
            d = self._r(g, half)
        return dW, d

    def table_grad(self, q, dX):
        g = np.zeros((self.levels * self.entries, 2))
        for l in range(self.levels):
            idx, wt = self.corners(q, l)
            for f in range(2):
                np.add.at(g[:, f], idx.ravel(), (wt.astype(np.float64) * dX[:, 2 * l + f][:, None]).ravel())
        return g

    @staticmethod
    def adam(p, m, v, g, step, lr, b1, b2, eps, sparse):
        """One Adam step (in place on float64 copies); g already scaled."""
        upd = np.ones(p.shape, bool) if not sparse else g != 0
        m[upd] = b1 * m[upd] + (1 - b1) * g[upd]
        v[upd] = b2 * v[upd] + (1 - b2) * g[upd] ** 2
        c1, c2 = 1 - b1 ** step, 1 - b2 ** step
        p[upd] -= lr * (m[upd] / c1) / (np.sqrt(v[upd] / c2) + eps)
        return p, m, v
