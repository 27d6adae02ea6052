"""Converged-image check of the Cornell-box frame (256x256, maxPathLength 4,
Disney materials, 4096 spp): the HIP renderer against the CPU oracle, in its
float build (the renderer's own arithmetic) and its double build (the same
algorithm computing in double, libm's double transcendentals).

    python tests/converged_rmse.py gpu      OUT.npy          # on a GPU machine
    python tests/converged_rmse.py oracle   SEED OUT.npy     # CPU float build, seeds SEED..SEED+4095
    python tests/converged_rmse.py oracle64 SEED OUT.npy     # CPU double build
    python tests/converged_rmse.py compare GPU.npy F32_SAME.npy F32_INDEP.npy [F64_SAME.npy F64_INDEP.npy]

`compare` prints, for each pair, the RMSE over pixels and the difference of
the image means with its standard error.  "Same" images use the GPU's seeds
(0..4095, the same PCG streams); "indep" images use seeds 4096..8191.  The
RMSE of two independent renders is the Monte Carlo noise floor: an unbiased
renderer sits at it.  The mean difference is judged against its standard
error, estimated from the spread of the 256 block means (16x16 pixels) of an
independent pair.  Not collected by pytest: an oracle image takes minutes.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from stratum2_amd import abi, scenes  # noqa: E402

W = H = 256
SPP = 4096
MAX_DEPTH = 4


def params(seed):
    return abi.default_params(W, H, max_depth=MAX_DEPTH, spp=SPP, seed=seed)


def render_gpu(out):
    import torch
    from stratum2_amd.renderer import Context, Scene, TestRenderer
    ctx = Context(0)
    img = TestRenderer(Scene(ctx, scenes.cornell_box(0))).render(scenes.cornell_view(W, H), params(0))["hdr"]
    torch.cuda.synchronize()
    np.save(out, img.cpu().numpy()[..., :3])


def render_oracle(seed, out, fp64=False):
    from oracle.oracle import OracleScene
    img = OracleScene(scenes.cornell_box(0), fp64=fp64).render(scenes.cornell_view(W, H), params(seed),
                                                               want_path_length=False)["hdr"]
    np.save(out, img[..., :3])


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def block_means(a, n=16):
    a = a.astype(np.float64).mean(-1)
    return a.reshape(H // n, n, W // n, n).mean((1, 3)).ravel()


def mean_diff(a, b, se_pair):
    """Difference of the image means and its standard error: the spread of the
    block-mean differences of an independent pair, over sqrt(blocks)."""
    d = block_means(se_pair[0]) - block_means(se_pair[1])
    se = float(d.std(ddof=1) / np.sqrt(d.size))
    return float(a.astype(np.float64).mean() - b.astype(np.float64).mean()), se


def compare(g, f32_same, f32_indep, f64_same=None, f64_indep=None):
    g, a, b = np.load(g), np.load(f32_same), np.load(f32_indep)
    pairs = [("gpu vs float oracle, same streams", g, a), ("gpu vs float oracle, independent seeds", g, b),
             ("float oracle vs float oracle, independent seeds (noise floor)", a, b)]
    se_pair = (a, b)
    if f64_same:
        c, e = np.load(f64_same), np.load(f64_indep)
        pairs += [("gpu vs double oracle, same streams", g, c), ("gpu vs double oracle, independent seeds", g, e),
                  ("float oracle vs double oracle, same streams", a, c),
                  ("double oracle vs double oracle, independent seeds (noise floor)", c, e)]
    m = float(a.astype(np.float64).mean())
    for name, x, y in pairs:
        v = rmse(x, y)
        d, se = mean_diff(x, y, se_pair)
        print(f"{name}: RMSE {v:.6g} ({v / m:.4%} of the mean {m:.5g}); mean difference {d:+.3e} "
              f"(standard error of an independent pair {se:.2e}, {d / se:+.2f} of it)")


if __name__ == "__main__":
    if sys.argv[1] == "gpu":
        render_gpu(sys.argv[2])
    elif sys.argv[1] == "oracle":
        render_oracle(int(sys.argv[2]), sys.argv[3])
    elif sys.argv[1] == "oracle64":
        render_oracle(int(sys.argv[2]), sys.argv[3], fp64=True)
    else:
        compare(*sys.argv[2:7])
