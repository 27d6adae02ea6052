"""CPU checks that pin the oracle's alias-table environment sampler
(ST_FLAG_ENV_ALIAS, the north-star replacement of the reference's mip warp,
src/Shaders/common/lights.hlsli:49-139):

- the table reproduces the texel weights luminance * sin(theta) computed here
  independently in float64, and its uv pdf integrates to 1;
- texels drawn with the renderer's own 23-bit PCG uniforms (nextFloat of
  rng.hlsli:23-60) occur at the frequencies the pdf states (chi-square over
  pdf classes and over rows);
- points are uniform inside their texel (both jitters);
- the pdf a sample reports equals the pdf evaluated at its uv.

Each statistical check is also run on a planted mistake restated here on the
CPU: the sampler that takes column, accept test and jitter from one uniform
(rx * N split into integer and fraction, which leaves 23 - log2(N) bits for
the fraction), which the checks must reject."""
import math

import numpy as np
import pytest
from scipy import stats

from oracle.oracle import OracleScene
from stratum2_amd import scenes

M32 = np.uint64(0xFFFFFFFF)


def pcg4d_np(x, y, z, w):
    """pcg4d (rng.hlsli:23-35) on uint32 arrays."""
    u = lambda a: (a.astype(np.uint64) & M32)  # noqa: E731
    x, y, z, w = [(u(a) * np.uint64(1664525) + np.uint64(1013904223)) & M32 for a in (x, y, z, w)]
    x = (x + y * w) & M32; y = (y + z * x) & M32; z = (z + x * y) & M32; w = (w + y * z) & M32
    x ^= x >> np.uint64(16); y ^= y >> np.uint64(16); z ^= z >> np.uint64(16); w ^= w >> np.uint64(16)
    x = (x + y * w) & M32; y = (y + z * x) & M32; z = (z + x * y) & M32; w = (w + y * z) & M32
    return [a.astype(np.uint32) for a in (x, y, z, w)]


def next_float4(n, seed=0, offset=0):
    """One RandomSampler(seed, pixel).nextFloat() per pixel of a 4096-wide
    image: float4 of asfloat(0x3f800000 | v >> 9) - 1."""
    i = np.arange(n, dtype=np.uint64)
    px, py = (i % 4096).astype(np.uint32), (i // 4096).astype(np.uint32)
    s = np.full(n, seed, np.uint32)
    w = np.full(n, offset + 1, np.uint32)
    o = pcg4d_np(px, py, s, w)
    return np.stack([((0x3F800000 | (v >> 9)).astype(np.uint32).view(np.float32) - np.float32(1)) for v in o], -1)


@pytest.fixture(scope="module")
def sky():
    sd = scenes.SceneData()
    p, i = scenes.quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1))
    vi, n = sd.add_mesh(p, i)
    sd.add_mesh_instance(vi, n, sd.add_material(scenes.material_record()))
    img = scenes.sky_environment()
    sd.set_environment((1, 1, 1), img)
    osc = OracleScene(sd)
    H, W = img.shape[:2]
    prob, alias, pdf = osc.env_alias_table(W * H)
    return dict(osc=osc, img=img, W=W, H=H, prob=prob, alias=alias, pdf=pdf)


def weights_f64(img):
    H, W = img.shape[:2]
    lum = 0.2126 * img[..., 0].astype(np.float64) + 0.7152 * img[..., 1] + 0.0722 * img[..., 2]
    st = np.sin(math.pi * (np.arange(H) + 0.5) / H)
    w = np.maximum(lum, 0) * st[:, None]
    return (w / w.sum() * (W * H)).reshape(-1)


def test_alias_table_reproduces_texel_weights(sky):
    N = sky["W"] * sky["H"]
    ref = weights_f64(sky["img"])
    np.testing.assert_allclose(sky["pdf"], ref, rtol=2e-7, atol=0)
    assert abs(sky["pdf"].astype(np.float64).mean() - 1) < 1e-6  # uv pdf integrates to 1
    prob, alias = sky["prob"].astype(np.float64), sky["alias"]
    assert np.all((prob >= 0) & (prob <= 1)) and np.all(alias < N)
    # probability mass of each texel: its own column's keep-probability plus
    # what every column aliasing to it gives away
    mass = prob.copy()
    give = alias != np.arange(N)
    mass += np.bincount(alias[give], weights=1 - prob[give], minlength=N)
    np.testing.assert_allclose(mass, ref, rtol=1e-4, atol=1e-6)


def texel_of(uv, W, H):
    x = np.clip(np.floor(uv[:, 0] * W).astype(np.int64), 0, W - 1)
    y = np.clip(np.floor(uv[:, 1] * H).astype(np.int64), 0, H - 1)
    return y * W + x


def buggy_sample(sky, rnd):
    """Planted mistake: column, accept test and u jitter from one uniform."""
    W, H, N = sky["W"], sky["H"], sky["W"] * sky["H"]
    x = rnd[:, 0] * np.float32(N)
    i = np.minimum(x.astype(np.int64), N - 1)
    f = (x - i.astype(np.float32)).astype(np.float32)
    p = sky["prob"][i]
    acc = f < p
    jx = np.where(acc, f / p, (f - p) / (np.float32(1) - p)).astype(np.float32)
    i = np.where(acc, i, sky["alias"][i])
    jx = np.minimum(jx, np.float32(0.99999994))
    u = ((i % W).astype(np.float32) + jx) / np.float32(W)
    v = ((i // W).astype(np.float32) + rnd[:, 1]) / np.float32(H)
    return np.stack([u, v], -1)


def class_chi2(sky, tex, nclass=40):
    """Chi-square p-value of the texel frequencies grouped into classes of
    equal expected count by pdf (dim texels together, bright together)."""
    pdf = sky["pdf"].astype(np.float64)
    order = np.argsort(pdf, kind="stable")
    cum = np.cumsum(pdf[order]) / pdf.sum()
    cls_sorted = np.minimum((cum * nclass).astype(np.int64), nclass - 1)
    cls = np.empty_like(cls_sorted)
    cls[order] = cls_sorted
    expect = np.bincount(cls, weights=pdf, minlength=nclass) / pdf.sum() * len(tex)
    obs = np.bincount(cls[tex], minlength=nclass)
    keep = expect > 20
    chi = ((obs[keep] - expect[keep]) ** 2 / expect[keep]).sum()
    return stats.chi2.sf(chi, keep.sum() - 1)


def row_chi2(sky, tex):
    W, H = sky["W"], sky["H"]
    rows = sky["pdf"].astype(np.float64).reshape(H, W).sum(1)
    expect = rows / rows.sum() * len(tex)
    obs = np.bincount(tex // W, minlength=H)
    keep = expect > 20
    chi = ((obs[keep] - expect[keep]) ** 2 / expect[keep]).sum()
    return stats.chi2.sf(chi, keep.sum() - 1)


N_SAMPLES = 1 << 22


@pytest.fixture(scope="module")
def draws(sky):
    rnd = next_float4(N_SAMPLES, seed=7)
    uv, pdf = sky["osc"].env_alias_sample(rnd)
    return rnd, uv, pdf


def test_alias_frequencies_match_pdf(sky, draws):
    rnd, uv, _ = draws
    tex = texel_of(uv, sky["W"], sky["H"])
    assert class_chi2(sky, tex) > 1e-4
    assert row_chi2(sky, tex) > 1e-4
    # the planted single-uniform sampler over-draws dim texels: rejected
    bad = texel_of(buggy_sample(sky, rnd), sky["W"], sky["H"])
    assert class_chi2(sky, bad) < 1e-12


def test_alias_points_uniform_in_texel(sky, draws):
    rnd, uv, _ = draws
    W, H = sky["W"], sky["H"]
    for frac in ((uv[:, 0].astype(np.float64) * W) % 1.0, (uv[:, 1].astype(np.float64) * H) % 1.0):
        obs = np.bincount(np.minimum((frac * 64).astype(np.int64), 63), minlength=64)
        assert stats.chisquare(obs).pvalue > 1e-4
    bad = buggy_sample(sky, rnd)
    frac = (bad[:, 0].astype(np.float64) * W) % 1.0
    obs = np.bincount(np.minimum((frac * 64).astype(np.int64), 63), minlength=64)
    assert stats.chisquare(obs).pvalue < 1e-12  # the planted jitter takes at most 16 values


def test_alias_sample_pdf_equals_evaluated_pdf(sky, draws):
    _, uv, pdf = draws
    W, H = sky["W"], sky["H"]
    tex = texel_of(uv, W, H)
    ev = sky["pdf"][tex]
    # the jitter stops short of 1, so no point rounds onto the next texel
    assert np.array_equal(ev, pdf)
    for k in range(0, 2000, 97):
        assert sky["osc"].env_alias_pdf(uv[k]) == sky["pdf"][tex[k]]
