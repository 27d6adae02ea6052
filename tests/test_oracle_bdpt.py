"""CPU checks of the oracle's bidirectional mode (TestRenderer with gUseVC:
light vertex cache, VCM recursive MIS without merging).

With dVC / dVCM kept in fp32 (ST_FLAG_VC_FP32_MIS) the MIS weights of the four
strategies (emission hit, NEE, light-vertex connection, light subpath to the
camera) sum to one, so the bidirectional image estimates the path tracer's.
Both render jittered pixels: the light-to-camera strategy integrates over the
pixel area.  With the reference's fp16 storage the weights overflow on longer
paths and the image comes out darker, which the second test pins."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from stratum2_amd import abi, scenes

W = H = 64
JIT = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_PIXEL_JITTER
VC = abi.ST_FLAG_BDPT_LVC


def _render(sc, depth, spp, flags, **kw):
    p = abi.default_params(W, H, max_depth=depth, spp=spp, flags=flags, **kw)
    return sc.render(scenes.cornell_view(W, H), p)["hdr"][..., :3]


def _blocks(a):
    return a.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3))


@pytest.mark.parametrize("smooth", [False, True])
def test_bdpt_estimates_the_path_traced_image(smooth):
    """Lambert Cornell boxes (flat, and with a vertex-normal sphere), four
    segments: image means within 1%, 8x8 blocks within 5%."""
    sd = scenes.cornell_box_smooth(0, subdiv=1) if smooth else scenes.cornell_box(0)
    sc = OracleScene(sd)
    fl = JIT | abi.ST_FLAG_LAMBERTIAN
    pt = _render(sc, 4, 1024, fl)
    vc = _render(sc, 4, 256, fl | VC | abi.ST_FLAG_VC_FP32_MIS)
    assert np.isfinite(vc).all() and (vc >= 0).all()
    assert abs(vc.mean() / pt.mean() - 1) < 0.01, (vc.mean(), pt.mean())
    bp, bv = _blocks(pt), _blocks(vc)
    lit = bp > 0.01
    ratio = bv[lit] / bp[lit]
    print("block ratio range", ratio.min(), ratio.max())
    assert np.all(np.abs(ratio - 1) < 0.05)


def test_bdpt_disney_mean():
    """Glossy and metallic Disney materials: the image mean within 2% (their
    blocks are dominated by caustic noise in both estimators)."""
    sc = OracleScene(scenes.cornell_box(0))
    pt = _render(sc, 4, 1024, JIT)
    vc = _render(sc, 4, 256, JIT | VC | abi.ST_FLAG_VC_FP32_MIS)
    assert abs(vc.mean() / pt.mean() - 1) < 0.02, (vc.mean(), pt.mean())


def test_fp16_mis_storage_darkens_longer_paths():
    """The reference's fp16 dVC / dVCM overflow on three-segment paths of this
    box at one light subpath per pixel (the contribution of those paths comes
    out about 8% low), less so with fewer light subpaths (their count enters
    the camera's dVCM); fp32 storage removes the loss.  Paths of one and two
    segments are unaffected."""
    sc = OracleScene(scenes.cornell_box(0))
    fl = JIT | abi.ST_FLAG_LAMBERTIAN
    seg3 = lambda flags, **kw: (_render(sc, 3, 256, flags, **kw) - _render(sc, 2, 256, flags, **kw)).mean()  # noqa: E731
    pt3 = seg3(fl)
    r16 = seg3(fl | VC) / pt3
    r16_quarter = seg3(fl | VC, light_subpath_factor=0.25) / pt3
    r32 = seg3(fl | VC | abi.ST_FLAG_VC_FP32_MIS) / pt3
    print("three-segment ratios: fp16", r16, "fp16 at 0.25 subpaths per pixel", r16_quarter, "fp32", r32)
    assert 0.87 < r16 < 0.95
    assert r16 < r16_quarter < 1.0
    assert abs(r32 - 1) < 0.015
    two_pt, two_vc = _render(sc, 2, 256, fl), _render(sc, 2, 256, fl | VC)
    assert abs(two_vc.mean() / two_pt.mean() - 1) < 0.003


def test_bdpt_one_segment_and_mode_precedence():
    """With one segment only emitters seen directly remain, identically; light
    tracing overrides the bidirectional flag (mLightTrace clears gUseVC)."""
    sc = OracleScene(scenes.cornell_box(1))
    a = _render(sc, 1, 2, abi.ST_FLAG_DEFAULT)
    b = _render(sc, 1, 2, abi.ST_FLAG_DEFAULT | VC)
    assert np.array_equal(a, b)
    lt = _render(sc, 3, 2, abi.ST_FLAG_DEFAULT | abi.ST_FLAG_TRACE_FROM_LIGHT)
    both = _render(sc, 3, 2, abi.ST_FLAG_DEFAULT | abi.ST_FLAG_TRACE_FROM_LIGHT | VC)
    assert np.array_equal(lt, both)


@pytest.mark.parametrize("extra", [dict(flags=abi.ST_FLAG_RESTIR_DI), dict(flags=abi.ST_FLAG_FLOAT_NEE),
                                   dict(flags=abi.ST_FLAG_DIRECT_ONLY), dict(min_path_length=2), "media"])
def test_bdpt_rejects_unsupported_combinations(extra):
    if extra == "media":
        sd = scenes.cornell_box(0)
        scenes.add_medium_box(sd, (0.5, 0.5, 0.5), (0.9, 0.9, 0.9), 0.0, (-0.5, 0.5, -0.5), (0.5, 1.5, 0.5))
        sc, kw = OracleScene(sd), {}
    else:
        sc, kw = OracleScene(scenes.cornell_box(0)), dict(extra)
    fl = abi.ST_FLAG_DEFAULT | VC | kw.pop("flags", 0)
    with pytest.raises(RuntimeError):
        _render(sc, 3, 1, fl, **kw)
