"""CPU checks of the oracle's light tracer (TestRenderer with mLightTrace).

The light tracer and the camera-path tracer estimate the same image: every
path of two or more segments that ends on a light, over non-singular vertices,
is found once by each.  Camera paths sample the pixel area when jittered (the
light tracer integrates over it), and --minPathLength 2 removes the one-segment
paths (emitters seen directly) that light tracing cannot make.  The Cornell box
light faces the world origin, so the reference's one-sided test against the
origin (SampleIllumination(rnd, 0)) drops no light sample here.
"""
import math

import numpy as np

from oracle.oracle import OracleScene, lib
from stratum2_amd import abi, scenes

FL = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_LAMBERTIAN


def _pair(sd, W, H, max_depth, cam_spp, light_spp, factor):
    sc = OracleScene(sd)
    view = scenes.cornell_view(W, H)
    cam = sc.render(view, abi.default_params(W, H, max_depth=max_depth, spp=cam_spp, min_path_length=2,
                                             flags=FL | abi.ST_FLAG_PIXEL_JITTER))["hdr"][..., :3]
    lt = sc.render(view, abi.default_params(W, H, max_depth=max_depth, spp=light_spp, light_subpath_factor=factor,
                                            flags=FL | abi.ST_FLAG_TRACE_FROM_LIGHT))["hdr"][..., :3]
    first = sc.render(view, abi.default_params(W, H, max_depth=1, spp=1, flags=FL))
    return cam, lt, first["visibility"][..., 0] & 0xFFFF


def test_light_tracing_estimates_the_camera_path_image():
    """Block means within 4% and image mean within 1.5% over three bounces, and
    on the smooth-shaded sphere the per-pixel error stays at the noise level
    (0.09 rms here); without the adjoint shading-normal correction of
    shading_data.hlsli:46-51 it is 0.22."""
    W = H = 64
    cam, lt, inst = _pair(scenes.cornell_box_smooth(0, subdiv=1), W, H, 3, 512, 128, 16.0)
    assert np.isfinite(lt).all() and (lt >= 0).all()
    assert abs(lt.mean() / cam.mean() - 1) < 0.015, (lt.mean(), cam.mean())
    blk = lambda a: a.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3))
    bc, bl = blk(cam), blk(lt)
    lit = bc > 0.01
    ratio = bl[lit] / bc[lit]
    print("block ratio range", ratio.min(), ratio.max())
    assert np.all(np.abs(ratio - 1) < 0.04)
    m = inst == 7  # the sphere with vertex normals
    assert m.sum() > 100
    rel = (lt[m] - cam[m]) / cam[m]
    rms = float(np.sqrt((rel ** 2).mean()))
    print("sphere rms rel", rms)
    assert rms < 0.14


def test_light_subpath_count_and_seeds():
    """mLightSubpathCount = max(1, uint(W*H*factor)) (TestRenderer.cpp:452): a
    factor that rounds to zero paths still traces one; more paths lower the
    noise, not the mean; each sample uses its own seed."""
    sd = scenes.cornell_box(1)
    sc = OracleScene(sd)
    W = H = 16
    view = scenes.cornell_view(W, H)
    base = dict(max_depth=3, flags=FL | abi.ST_FLAG_TRACE_FROM_LIGHT)
    one = sc.render(view, abi.default_params(W, H, spp=1, light_subpath_factor=1e-6, **base))
    assert one["ray_counts"][0] >= 1 and one["ray_counts"][0] <= 3
    a = sc.render(view, abi.default_params(W, H, spp=64, light_subpath_factor=4.0, **base))["hdr"]
    b = sc.render(view, abi.default_params(W, H, spp=16, light_subpath_factor=16.0, **base))["hdr"]
    assert abs(a[..., :3].mean() / b[..., :3].mean() - 1) < 0.03
    s0 = sc.render(view, abi.default_params(W, H, spp=1, seed=5, **base))["hdr"]
    s1 = sc.render(view, abi.default_params(W, H, spp=1, seed=6, **base))["hdr"]
    s01 = sc.render(view, abi.default_params(W, H, spp=2, seed=5, **base))["hdr"]
    assert not np.array_equal(s0, s1)
    assert np.allclose(s01, (s0 + s1) / 2, rtol=1e-6, atol=1e-7)
    assert np.all(s01[..., 3] == 1)


def test_light_tracing_ignores_camera_only_modes():
    """NEE, ReSTIR DI and the direct-only mode are switched off when light
    tracing (TestRenderer.cpp:119-124): their flags do not change the image."""
    sd = scenes.cornell_box(2)
    sc = OracleScene(sd)
    W = H = 16
    view = scenes.cornell_view(W, H)
    p = abi.default_params(W, H, spp=2, max_depth=3, flags=FL | abi.ST_FLAG_TRACE_FROM_LIGHT)
    ref = sc.render(view, p)["hdr"]
    p.flags = FL | abi.ST_FLAG_TRACE_FROM_LIGHT | abi.ST_FLAG_RESTIR_DI | abi.ST_FLAG_DIRECT_ONLY | abi.ST_FLAG_FLOAT_NEE
    assert np.array_equal(sc.render(view, p)["hdr"], ref)
    p.flags = (FL & ~abi.ST_FLAG_SAMPLE_DIRECT) | abi.ST_FLAG_TRACE_FROM_LIGHT
    assert np.array_equal(sc.render(view, p)["hdr"], ref)


def test_adjoint_refraction_is_the_reversed_bsdf():
    """Light paths evaluate the adjoint BSDF: f*(i -> o) = f(o -> i).  With the
    reference's eval carrying |cos out|, the rough-glass transmission of
    disney.hlsli:110-138 with eta_factor = 1/eta^2 (Adjoint) must satisfy
    f*(i, o) / |o.z| = f(o, i) / |i.z|; the camera-path eval misses it by eta^2."""
    L = lib()
    eta = 1.5
    m = np.array(list(scenes.material_record(base_color=(0.9, 0.7, 0.5), specular=0.0, metallic=0.0, roughness=0.6,
                                              transmission=1.0, eta=eta)[:4]) + [0], np.uint32)
    eta_q = 1 + round((eta - 1) * 255) / 255  # the record stores eta - 1 in 8 bits
    rng = np.random.default_rng(7)
    fa, fr, fc = np.zeros(5, np.float32), np.zeros(5, np.float32), np.zeros(5, np.float32)
    n = 0
    for _ in range(400):
        i = rng.normal(size=3)
        o = rng.normal(size=3)
        i[2], o[2] = abs(i[2]) + 0.2, -abs(o[2]) - 0.2
        i = (i / np.linalg.norm(i)).astype(np.float32)
        o = (o / np.linalg.norm(o)).astype(np.float32)
        L.or_bsdf_eval_adjoint(m.ctypes.data, 0, 1, i.ctypes.data, o.ctypes.data, fa.ctypes.data)
        L.or_bsdf_eval_adjoint(m.ctypes.data, 0, 0, o.ctypes.data, i.ctypes.data, fr.ctypes.data)
        L.or_bsdf_eval_adjoint(m.ctypes.data, 0, 0, i.ctypes.data, o.ctypes.data, fc.ctypes.data)
        a, r, c = fa[:3] / abs(o[2]), fr[:3] / abs(i[2]), fc[:3] / abs(o[2])
        if not (r.max() > 1e-3):
            continue
        n += 1
        assert np.allclose(a, r, rtol=2e-4), (i, o, a, r)
        assert np.allclose(c / r, eta_q * eta_q, rtol=2e-4)
    assert n > 100


def test_adjoint_is_camera_eval_without_refraction_and_shading_normals():
    """Reflection and Lambert lobes have no adjoint form: both evals agree
    bit for bit (only refraction and the shading-normal ratio differ)."""
    L = lib()
    rng = np.random.default_rng(3)
    fa, fc = np.zeros(5, np.float32), np.zeros(5, np.float32)
    for lam in (0, 1):
        m = np.array(list(scenes.material_record(base_color=(0.6, 0.5, 0.4), specular=0.7, metallic=0.3, roughness=0.4,
                                                  sheen=0.5, clearcoat=0.4, transmission=0.5)[:4]) + [0], np.uint32)
        for _ in range(200):
            i = rng.normal(size=3)
            o = rng.normal(size=3)
            i[2], o[2] = abs(i[2]), abs(o[2])
            i = (i / np.linalg.norm(i)).astype(np.float32)
            o = (o / np.linalg.norm(o)).astype(np.float32)
            L.or_bsdf_eval_adjoint(m.ctypes.data, lam, 1, i.ctypes.data, o.ctypes.data, fa.ctypes.data)
            L.or_bsdf_eval_adjoint(m.ctypes.data, lam, 0, i.ctypes.data, o.ctypes.data, fc.ctypes.data)
            assert np.array_equal(fa, fc)
    assert math.isfinite(float(fa[0]))


def test_light_facing_away_from_the_origin_emits_nothing():
    """SampleIllumination(rnd, 0) tests the light's one-sidedness against the
    world origin (lights.hlsli:252-256): a Cornell box moved below the origin
    has its ceiling light facing away from it, and light tracing renders black."""
    sd = scenes.cornell_box(0)
    for i in range(len(sd.transforms)):
        t = np.array(sd.transforms[i], np.float32)
        t[1, 3] -= 3.0
        sd.transforms[i] = t
    W = H = 16
    view = scenes.make_view(W, H, position=(0.0, -2.0, 2.4))
    sc = OracleScene(sd)
    r = sc.render(view, abi.default_params(W, H, spp=2, max_depth=3, flags=FL | abi.ST_FLAG_TRACE_FROM_LIGHT))
    assert np.all(r["hdr"][..., :3] == 0) and r["ray_counts"][0] == 0
    cam = sc.render(view, abi.default_params(W, H, spp=2, max_depth=3, flags=FL))["hdr"]
    assert cam[..., :3].mean() > 0.01  # the camera paths still see it lit


def test_light_tracing_through_media_matches_camera_paths():
    """A scattering medium (sigma 0.6, albedo 0.8, g 0.3) around the ceiling
    light: light paths start inside it (GetMediumIndex, scene.hlsli:32-40),
    track through it (delta tracking), scatter in it (phase function toward
    the camera) and connect through it (ratio tracking).  Image mean within 2%
    and block means within 12% of the jittered camera paths; starting the
    light paths outside the medium instead puts blocks off by up to 90%.  The
    medium is thin because the camera paths' NEE rays stop short of the light
    (AdjustIntersectionDistance), which in a dense medium brightens them by
    the transmittance of the last centimetre."""
    sd = scenes.cornell_box(0)
    scenes.add_medium_box(sd, (0.6, 0.6, 0.6), (0.8, 0.8, 0.8), 0.3, (-0.7, 0.9, -0.7), (0.7, 1.99, 0.7))
    W = H = 48
    view = scenes.cornell_view(W, H)
    sc = OracleScene(sd)
    cam = sc.render(view, abi.default_params(W, H, max_depth=4, spp=384, min_path_length=2,
                                             flags=FL | abi.ST_FLAG_PIXEL_JITTER))["hdr"][..., :3]
    lt = sc.render(view, abi.default_params(W, H, max_depth=4, spp=48, light_subpath_factor=16.0,
                                            flags=FL | abi.ST_FLAG_TRACE_FROM_LIGHT))["hdr"][..., :3]
    assert np.isfinite(lt).all()
    assert abs(lt.mean() / cam.mean() - 1) < 0.02, (lt.mean(), cam.mean())
    blk = lambda a: a.reshape(6, 8, 6, 8, 3).mean(axis=(1, 3))
    ratio = blk(lt) / blk(cam)
    print("block ratio range", ratio.min(), ratio.max())
    assert np.all(np.abs(ratio - 1) < 0.12)
