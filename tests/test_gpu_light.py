"""GPU parity of light tracing (TestRenderer with mLightTrace) against the
oracle on identical PCG streams.  Same bars as the camera paths: at least 99%
of pixel channels within 1e-4 relative, image mean within 1%, ray counts within
1%.  Per-pixel sums are 32-bit integer atomics, so the GPU image does not depend
on the order of the light paths: waves of 1 or several samples agree bit for bit."""
import numpy as np
import pytest

from stratum2_amd import abi, scenes

pytestmark = pytest.mark.gpu
FL = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_TRACE_FROM_LIGHT


def _render(gpu_ctx, sd, view, params):
    import torch
    from stratum2_amd.renderer import Scene, TestRenderer
    out = TestRenderer(Scene(gpu_ctx, sd)).render(view, params)
    torch.cuda.synchronize()
    return {k: (v.cpu().numpy() if v is not None else None) for k, v in out.items()}


def _both(gpu_ctx, sd, view, params):
    from oracle.oracle import OracleScene
    return _render(gpu_ctx, sd, view, params), OracleScene(sd).render(view, params)


def _check(g, r):
    hg, hr = g["hdr"][..., :3], r["hdr"][..., :3]
    assert np.isfinite(hg).all()
    assert np.all(g["hdr"][..., 3] == 1)
    ok = np.abs(hg - hr) <= 1e-4 * np.maximum(np.abs(hr), 1e-3)
    print(f"pixel channels within 1e-4 rel: {ok.mean():.6f}, bit-identical {(hg == hr).mean():.6f}; "
          f"mean gpu {hg.mean():.6f} oracle {hr.mean():.6f}")
    assert ok.all()
    rc = g["ray_counts"].view(np.uint64)
    assert int(rc[0]) == int(r["ray_counts"][0]) and int(rc[1]) == int(r["ray_counts"][1])


@pytest.mark.parametrize("case", ["lambert", "disney", "smooth_normals"])
def test_light_tracing_cornell(gpu_ctx, case):
    W = H = 96
    if case == "lambert":
        sd, fl, factor = scenes.cornell_box(3), FL | abi.ST_FLAG_LAMBERTIAN, 1.0
    elif case == "disney":
        sd, fl, factor = scenes.cornell_box(0), FL, 2.0
    else:
        sd, fl, factor = scenes.cornell_box_smooth(1, subdiv=2), FL, 0.5
    params = abi.default_params(W, H, max_depth=5, spp=4, flags=fl, light_subpath_factor=factor)
    g, r = _both(gpu_ctx, sd, scenes.cornell_view(W, H), params)
    assert r["hdr"][..., :3].mean() > 0.01
    _check(g, r)


def test_light_tracing_environment_and_emitters(gpu_ctx):
    """Environment light paths start on the scene sphere's disc; emissive
    triangles facing away from the origin start none."""
    sd = scenes.spheres_field(seed=2, n_spheres=40, n_emitters=8)
    W, H = 96, 64
    view = scenes.make_view(W, H, position=(0.0, 3.0, 14.0), orientation=scenes.axis_angle_quat((1, 0, 0), -10.0))
    params = abi.default_params(W, H, max_depth=4, spp=2, flags=FL)
    g, r = _both(gpu_ctx, sd, view, params)
    _check(g, r)


def test_light_tracing_min_path_length_and_odd_sizes(gpu_ctx):
    sd = scenes.cornell_box(2)
    W, H = 45, 37  # not a multiple of the 8x8 tile: scanline pixel order
    params = abi.default_params(W, H, max_depth=4, spp=3, min_path_length=3, flags=FL | abi.ST_FLAG_LAMBERTIAN)
    g, r = _both(gpu_ctx, sd, scenes.cornell_view(W, H), params)
    _check(g, r)


def test_light_tracing_independent_of_samples_per_wave(gpu_ctx):
    sd = scenes.cornell_box(1)
    W = H = 64
    imgs = []
    for k in (1, 3, 0):
        p = abi.default_params(W, H, max_depth=4, spp=6, flags=FL, samples_per_wave=k)
        imgs.append(_render(gpu_ctx, sd, scenes.cornell_view(W, H), p)["hdr"])
    assert np.array_equal(imgs[0], imgs[1]) and np.array_equal(imgs[0], imgs[2])


def test_light_tracing_rejects_bad_factor(gpu_ctx):
    from stratum2_amd.renderer import Scene, TestRenderer
    sc = Scene(gpu_ctx, scenes.cornell_box(0))
    with pytest.raises(RuntimeError, match="light_subpath_factor"):
        TestRenderer(sc).render(scenes.cornell_view(16, 16), abi.default_params(16, 16, flags=FL, light_subpath_factor=-1.0))


def test_light_tracing_medium_around_the_light(gpu_ctx):
    """Light paths start inside a homogeneous medium (GetMediumIndex), scatter
    in it and connect to the camera through it (ratio tracking)."""
    sd = scenes.cornell_box(0)
    scenes.add_medium_box(sd, (1.5, 1.5, 1.5), (0.8, 0.8, 0.8), 0.3, (-0.7, 0.9, -0.7), (0.7, 1.99, 0.7))
    W = H = 64
    params = abi.default_params(W, H, max_depth=5, spp=2, flags=FL | abi.ST_FLAG_LAMBERTIAN)
    g, r = _both(gpu_ctx, sd, scenes.cornell_view(W, H), params)
    assert r["hdr"][..., :3].mean() > 0.01
    _check(g, r)


def test_light_tracing_emitters_in_a_medium(gpu_ctx):
    """Emissive spheres inside a large thin medium box: light paths start in
    the medium, scatter in it and among the spheres."""
    sd = scenes.spheres_field(seed=2, n_spheres=40, n_emitters=8, env=False)
    scenes.add_medium_box(sd, (0.3, 0.3, 0.3), (0.9, 0.9, 0.9), 0.5, (-6, 0, -6), (6, 4, 6))
    W, H = 96, 64
    view = scenes.make_view(W, H, position=(0.0, 3.0, 14.0), orientation=scenes.axis_angle_quat((1, 0, 0), -10.0))
    params = abi.default_params(W, H, max_depth=4, spp=2, flags=FL)
    g, r = _both(gpu_ctx, sd, view, params)
    assert r["hdr"][..., :3].mean() > 0
    _check(g, r)
