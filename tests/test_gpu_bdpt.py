"""GPU parity of the bidirectional mode (TestRenderer with gUseVC: light vertex
cache, VCM recursive MIS) against the oracle on identical PCG streams.  The
cache is ordered by light subpath and depth on both sides, so a camera vertex
picks the same light vertex.  Same bars as the other modes: at least 99% of
pixel channels within 1e-4 relative, image mean within 1%, ray counts within
1%; the primary visibility is bit-exact."""
import numpy as np
import pytest

from stratum2_amd import abi, scenes

pytestmark = pytest.mark.gpu
VC = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_BDPT_LVC


def _both(gpu_ctx, sd, view, params):
    import torch
    from oracle.oracle import OracleScene
    from stratum2_amd.renderer import Scene, TestRenderer
    out = TestRenderer(Scene(gpu_ctx, sd)).render(view, params, want_path_length=True)
    torch.cuda.synchronize()
    g = {k: (v.cpu().numpy() if v is not None else None) for k, v in out.items()}
    return g, OracleScene(sd).render(view, params)


def _check(g, r):
    hg, hr = g["hdr"][..., :3], r["hdr"][..., :3]
    assert np.isfinite(hg).all()
    assert np.array_equal(g["visibility"].view(np.uint32)[..., 0], r["visibility"][..., 0]), "primary hit ids differ"
    ok = np.abs(hg - hr) <= 1e-4 * np.maximum(np.abs(hr), 1e-3)
    print(f"pixel channels within 1e-4 rel: {ok.mean():.6f}, bit-identical {(hg == hr).mean():.6f}; "
          f"mean gpu {hg.mean():.6f} oracle {hr.mean():.6f}")
    assert ok.all()
    assert np.array_equal(g["path_length"].view(np.uint32), r["path_length"])
    rc = g["ray_counts"].view(np.uint64)
    assert int(rc[0]) == int(r["ray_counts"][0]) and int(rc[1]) == int(r["ray_counts"][1])


@pytest.mark.parametrize("case", ["lambert", "disney", "smooth_normals", "fp32_mis"])
def test_bdpt_cornell(gpu_ctx, case):
    W = H = 64
    fl, factor, sd = VC, 1.0, scenes.cornell_box(0)
    if case == "lambert":
        fl, sd = VC | abi.ST_FLAG_LAMBERTIAN, scenes.cornell_box(3)
    elif case == "smooth_normals":
        sd, factor = scenes.cornell_box_smooth(1, subdiv=2), 0.5
    elif case == "fp32_mis":
        fl, factor = VC | abi.ST_FLAG_VC_FP32_MIS | abi.ST_FLAG_PIXEL_JITTER, 2.0
    params = abi.default_params(W, H, max_depth=5, spp=4, flags=fl, light_subpath_factor=factor)
    g, r = _both(gpu_ctx, sd, scenes.cornell_view(W, H), params)
    assert r["hdr"][..., :3].mean() > 0.01
    _check(g, r)


def test_bdpt_environment_emitters_and_ragged_image(gpu_ctx):
    """Environment and emissive triangles (alias-free mip-warp sampling), an
    image that is not a multiple of the 8x8 path tiles, three segments."""
    sd = scenes.spheres_field(seed=2, n_spheres=40, n_emitters=8)
    W, H = 45, 37
    params = abi.default_params(W, H, max_depth=3, spp=3, flags=VC, light_subpath_factor=1.0)
    g, r = _both(gpu_ctx, sd, scenes.spheres_view(W, H), params)
    _check(g, r)


def test_bdpt_samples_rotate_through_the_wave_slots(gpu_ctx):
    """Six samples: waves rotate through the four slots, each with its own light
    vertex cache; the image and the per-sample path lengths still match."""
    W = H = 32
    sd = scenes.cornell_box(4)
    params = abi.default_params(W, H, max_depth=4, spp=6, flags=VC | abi.ST_FLAG_LAMBERTIAN, light_subpath_factor=1.5)
    g, r = _both(gpu_ctx, sd, scenes.cornell_view(W, H), params)
    _check(g, r)


def test_bdpt_rejects_unsupported(gpu_ctx):
    from stratum2_amd.renderer import Scene, TestRenderer
    W = H = 16
    sd = scenes.cornell_box(0)
    scenes.add_medium_box(sd, (0.5, 0.5, 0.5), (0.9, 0.9, 0.9), 0.0, (-0.5, 0.5, -0.5), (0.5, 1.5, 0.5))
    with pytest.raises(RuntimeError):
        TestRenderer(Scene(gpu_ctx, sd)).render(scenes.cornell_view(W, H), abi.default_params(W, H, flags=VC))
    tr = TestRenderer(Scene(gpu_ctx, scenes.cornell_box(0)))
    for kw in (dict(flags=VC | abi.ST_FLAG_RESTIR_DI), dict(flags=VC, min_path_length=2)):
        with pytest.raises(RuntimeError):
            tr.render(scenes.cornell_view(W, H), abi.default_params(W, H, **kw))
