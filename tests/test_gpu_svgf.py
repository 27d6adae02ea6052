"""GPU parity of the SVGF denoiser (st_denoise) against the oracle
(OracleDenoiser over oracle/svgf_oracle.c) on the same frames.

Bars: temporal accumulation uses only IEEE-rounded operations (add, multiply,
divide, square root), so without filter passes the output and the
accumulated sample counts must be bit-identical.  The variance estimate and the
filter passes use expf and powf (ocml on the GPU, glibc on the host, both
within 1-2 ulp), so there every channel must agree to 1e-4 relative (absolute
floor 1e-6) and 99.9% of channels to 1e-5."""
import numpy as np
import pytest

from stratum2_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _random_frames(seed, W, H, n, inst_choices=(1, 2, 0xFFFF)):
    from oracle.oracle import lib
    rng = np.random.default_rng(seed)
    frames = []
    normals = rng.normal(size=(4, 3))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    packed = [lib().or_pack_normal(np.ascontiguousarray(v, np.float32).ctypes.data) for v in normals]
    for _ in range(n):
        rad = np.ones((H, W, 4), np.float32)
        rad[..., :3] = rng.gamma(0.7, 1.0, (H, W, 3))
        bad = rng.uniform(size=(H, W)) < 0.003
        rad[bad, 1] = np.nan
        vis = np.zeros((H, W, 2), np.uint32)
        vis[..., 0] = rng.choice(inst_choices, size=(H, W), p=[0.6, 0.3, 0.1]) | (rng.integers(0, 100, (H, W)) << 16)
        vis[..., 1] = np.array(packed, np.uint32)[rng.integers(0, 4, (H, W))]
        depth = np.zeros((H, W, 4), np.float32)
        depth[..., 0] = rng.uniform(1, 3, (H, W))
        depth[..., 1] = depth[..., 0] + rng.normal(0, 1.5, (H, W))
        depth[..., 2:] = rng.uniform(0.2, 1.5, (H, W, 2))
        depth[vis[..., 0] == 0xFFFF, 0] = np.inf
        x, y = np.meshgrid(np.arange(W), np.arange(H))
        shift = rng.normal(0, 0.7, (H, W, 2))
        uv = np.stack([(x + 0.5 + shift[..., 0]) / W, (y + 0.5 + shift[..., 1]) / H], -1).astype(np.float32)
        uv[rng.uniform(size=(H, W)) < 0.01] = (-3.5, 7.25)
        albedo = np.ones((H, W, 4), np.float32)
        albedo[..., :3] = rng.uniform(0, 1, (H, W, 3))
        frames.append(dict(hdr=rad, albedo=albedo, prev_uv=uv, visibility=vis, depth=depth))
    return frames


def _to_device(f):
    import torch
    d = {}
    for k, v in f.items():
        t = torch.from_numpy(np.ascontiguousarray(v.view(np.int32) if v.dtype == np.uint32 else v))
        d[k] = t.cuda()
    return d


def _compare(g, r, exact):
    if exact:
        same = (g == r) | (np.isnan(g) & np.isnan(r))
        assert same.all(), f"{(~same).sum()} channels differ; first at {np.argwhere(~same)[:3].tolist()}"
        return
    # the first frame of a history passes the radiance through, NaNs included
    assert np.array_equal(np.isnan(g), np.isnan(r))
    g, r = np.nan_to_num(g, nan=0.0), np.nan_to_num(r, nan=0.0)
    err = np.abs(g.astype(np.float64) - r)
    scale = np.maximum(np.abs(r.astype(np.float64)), 1e-2)
    rel = err / scale
    print(f"max rel {rel.max():.3g}, within 1e-5: {(rel <= 1e-5).mean():.5f}")
    assert (err <= 1e-4 * scale + 1e-6).all()
    assert (rel <= 1e-5).mean() >= 0.999


CONFIGS = [
    dict(flags=abi.ST_DENOISE_DEFAULT),
    dict(flags=abi.ST_DENOISE_DEFAULT | abi.ST_DENOISE_SHADOW_HYSTERESIS, history_limit=3),
    dict(flags=0),
    dict(flags=abi.ST_DENOISE_REPROJECTION | abi.ST_DENOISE_CHECK_DEPTH, history_limit=2),
    dict(flags=abi.ST_DENOISE_DEFAULT, atrous_iterations=3, filter_type=abi.ST_FILTER_BOX3),
    dict(flags=abi.ST_DENOISE_DEFAULT, atrous_iterations=4, filter_type=abi.ST_FILTER_ATROUS, history_limit=16,
         history_tap=1),
    dict(flags=abi.ST_DENOISE_REPROJECTION, atrous_iterations=2, filter_type=abi.ST_FILTER_BOX5, history_limit=8,
         history_tap=2, variance_boost_length=4.0),
    dict(flags=abi.ST_DENOISE_DEFAULT, atrous_iterations=3, filter_type=abi.ST_FILTER_SUBSAMPLED, history_limit=4),
    dict(flags=abi.ST_DENOISE_CHECK_NORMAL, atrous_iterations=3, filter_type=abi.ST_FILTER_BOX3_SUBSAMPLED,
         history_limit=5, history_tap=3),
    dict(flags=abi.ST_DENOISE_DEFAULT, atrous_iterations=2, filter_type=abi.ST_FILTER_BOX5_SUBSAMPLED,
         history_limit=100),
]


@pytest.mark.parametrize("ci", range(len(CONFIGS)))
def test_denoiser_matches_oracle_on_random_frames(gpu_ctx, ci):
    import torch
    from oracle.oracle import OracleDenoiser
    from stratum2_amd.renderer import Denoiser
    cfg = CONFIGS[ci]
    W, H = 75, 43   # ragged against the 64x4 tiles
    view = scenes.make_view(W, H)
    frames = _random_frames(100 + ci, W, H, 5)
    gden, oden = Denoiser(gpu_ctx, W, H, **cfg), OracleDenoiser(W, H)
    imap = np.array([0, 1, 2], np.uint32)
    for k, f in enumerate(frames):
        if k == 3:
            gden.reset_accumulation()
            oden.reset()
        m = torch.from_numpy(imap.view(np.int32)).cuda() if k % 2 else None
        g = gden.denoise(view, _to_device(f), instance_index_map=m)
        torch.cuda.synchronize()
        r = oden.denoise(view, f["hdr"], f["albedo"], f["prev_uv"], f["visibility"], f["depth"], gden.params,
                         instance_index_map=imap if k % 2 else None)
        print(f"config {ci} frame {k}: frames {gden.accumulated_frames}/{oden.accumulated_frames}")
        assert gden.accumulated_frames == oden.accumulated_frames
        _compare(g.cpu().numpy(), r, exact=cfg.get("atrous_iterations", 0) == 0)


def test_denoiser_on_rendered_frames_then_tonemap(gpu_ctx):
    """TestRenderer's loop with the denoiser (TestRenderer.cpp:463-470,
    656-670): each frame seeded with the accumulated frame count, denoised with
    albedo demodulation and three box filter passes, then tonemapped with the
    albedo."""
    import torch
    from oracle.oracle import OracleDenoiser, tonemap
    from stratum2_amd.renderer import Denoiser, Scene, TestRenderer
    W = H = 64
    sc = Scene(gpu_ctx, scenes.cornell_box(0))
    view = scenes.cornell_view(W, H)
    gden = Denoiser(gpu_ctx, W, H, atrous_iterations=3, history_limit=32)
    oden = OracleDenoiser(W, H)
    tr = TestRenderer(sc)
    for _ in range(4):
        seed = gden.accumulated_frames
        frame = tr.render(view, abi.default_params(W, H, max_depth=4, spp=1, seed=seed), want_aux=True)
        out = gden.denoise(view, frame)
        torch.cuda.synchronize()
        host = {k: frame[k].cpu().numpy() for k in ("hdr", "albedo", "prev_uv", "visibility", "depth")}
        r = oden.denoise(view, host["hdr"], host["albedo"], host["prev_uv"], host["visibility"], host["depth"],
                         gden.params)
        _compare(out.cpu().numpy(), r, exact=False)
    ldr = torch.empty_like(out)
    abi.check(gpu_ctx.lib, gpu_ctx.lib.st_tonemap(gpu_ctx.h, out.data_ptr(), frame["albedo"].data_ptr(), W, H, 0.0,
                                                  abi.ST_TONEMAP_RAW, 1, ldr.data_ptr()), "st_tonemap")
    torch.cuda.synchronize()
    want, _ = tonemap(out.cpu().numpy(), host["albedo"], exposure=0.0, mode=abi.ST_TONEMAP_RAW, gamma=True)
    np.testing.assert_allclose(ldr.cpu().numpy(), want, rtol=1e-5, atol=1e-6)


def test_denoiser_rejects_bad_arguments(gpu_ctx):
    import torch
    from stratum2_amd.renderer import Denoiser
    W, H = 32, 16
    f = _to_device(_random_frames(1, W, H, 1)[0])
    with pytest.raises(RuntimeError):
        Denoiser(gpu_ctx, W, H, filter_type=9).denoise(scenes.make_view(W, H), f)
    with pytest.raises(RuntimeError):
        Denoiser(gpu_ctx, W, H).denoise(scenes.make_view(W + 1, H), f)
    with pytest.raises(RuntimeError):
        Denoiser(gpu_ctx, W, H, atrous_iterations=17).denoise(scenes.make_view(W, H), f)
    with pytest.raises(ValueError):
        Denoiser(gpu_ctx, W + 1, H).denoise(scenes.make_view(W + 1, H), f)
    torch.cuda.synchronize()
