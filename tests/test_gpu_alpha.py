"""Alpha-masked triangles on the GPU (the traversal kernels' ALPHA variants,
launched only for scenes whose mesh materials carry a mask) against the
oracle on identical inputs: ray queries bit for bit, camera paths at the
per-pixel bar of the other camera-path tests."""
import numpy as np
import pytest

from stratum2_amd import abi, scenes

pytestmark = pytest.mark.gpu


def test_alpha_rays_vs_oracle(gpu_ctx):
    import torch
    from oracle.oracle import OracleScene
    from stratum2_amd.renderer import Scene
    sd = scenes.cornell_alpha(0)
    rng = np.random.default_rng(11)
    n = 20000
    o = rng.uniform([-0.95, 0.05, -0.95], [0.95, 1.95, 0.95], (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 3], rays[:, 4:7] = o, np.inf, d
    rays[::7, 3] = 0.3
    sc = Scene(gpu_ctx, sd)
    rt = torch.from_numpy(rays).cuda()
    gc = sc.trace_rays(rt).cpu().numpy().view(np.uint32)
    ga = sc.trace_rays(rt, closest=False).cpu().numpy().view(np.uint32)
    torch.cuda.synchronize()
    osc = OracleScene(sd)
    rc = osc.trace_rays(rays)
    ra = osc.trace_rays(rays, closest=False)
    masked = (rc[:, 1] & 0xFFFF) >= len(sd.instances) - 2
    assert masked.sum() > 300
    assert np.array_equal(gc, rc)
    assert np.array_equal(ga, ra)


@pytest.mark.parametrize("case", ["disney", "lambert_ggx", "ris", "ragged"])
def test_alpha_cornell_vs_oracle(gpu_ctx, case):
    """The Cornell box with a masked panel in view and a masked grille under
    the light (camera rays, shadow rays and bounces all take the test)."""
    from test_gpu_spheres import _both, _check
    sd = scenes.cornell_alpha(0)
    W = H = 96
    flags = abi.ST_FLAG_DEFAULT
    if case == "lambert_ggx":
        flags |= abi.ST_FLAG_LAMBERT_GGX
    elif case == "ris":
        flags |= abi.ST_FLAG_RESTIR_DI
    elif case == "ragged":
        W, H = 45, 37
    params = abi.default_params(W, H, flags=flags, max_depth=5, spp=4)
    print(f"case {case}", flush=True)
    g, r = _both(gpu_ctx, sd, scenes.cornell_view(W, H), params)
    vis = g["visibility"].view(np.uint32)[..., 0] & 0xFFFF
    assert (vis == len(sd.instances) - 2).mean() > 0.02  # the panel is in view
    _check(g, r, spheres=False)


def test_alpha_update_refused(gpu_ctx):
    from stratum2_amd.renderer import Scene
    sd = scenes.cornell_alpha(0)
    sc = Scene(gpu_ctx, sd)
    with pytest.raises(RuntimeError, match="alpha"):
        sc.update(np.stack([np.array(t, np.float32) for t in sd.transforms]))
