"""Material images on the GPU (base colour and packed-word modulation at the
shading point's uv, trilinear over the image's mip chain at the lod of its
texture-coordinate screen size; kernels instantiated for scenes with images)
against the oracle on identical PCG streams, at the camera paths' per-pixel
bar."""
import numpy as np
import pytest

from stratum2_amd import abi, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["disney", "lambert_ggx", "ris", "ragged_alias"])
def test_textured_cornell_vs_oracle(gpu_ctx, case):
    from test_gpu_spheres import _both, _check
    sd = scenes.cornell_textured(0)
    W = H = 96
    flags = abi.ST_FLAG_DEFAULT
    kw = dict(max_depth=5, spp=4)
    if case == "lambert_ggx":
        flags |= abi.ST_FLAG_LAMBERT_GGX
    elif case == "ris":
        flags |= abi.ST_FLAG_RESTIR_DI
    elif case == "ragged_alias":
        W, H = 45, 37
        sd.set_environment(image=scenes.sky_environment(64, 32, sun_peak=50.0))
        flags |= abi.ST_FLAG_ENV_ALIAS
        kw.update(rr_start_depth=3)
    params = abi.default_params(W, H, flags=flags, **kw)
    print(f"case {case}", flush=True)
    g, r = _both(gpu_ctx, sd, scenes.cornell_view(W, H), params)
    _check(g, r, spheres=False)


def test_textured_refusals(gpu_ctx):
    from stratum2_amd.renderer import Scene, TestRenderer
    sc = Scene(gpu_ctx, scenes.cornell_textured(0))
    for fl in (abi.ST_FLAG_RESTIR_DI | abi.ST_FLAG_RESTIR_DI_REUSE, abi.ST_FLAG_RESTIR_DI | abi.ST_FLAG_RESTIR_DI_HASHGRID,
               abi.ST_FLAG_TRACE_FROM_LIGHT, abi.ST_FLAG_BDPT_LVC):
        with pytest.raises(RuntimeError, match="material images"):
            TestRenderer(sc).render(scenes.cornell_view(16, 16),
                                    abi.default_params(16, 16, max_depth=3, spp=1, flags=abi.ST_FLAG_DEFAULT | fl))
    for word in (5, 9):  # alpha masks are taken (the alpha-mask tests)
        sd = scenes.cornell_box(0)
        sd.images.append(np.ones((4, 4, 4), np.float32))
        w = scenes.material_record()
        w[word] = len(sd.images) - 1
        vi, n = sd.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([0, 1, 2], np.uint32))
        sd.add_mesh_instance(vi, n, sd.add_material(w))
        with pytest.raises(RuntimeError):
            Scene(gpu_ctx, sd)
