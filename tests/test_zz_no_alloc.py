"""st_trace_rays_noalloc (allocation-free raw ray queries over a fixed device
scratch) against st_trace_rays, through the C ABI: bit-exact hits, two calls
in a row with the second smaller, closest and any-hit, on a shallow scene and
on one whose traversal stacks spill past their LDS part into the scratch."""
import numpy as np
import pytest

from stratum2_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _lib():
    return abi.load_library()


def _random_rays(rng, n, scale, centre):
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(-1, 1, (n, 3)) * scale + centre
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = np.where(rng.uniform(size=n) < 0.8, np.inf, rng.uniform(0.01, 2.0, n))
    return rays


def _stacked_scene(rng, k=60000):
    """Nearly parallel triangles stacked along -z: rays down the stack push
    three entries per tree level, past the 16 LDS entries of each lane."""
    z = -1.0 - 0.0005 * np.arange(k)
    j = rng.uniform(-0.3, 0.3, (k, 3))
    pos = np.zeros((k, 3, 3), np.float32)
    pos[:, 0] = np.stack([-2 + j[:, 0], -2 + j[:, 1], z], -1)
    pos[:, 1] = np.stack([2 + j[:, 2], -2 + j[:, 1], z], -1)
    pos[:, 2] = np.stack([j[:, 0], 2 + j[:, 2], z], -1)
    sd = scenes.SceneData()
    vi, n = sd.add_mesh(pos.reshape(-1, 3), np.arange(3 * k, dtype=np.uint32))
    sd.add_mesh_instance(vi, n, sd.add_material(scenes.material_record()))
    return sd


def _stacked_rays(rng, nr):
    o = np.zeros((nr, 3))
    o[:, :2] = rng.uniform(-0.5, 0.5, (nr, 2))
    o[nr // 2:, 2] = rng.uniform(-25, -2, nr - nr // 2)
    d = np.stack([rng.normal(0, 0.02, nr), rng.normal(0, 0.02, nr), -np.ones(nr)], -1)
    d[nr // 2::2, 2] = 1.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((nr, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, np.inf, d
    return rays


def _both(lib, sc, rays, closest):
    import torch
    n = len(rays)
    gr = torch.from_numpy(rays).cuda()
    ref = sc.trace_rays(gr, closest).cpu().numpy()
    out = torch.full((n, 4) if closest else (n,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    st = lib.st_trace_rays_noalloc(sc.h, gr.data_ptr(), n, int(closest), out.data_ptr())
    assert st == abi.ST_OK, lib.st_last_error()
    return ref, out.cpu().numpy()


@pytest.mark.parametrize("which", ["cornell", "stacked"])
def test_noalloc_trace_equals_trace_rays(gpu_ctx, which):
    from stratum2_amd.renderer import Scene
    lib = _lib()
    rng = np.random.default_rng(21)
    if which == "cornell":
        sd = scenes.cornell_box(0)
        big, small = _random_rays(rng, 200000, 0.9, (0, 1, 0)), _random_rays(rng, 3001, 0.9, (0, 1, 0))
    else:
        sd = _stacked_scene(rng)
        big, small = _stacked_rays(rng, 40000), _stacked_rays(rng, 1500)
    sc = Scene(gpu_ctx, sd)
    if which == "stacked":
        assert 3 * sc.stats()["depth"] > 16  # the stacks reach the scratch's spill region
    for closest in (True, False):
        for rays in (big, small):  # two calls in a row, the second smaller
            ref, got = _both(lib, sc, rays, closest)
            assert np.array_equal(ref, got), (which, closest, len(rays))
    if which == "stacked":  # the device-built tree is deeper: the grid shrinks to fit the scratch
        sc.update(np.stack(sd.transforms))
        for rays in (big, small):
            ref, got = _both(lib, sc, rays, True)
            assert np.array_equal(ref, got)


def test_noalloc_argument_errors(gpu_ctx):
    from stratum2_amd.renderer import Scene
    lib = _lib()
    sc = Scene(gpu_ctx, scenes.cornell_box(0))
    assert lib.st_trace_rays_noalloc(None, None, 0, 1, None) == abi.ST_ERROR_INVALID_ARGUMENT
    assert lib.st_trace_rays_noalloc(sc.h, None, 5, 1, None) == abi.ST_ERROR_INVALID_ARGUMENT
    assert lib.st_trace_rays_noalloc(sc.h, 16, 0x80000000, 1, 16) == abi.ST_ERROR_LIMIT
    assert lib.st_trace_rays_noalloc(sc.h, None, 0, 1, None) == abi.ST_OK
