"""CPU checks that pin the oracle's restatement of st_render_cached (paths
ending in the radiance cache at a fixed path length; training paths that go
on and record (query, target) rows).  The reference has no cache; what fixes
the restatement is:

- with one training path per path (stride 1) nothing ends in the cache, so
  the image, path lengths and ray counts are the plain render's bit for bit
  (training paths consume the same random numbers and take the same turns),
  with and without screen-space reuse and in media;
- in a closed furnace (every wall emits Le, Lambert albedo a, no other
  light) the radiance a vertex at path length d scatters towards the
  previous vertex, for paths truncated at maximum depth D, is
  Le (a + a^2 + ... + a^(D - d)): the training targets' mean must match it
  (it would be one Le too high if the vertex's own emission were not in the
  gathered radiance, and too low by the NEE share if the vertex's NEE were);
- with the cache answering 0 the paths that end in it keep exactly the
  emission of their first d vertices (Le (1 + ... + a^(d-1)) in the white
  furnace); answering the closed form above restores the full-depth mean.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene, cached_image, training_rows
from stratum2_amd import abi, scenes

LE = 0.5


def _furnace(albedo=1.0):
    sd = scenes.SceneData()
    v, f = scenes.icosphere(2)
    v = v.astype(np.float32) * 3
    f = scenes._orient_outward(v, f, np.zeros(3))[:, [0, 2, 1]]  # facing inward
    vi, n = sd.add_mesh(v, f.reshape(-1).astype(np.uint32))
    sd.add_mesh_instance(vi, n, sd.add_material(scenes.material_record(base_color=(albedo,) * 3, emission=(LE, LE, LE))))
    return sd


def _same(a, b):
    assert np.array_equal(a["hdr"], b["hdr"])
    assert np.array_equal(a["path_length"], b["path_length"])
    assert np.array_equal(a["ray_counts"], b["ray_counts"])


def test_stride_one_is_the_plain_render_cornell():
    sd = scenes.cornell_box(0)
    W = H = 32
    view = scenes.cornell_view(W, H)
    o = OracleScene(sd)
    for flags in (abi.ST_FLAG_DEFAULT, abi.ST_FLAG_DEFAULT | abi.ST_FLAG_LAMBERTIAN):
        p = abi.default_params(W, H, max_depth=4, spp=2, flags=flags, rr_start_depth=2)
        plain = o.render(view, p)
        for depth in (1, 2, 3):
            res, recs = o.render_cached(view, p, depth, 1, frame=7)
            _same(res, plain)
            assert not (recs[..., 0] == 1).any()
            assert (recs[..., 0] == 2).any()
            # every path's record splits its radiance into fp32 and fixed-point parts
            tot = (recs[..., 19:22] + recs[..., 22:25]).mean(0)
            assert np.allclose(tot, plain["hdr"][..., :3], rtol=1e-5, atol=1e-6)


def test_stride_one_is_the_plain_render_headline_modes():
    sd = scenes.spheres_field(seed=4, n_spheres=30, n_emitters=12)
    W, H = 48, 32
    view = scenes.make_view(W, H, position=(0.0, 3.0, 12.0), orientation=scenes.axis_angle_quat((1, 0, 0), -10.0))
    o = OracleScene(sd)
    flags = (abi.ST_FLAG_DEFAULT | abi.ST_FLAG_ENV_ALIAS | abi.ST_FLAG_RESTIR_DI | abi.ST_FLAG_RESTIR_DI_REUSE)
    p = abi.default_params(W, H, max_depth=8, spp=2, flags=flags, rr_start_depth=3, candidate_samples=32, max_m=3.0,
                           spatial_neighbours=5, spatial_radius=30.0)
    plain = o.render(view, p)
    res, recs = o.render_cached(view, p, 3, 1, frame=1)
    _same(res, plain)


def test_stride_one_is_the_plain_render_media():
    sd, _ = scenes.cornell_cloud(0, n=16)
    W = H = 24
    view = scenes.cornell_view(W, H)
    o = OracleScene(sd)
    p = abi.default_params(W, H, max_depth=6, spp=1, max_null_collisions=1000)
    plain = o.render(view, p)
    res, recs = o.render_cached(view, p, 2, 1)
    _same(res, plain)
    # some cache-depth vertices are medium vertices: normal 0
    rec2 = recs[recs[..., 0] == 2]
    assert (np.abs(rec2[:, 7:10]).sum(1) == 0).any()


def test_selection_one_training_path_per_run():
    sd = scenes.cornell_box(0)
    W = H = 16
    o = OracleScene(sd)
    p = abi.default_params(W, H, max_depth=4, spp=3)
    stride = 5
    _, recs = o.render_cached(scenes.cornell_view(W, H), p, 2, stride, frame=3)
    kinds = recs[..., 0].reshape(-1)
    g = np.nonzero(kinds == 2)[0]
    assert len(np.unique(g // stride)) == len(g)  # at most one training path per run
    from oracle.oracle import lib
    L = lib()
    for gi in g:
        row = gi // stride
        assert gi % stride == L.or_pcg(row ^ L.or_pcg(3)) % stride


@pytest.mark.parametrize("depth", [2, 3])
def test_white_furnace_targets_and_images(depth):
    D = 5
    sd = _furnace(1.0)
    W = H = 24
    view = scenes.make_view(W, H, position=(0, 0, 0))
    p = abi.default_params(W, H, max_depth=D, spp=8, flags=abi.ST_FLAG_DEFAULT | abi.ST_FLAG_LAMBERTIAN)
    stride = 4
    o = OracleScene(sd)
    res, recs = o.render_cached(view, p, depth, stride, frame=2)
    q, t = training_rows(recs, stride)
    tg = t[t[:, 3] > 0, :3]
    assert len(tg) > 500
    # a = 1: every later vertex adds Le, so L_s = Le (D - depth)
    assert abs(tg.mean() - LE * (D - depth)) < 0.02 * LE * D, tg.mean()
    # the cache answering 0: ended paths keep Le per vertex up to depth
    zero = cached_image(recs)
    assert np.array_equal(zero[..., :3], res["hdr"][..., :3])
    train_frac = (recs[..., 0] == 2).mean()
    want0 = LE * (depth * (1 - train_frac) + D * train_frac)
    assert abs(zero[..., :3].mean() - want0) < 0.02 * LE * D
    # answering the closed form restores the full-depth mean Le D
    n1 = int((recs[..., 0] == 1).sum())
    full = cached_image(recs, np.full((n1, 3), LE * (D - depth), np.float32))
    assert abs(full[..., :3].mean() - LE * D) < 0.02 * LE * D


def test_grey_furnace_targets_closed_form():
    """Albedo 128/255 with Russian roulette at every vertex: the targets' mean
    is Le (a + ... + a^(D - d)) within 4 standard errors."""
    D, depth = 6, 2
    a = 128 / 255
    sd = _furnace(0.5)
    W = H = 32
    view = scenes.make_view(W, H, position=(0, 0, 0))
    p = abi.default_params(W, H, max_depth=D, spp=16, flags=abi.ST_FLAG_DEFAULT | abi.ST_FLAG_LAMBERTIAN)
    _, recs = OracleScene(sd).render_cached(view, p, depth, 2, frame=5)
    _, t = training_rows(recs, 2)
    tg = t[t[:, 3] > 0, :3].mean(1)
    want = LE * sum(a ** j for j in range(1, D - depth + 1))
    se = tg.std() / np.sqrt(len(tg))
    assert abs(tg.mean() - want) < 4 * se + 1e-4, (tg.mean(), want, se)


def test_refusals():
    sd = scenes.cornell_box(0)
    o = OracleScene(sd)
    view = scenes.cornell_view(8, 8)
    p = abi.default_params(8, 8, max_depth=3, spp=1)
    with pytest.raises(RuntimeError):
        o.render_cached(view, p, 3, 1)  # depth must be below max_depth
    with pytest.raises(RuntimeError):
        o.render_cached(view, p, 0, 1)
    p.flags |= abi.ST_FLAG_RESTIR_DI
    with pytest.raises(RuntimeError):
        o.render_cached(view, p, 2, 1)  # RIS at every vertex is not combined with the cache
