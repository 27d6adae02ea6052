"""Pins the oracle's path-replay adjoint (or_adjoint), the checker of
st_adjoint, by means independent of its bookkeeping:

- albedo: with roulette off nothing but the path weights depends on the
  per-triangle albedo, so the fixed-seed central difference of
  sum(d_hdr * hdr) over the albedo table (eps 1e-3) is the gradient;
- density: the adjoint differentiates the path weights with the tracking
  decisions held fixed.  The oracle's weight grid (or_scene_set_weight_grid)
  feeds the weights while the decisions read the unchanged grid, so the
  fixed-seed central difference over one voxel of the weight grid (eps 1e-3)
  is that derivative, whatever events the tracking sampled;
- unbiasedness of the detached derivative, which no fixed-seed difference
  shows: through a uniform slab of density rho under a looser majorant and
  with albedo 0, E[L] = L_env exp(-sigma rho d), so the gradient summed over
  every voxel must average to -sigma d L_env exp(-sigma rho d);
- gradients of triangles and voxels that no path reaches are zero.

A wrong log-derivative (real vs null, a missing density factor), a wrong
"radiance still to come" (an off-by-one vertex, NEE counted before or after
its vertex) or a gradient credited to the wrong triangle or voxel fails one
of these."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from stratum2_amd import abi, nanovdb, scenes

FLAGS = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_LAMBERTIAN | abi.ST_FLAG_FLOAT_NEE
NO_RR = 255  # roulette from path length 255: never within maxPathLength


def _params(W, H, spp, depth, seed=3, flags=FLAGS, rr=NO_RR):
    return abi.default_params(W, H, max_depth=depth, spp=spp, flags=flags, rr_start_depth=rr, seed=seed)


def _loss(o, view, p, dh):
    hdr = o.render(view, p, want_path_length=False)["hdr"].astype(np.float64)
    return float((hdr * dh).sum())


def test_albedo_gradient_equals_fixed_seed_central_difference():
    rng = np.random.default_rng(0)
    o = OracleScene(scenes.cornell_box(0))
    T = o.triangle_count()
    alb = rng.uniform(0.2, 0.9, (T, 3)).astype(np.float32)
    o.set_triangle_albedo(alb)
    W = H = 24
    view, p = scenes.cornell_view(W, H), _params(W, H, spp=2, depth=4)
    dh = rng.uniform(0.5, 1.5, (H, W, 4)).astype(np.float32)
    ga, _ = o.adjoint(view, p, dh)
    assert ga.shape == (T, 3) and np.isfinite(ga).all()
    flat = np.abs(ga).ravel()
    picks = np.argsort(flat)[::-1][:32]
    assert flat[picks[-1]] > 0
    eps = 1e-3
    for e in picks:
        t, ch = divmod(int(e), 3)
        lo, hi = alb.copy(), alb.copy()
        lo[t, ch] -= eps
        hi[t, ch] += eps
        o.set_triangle_albedo(hi)
        lp = _loss(o, view, p, dh)
        o.set_triangle_albedo(lo)
        lm = _loss(o, view, p, dh)
        fd = (lp - lm) / (float(hi[t, ch]) - float(lo[t, ch]))
        assert abs(fd - ga[t, ch]) <= 2e-3 * abs(ga[t, ch]) + 2e-4 * flat.max(), (t, ch, fd, ga[t, ch])
    o.set_triangle_albedo(alb)
    # the emitter's two triangles (instance 5) are never shaded as reflectors
    # facing the camera from behind; triangles the camera's paths never reach
    # at all get exactly zero
    assert (flat == 0).sum() > 0


def test_density_gradient_equals_weight_grid_central_difference():
    rng = np.random.default_rng(1)
    sd, offs = scenes.cornell_cloud(seed=0, n=16)
    o = OracleScene(sd)
    grid = sd.volumes[0]
    words = grid.nbytes // 4
    W = H = 24
    view, p = scenes.cornell_view(W, H), _params(W, H, spp=2, depth=4)
    dh = rng.uniform(0.5, 1.5, (H, W, 4)).astype(np.float32)
    _, gg = o.adjoint(view, p, dh, albedo=False, grid_vol=0, grid_words=words)
    assert np.isfinite(gg).all()
    slot = offs.ravel() // 4
    gv = gg[slot]
    # nothing but voxel values is read inside the grid's box
    other = np.ones(words, bool)
    other[slot] = False
    assert np.all(gg[other] == 0)
    vals = np.frombuffer(grid.tobytes(), "<f4")
    vmax = vals[nanovdb.ROOT_MAX_OFFSET // 4]
    cand = np.where(vals[slot] < vmax - 2e-3)[0]
    picks = cand[np.argsort(np.abs(gv[cand]))[::-1][:32]]
    assert np.abs(gv[picks[-1]]) > 0
    eps = 1e-3
    for v in picks:
        fd = []
        for sgn in (1, -1):
            wg = grid.copy()
            f = wg.view("<f4")
            f[slot[v]] += sgn * eps
            o.set_weight_grid(0, wg)
            fd.append(_loss(o, view, p, dh))
        o.set_weight_grid(0, None)
        d = (fd[0] - fd[1]) / (2 * eps)
        assert abs(d - gv[v]) <= 5e-3 * abs(gv[v]) + 1e-3 * np.abs(gv).max(), (v, d, gv[v])


def test_slab_gradient_is_unbiased_against_the_closed_form():
    # a 4 x 4 x 4 grid of value rho = 0.5 whose root maximum (the majorant)
    # is 1.0, filling the box |x|, |y| < 2, z in [-1.5, -0.5]; albedo 0 (a
    # real collision ends the path); a white environment seen along -z
    # through 1.0 of medium by a camera of 1 degree field of view (every ray
    # within 0.71 degrees of the axis: d within 1e-4 of 1).
    # E[L] = exp(-sigma * rho * d)
    rho, sigma, d = 0.5, 1.5, 1.0
    grid, offs = nanovdb.write_float_grid(np.full((4, 4, 4), rho, np.float32), all_leaves=True, return_offsets=True)
    grid.view("<f4")[nanovdb.ROOT_MAX_OFFSET // 4] = 1.0
    sd = scenes.SceneData()
    scenes.add_medium_box(sd, (sigma,) * 3, (0, 0, 0), 0.0, (-2, -2, -1.5), (2, 2, -0.5), grid=grid)
    sd.set_environment(emission=(1, 1, 1))
    o = OracleScene(sd)
    W = H = 16
    view = scenes.make_view(W, H, position=(0, 0, 0), fovy_deg=1.0)
    words = grid.nbytes // 4
    dh = np.ones((H, W, 4), np.float32)
    means, grads = [], []
    for seed in range(12):
        p = _params(W, H, spp=64, depth=3, seed=1000 * seed, flags=abi.ST_FLAG_DEFAULT | abi.ST_FLAG_FLOAT_NEE)
        hdr = o.render(view, p, want_path_length=False)["hdr"]
        means.append(hdr[..., 0].mean())
        _, gg = o.adjoint(view, p, dh, albedo=False, grid_vol=0, grid_words=words)
        grads.append(gg[offs.ravel() // 4].sum() / (3 * W * H))  # per pixel and channel
    e = np.exp(-sigma * rho * d)
    m, g = np.array(means), np.array(grads)
    assert abs(m.mean() - e) < 4 * m.std(ddof=1) / np.sqrt(len(m)) + 1e-4, (m.mean(), e)
    want = -sigma * d * e
    assert abs(g.mean() - want) < 4 * g.std(ddof=1) / np.sqrt(len(g)) + 1e-4, (g.mean(), want)
    assert abs(g.mean() - want) < 0.05 * abs(want)


def test_adjoint_refuses_what_it_does_not_replay():
    o = OracleScene(scenes.cornell_box(0))
    view, p = scenes.cornell_view(8, 8), _params(8, 8, spp=1, depth=3)
    dh = np.ones((8, 8, 4), np.float32)
    with pytest.raises(RuntimeError, match="per-triangle albedo"):
        o.adjoint(view, p, dh)  # no albedo table
    o.set_triangle_albedo(np.full((o.triangle_count(), 3), 0.5, np.float32))
    bad = _params(8, 8, spp=1, depth=3, flags=FLAGS | abi.ST_FLAG_RESTIR_DI)
    with pytest.raises(RuntimeError, match="plain NEE"):
        o.adjoint(view, bad, dh)
    with pytest.raises(RuntimeError, match="LAMBERTIAN"):
        o.render(view, _params(8, 8, spp=1, depth=3, flags=abi.ST_FLAG_DEFAULT))
