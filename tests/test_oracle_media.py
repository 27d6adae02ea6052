"""CPU checks that pin the oracle's participating-media code: the NanoVDB
reader against the reference's own PNanoVDB.h (compiled in C mode when the
reference tree is present) and against the dense array the grid was written
from; the Henyey-Greenstein sampler against its pdf; delta tracking, ratio
tracking and the medium-vertex MIS against closed forms (Beer-Lambert chords,
the E3 exponential integral of a lit slab, a scattering furnace)."""
import ctypes as C
import math
import os

import numpy as np
import pytest
from scipy.special import expn

from oracle.oracle import OracleScene, lib
from stratum2_amd import abi, nanovdb, scenes

REF_SO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref_nanovdb.so")


def _ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libref_nanovdb.so not built (reference tree absent)")
    L = C.CDLL(REF_SO)
    L.ref_read_float.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_int]
    L.ref_read_float.restype = C.c_float
    L.ref_root_info.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
    L.ref_world_to_index.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int, C.c_void_p]
    return L


def _sparse_dense(seed=3, shape=(40, 24, 150)):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.1, 2.0, shape).astype(np.float32)
    d[rng.uniform(size=shape) < 0.7] = 0
    d[:, :, 20:60] = 0  # whole empty leaves and an empty lower-node column
    return d


def _read_all(grid, pts):
    L = lib()
    return np.array([L.or_nvdb_read(grid.ctypes.data, int(i), int(j), int(k)) for i, j, k in pts], np.float32)


def test_nanovdb_reader_returns_the_dense_values():
    d = _sparse_dense()
    g = nanovdb.write_float_grid(d, voxel_size=0.37)
    rng = np.random.default_rng(0)
    inside = np.argwhere(np.ones(d.shape, bool))
    pts = inside[rng.choice(len(inside), 6000, replace=False)]
    assert np.array_equal(_read_all(g, pts), d[tuple(pts.T)])
    # outside the written block (other leaves, other lower nodes, negative coordinates): background
    out = np.array([[-1, 0, 0], [0, -5, 3], [40, 0, 0], [0, 24, 0], [200, 7, 9], [5000, 1, 1], [-4097, 2, 2]])
    assert np.all(_read_all(g, out) == 0)


def test_nanovdb_reader_matches_reference_pnanovdb():
    R = _ref_lib()
    d = _sparse_dense(seed=5)
    g = nanovdb.write_float_grid(d, voxel_size=0.37)
    rng = np.random.default_rng(1)
    pts = np.concatenate([rng.integers(-10, 160, (4000, 3)), np.argwhere(d > 0)[:2000]])
    ours = _read_all(g, pts)
    ref = np.array([R.ref_read_float(g.ctypes.data, g.nbytes, int(i), int(j), int(k)) for i, j, k in pts], np.float32)
    assert np.array_equal(ours, ref)
    mx = np.zeros(1, np.float32)
    bb = np.zeros(6, np.int32)
    R.ref_root_info(g.ctypes.data, g.nbytes, mx.ctypes.data, bb.ctypes.data)
    act = np.argwhere(d != 0)
    assert mx[0] == d.max()
    assert bb.tolist() == act.min(0).tolist() + act.max(0).tolist()
    L = lib()
    for _ in range(200):
        p = rng.normal(size=3).astype(np.float32) * 10
        for dirflag in (0, 1):
            a = np.zeros(3, np.float32)
            b = np.zeros(3, np.float32)
            L.or_nvdb_world_to_index(g.ctypes.data, p.ctypes.data, dirflag, a.ctypes.data)
            R.ref_world_to_index(g.ctypes.data, g.nbytes, p.ctypes.data, dirflag, b.ctypes.data)
            assert np.array_equal(a, b)


@pytest.mark.parametrize("g", [-0.6, 0.0, 0.3, 0.85])
def test_henyey_greenstein_sampling_matches_its_pdf(g):
    """Histogram of cos(theta) of sampled directions vs 2*pi*p(cos); the pdf
    integrates to 1 and has mean cosine g about the propagation direction
    -dirIn (which fixes the HG formula and its sign convention)."""
    L = lib()
    rng = np.random.default_rng(7)
    din = np.array([0.36, -0.48, 0.8], np.float32)
    n = 60000
    cos = np.empty(n)
    out = np.zeros(3, np.float32)
    for i, (r0, r1) in enumerate(rng.uniform(size=(n, 2)).astype(np.float32)):
        L.or_hg_sample(g, float(r0), float(r1), din.ctypes.data, out.ctypes.data)
        assert abs(np.linalg.norm(out) - 1) < 1e-5
        cos[i] = float(out @ din)
    edges = np.linspace(-1, 1, 41)
    hist, _ = np.histogram(cos, edges)
    mu = np.linspace(-1, 1, 4001)
    p = np.array([L.or_hg_eval(g, din.ctypes.data, (din * np.float32(m) + np.float32(math.sqrt(max(0, 1 - m * m))) *
                                                   np.array([0.8, 0.6, 0.0], np.float32)).astype(np.float32).ctypes.data)
                  for m in mu])
    dens = 2 * math.pi * p
    assert abs(np.trapezoid(dens, mu) - 1) < 2e-3
    assert abs(np.trapezoid(dens * mu, mu) + g) < 2e-3
    expect = np.array([np.trapezoid(dens[(mu >= a) & (mu <= b)], mu[(mu >= a) & (mu <= b)]) for a, b in zip(edges[:-1], edges[1:])]) * n
    ok = expect > 50
    chi2 = (((hist - expect) ** 2) / expect)[ok].sum()
    assert chi2 < 2.0 * ok.sum()


# ----------------------------------------------------------------- scenes
def _camera_dirs(view, W, H):
    """Pixel-centre camera directions (float64) of make_view views."""
    px, py = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    u = (px - view.image_min[0]) / (view.image_max[0] - view.image_min[0])
    v = (py - view.image_min[1]) / (view.image_max[1] - view.image_min[1])
    cx, cy = 2 * u - 1, -(2 * v - 1)
    pr = view.projection
    sg = np.sign(pr.near_plane)
    bp = np.stack([pr.near_plane * (cx * sg - pr.offset[0]) / pr.scale[0],
                   pr.near_plane * (cy * sg - pr.offset[1]) / pr.scale[1],
                   np.full_like(cx, pr.near_plane)], -1)
    bp /= np.linalg.norm(bp, axis=-1, keepdims=True)
    R = np.array([[view.camera_to_world.m[r][c] for c in range(3)] for r in range(3)])
    d = bp @ R.T
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def _box_chord(o, d, lo, hi):
    with np.errstate(divide="ignore", invalid="ignore"):
        t0, t1 = (lo - o) / d, (hi - o) / d
    tn = np.minimum(t0, t1).max(-1)
    tf = np.maximum(t0, t1).min(-1)
    return np.clip(tf - np.maximum(tn, 0), 0, None)


@pytest.mark.parametrize("use_grid", [False, True])
def test_absorbing_medium_chord_transmittance(use_grid):
    """Camera rays through an absorbing box under a constant environment:
    E[pixel] = exp(-sigma_c * chord) per channel (Beer-Lambert), with a
    chromatic sigma so the spectral-MIS average of deltaTrack's pdfs matters."""
    sd = scenes.SceneData()
    sd.set_environment(emission=(1.0, 1.0, 1.0))
    sigma = np.array([0.6, 1.5, 3.0])
    lo, hi = np.array([-0.5, -0.4, -0.6]), np.array([0.5, 0.4, 0.6])
    scenes.add_medium_box(sd, tuple(sigma), (0.0, 0.0, 0.0), 0.0, lo, hi, scenes.layered_grid([1.0] * 4) if use_grid else None)
    W = H = 24
    view = scenes.make_view(W, H, position=(0.0, 0.0, 1.4))
    p = abi.default_params(W, H, max_depth=4, spp=384)
    out = OracleScene(sd).render(view, p, want_path_length=False)
    d = _camera_dirs(view, W, H)
    chord = _box_chord(np.array([0.0, 0.0, 1.4]), d, lo, hi)
    expect = np.exp(-chord[..., None] * sigma)
    got = out["hdr"][..., :3].astype(np.float64)
    assert (chord > 0).sum() > 100
    # one sample is exp(-sigma d) / exp(-sigma_c d) with probability exp(-sigma_c d) / 3
    # for each hero channel c: heavy-tailed per pixel, so compare the image means
    m = chord > 0
    rel = got[m].mean(0) / expect[m].mean(0) - 1
    assert np.abs(rel).max() < 0.02, rel
    assert np.allclose(got[chord == 0], 1.0)


def _floor(sd, half=50.0, albedo=0.8):
    pos = np.array([[-half, 0, -half], [half, 0, -half], [half, 0, half], [-half, 0, half]], np.float32)
    idx = np.array([0, 2, 1, 0, 3, 2], np.uint32)
    v = pos[idx.reshape(2, 3)]
    n = np.cross(v[:, 0] - v[:, 2], v[:, 1] - v[:, 2])
    assert np.all(n[:, 1] > 0)  # geometric normal (shading_data.hlsli:142-151) points up
    vi, nt = sd.add_mesh(pos, idx)
    mat = sd.add_material(scenes.material_record(base_color=(albedo,) * 3))
    sd.add_mesh_instance(vi, nt, mat)


@pytest.mark.parametrize("use_grid", [False, True])
def test_lit_floor_under_absorbing_slab(use_grid):
    """Lambert floor (albedo a) under an absorbing slab of optical thickness
    tau, constant unit environment, direct illumination only:
    L = 2 a E3(tau) * exp(-tau / mu_cam).  With the grid the slab is two
    layers of densities 1 and 0.25 under one majorant, so NEE goes through
    ratio tracking with real null-collision weights (tau = 0.625 sigma h)."""
    sd = scenes.SceneData()
    sd.set_environment(emission=(1.0, 1.0, 1.0))
    _floor(sd)
    h = 0.2
    sigma = np.array([1.0, 2.5, 5.0])
    grid = scenes.layered_grid([0.25, 1.0]) if use_grid else None
    scenes.add_medium_box(sd, tuple(sigma), (0.0, 0.0, 0.0), 0.0, (-50, 0.2, -50), (50, 0.2 + h, 50), grid)
    tau = sigma * h * (0.625 if use_grid else 1.0)
    W = H = 16
    view = scenes.make_view(W, H, position=(0.0, 5.0, 0.0), orientation=scenes.axis_angle_quat((1, 0, 0), 90.0))
    p = abi.default_params(W, H, max_depth=2, spp=512,
                           flags=abi.ST_FLAG_DEFAULT | abi.ST_FLAG_LAMBERTIAN | abi.ST_FLAG_DIRECT_ONLY)
    out = OracleScene(sd).render(view, p, want_path_length=False)
    d = _camera_dirs(view, W, H)
    assert np.all(d[..., 1] < -0.7)  # looking down onto the floor
    mu_cam = -d[..., 1]
    a = scenes.unorm8(0.8) / 255.0
    expect = 2 * a * expn(3, tau)[None, None, :] * np.exp(-tau[None, None, :] / mu_cam[..., None])
    got = out["hdr"][..., :3].astype(np.float64)
    rel = got.mean((0, 1)) / expect.mean((0, 1)) - 1
    assert np.abs(rel).max() < 0.03, rel


@pytest.mark.parametrize("paths", ["direct", "full"])
def test_ris_in_media_keeps_expectation(paths):
    """RIS over 32 candidates at surface and medium vertices (ReSTIR DI
    without reuse) estimates the image of one-sample NEE: a Cornell box with
    a scattering block, paired over 8 seed ranges."""
    from scipy import stats
    sd = scenes.cornell_box(0)
    scenes.add_medium_box(sd, (0.8, 1.0, 1.3), (0.8, 0.8, 0.8), 0.3, (-0.55, 0.45, -0.7), (0.55, 1.35, 0.3))
    W = H = 32
    view = scenes.cornell_view(W, H)
    osc = OracleScene(sd)
    flags = abi.ST_FLAG_DEFAULT | (abi.ST_FLAG_DIRECT_ONLY if paths == "direct" else 0)
    means = {}
    for mode in ("nee", "ris"):
        fl = flags | (abi.ST_FLAG_RESTIR_DI if mode == "ris" else 0)
        means[mode] = np.array([osc.render(view, abi.default_params(W, H, max_depth=4, spp=16, seed=1000 * k, flags=fl),
                                           want_path_length=False)["hdr"][..., :3].mean() for k in range(8)])
    t = stats.ttest_rel(means["nee"], means["ris"])
    print(means["nee"].mean(), means["ris"].mean(), t.pvalue)
    assert t.pvalue > 1e-3


@pytest.mark.parametrize("g", [0.0, 0.3])
def test_scattering_medium_furnace(g):
    """A white (albedo 1) scattering medium inside a closed box whose walls
    emit Le and reflect nothing: every path ends on a wall, so L = Le for
    every pixel.  Exercises free-flight sampling, HG sampling, NEE from
    medium vertices through transmittance, and its MIS with emission hits."""
    sd = scenes.SceneData()
    scenes.emissive_inward_cube(sd, 1.0, (1.0, 1.0, 1.0))
    scenes.add_medium_box(sd, (1.5, 1.5, 1.5), (1.0, 1.0, 1.0), g, (-0.99,) * 3, (0.99,) * 3)
    W = H = 16
    view = scenes.make_view(W, H, position=(0.1, -0.2, 0.3))
    p = abi.default_params(W, H, max_depth=64, spp=128)
    out = OracleScene(sd).render(view, p)
    got = out["hdr"][..., :3].astype(np.float64)
    assert abs(got.mean() - 1) < 0.02, got.mean()
    assert out["path_length"].max() > 3  # paths did scatter
    # RIS at the medium and wall vertices keeps L = Le
    p = abi.default_params(W, H, max_depth=64, spp=64, flags=abi.ST_FLAG_DEFAULT | abi.ST_FLAG_RESTIR_DI, candidate_samples=8)
    got = OracleScene(sd).render(view, p, want_path_length=False)["hdr"][..., :3].astype(np.float64)
    assert abs(got.mean() - 1) < 0.02, got.mean()
