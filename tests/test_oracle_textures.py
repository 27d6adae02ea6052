"""Material images in the oracle (LoadMaterial with its image indices,
material.hlsli:22-45; SampleImage4, scene.hlsli:101-109, with the static
sampler's linear filtering and repeat addressing, TestRenderer.cpp:54-57),
pinned by what the definitions fix:

- an image of one constant colour c equals the untextured material whose
  base colour is the 8-bit rounding of base x c, bit for bit over an image;
- at a texel centre of level k with a screen size that selects exactly lod
  k, the sample is the 2^k x 2^k box average of the base level (the mip
  chain, the lod and the addressing together), and halfway between two
  levels it is their mean;
- repeat addressing: u and u + 1 sample the same value;
- the packed-word modulation turns roughness through smoothness
  (1 - (1 - r)(1 - t)) and scales the other channels.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from stratum2_amd import abi, scenes


def test_constant_image_equals_requantised_base_colour():
    c = np.array([0.6, 0.3, 0.9, 1.0], np.float32)
    base = np.array([0.8, 0.7, 0.5])
    a = scenes.cornell_box(0)
    b = scenes.cornell_box(0)
    a.images.append(np.tile(c, (8, 8, 1)))
    for sd, textured in ((a, True), (b, False)):
        w = scenes.material_record(base_color=tuple(base), specular=0.3, roughness=0.6)
        if textured:
            w[4] = len(sd.images) - 1
        else:
            stored = np.array([(int(w[0]) >> (8 * k)) & 0xFF for k in range(3)], np.float32) / np.float32(255)
            q = np.floor(np.clip(stored * c[:3], 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint32)
            w[0] = (w[0] & 0xFF000000) | q[0] | (q[1] << 8) | (q[2] << 16)
        p, i = scenes.quad((-0.9, 0.003, -0.9), (0.9, 0.003, -0.9), (0.9, 0.003, 0.9), (-0.9, 0.003, 0.9))
        f = scenes._tri_normal_ok(p, i.reshape(-1, 3), (0, 1, 0))
        vi, n = sd.add_mesh(p, f.reshape(-1), texcoords=np.array([[0, 0], [3, 0], [3, 3], [0, 3]], np.float32))
        sd.add_mesh_instance(vi, n, sd.add_material(w))
    W = H = 32
    view = scenes.cornell_view(W, H)
    p = abi.default_params(W, H, max_depth=4, spp=2)
    ra = OracleScene(a).render(view, p)
    rb = OracleScene(b).render(view, p)
    assert np.array_equal(ra["hdr"], rb["hdr"]) and np.array_equal(ra["path_length"], rb["path_length"])


def _texture_probe(img, u, v, uvs):
    """The oracle's SampleImage4 of img at (u, v) with screen size uvs, through
    the base colour of a white material on a quad (uv = (u, v) everywhere is
    not possible; instead the sampler is reached through or_tex_sample)."""
    from oracle.oracle import lib
    import ctypes as C
    L = lib()
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros(4, np.float32)
    L.or_tex_sample.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_void_p]
    L.or_tex_sample(img.ctypes.data, img.shape[1], img.shape[0], u, v, uvs, out.ctypes.data)
    return out


def test_mip_levels_lod_and_addressing():
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, (32, 64, 4)).astype(np.float32)
    H, W = img.shape[:2]
    for k in range(0, 5):
        s = 2 ** k
        lvl = img.reshape(H // s, s, W // s, s, 4).astype(np.float64).mean((1, 3))
        for (tx, ty) in ((0, 0), (3, 1), (W // s - 1, H // s - 1)):
            u, v = (tx + 0.5) / (W // s), (ty + 0.5) / (H // s)
            got = _texture_probe(img, u, v, s / W)  # lod = log2(uvs * max(w, h)) = k
            np.testing.assert_allclose(got, lvl[ty, tx], rtol=2e-5, atol=2e-6)
            np.testing.assert_allclose(_texture_probe(img, u + 1, v - 2, s / W), got, rtol=2e-6, atol=1e-7)
    # halfway between levels 1 and 2: the mean of the two levels' texel-centred samples
    l1 = img.reshape(H // 2, 2, W // 2, 2, 4).astype(np.float64).mean((1, 3))
    l2 = img.reshape(H // 4, 4, W // 4, 4, 4).astype(np.float64).mean((1, 3))
    u, v = (2 + 0.5) / (W // 4), (1 + 0.5) / (H // 4)  # a level-2 texel centre: four level-1 texels meet there
    want2 = l2[1, 2]
    want1 = l1[2:4, 4:6].mean((0, 1))
    got = _texture_probe(img, u, v, 2 ** 1.5 / W)
    np.testing.assert_allclose(got, 0.5 * (want1 + want2), rtol=2e-4, atol=2e-5)


def test_textured_scene_renders_and_differs():
    sd = scenes.cornell_textured(0)
    W = H = 48
    view = scenes.cornell_view(W, H)
    r = OracleScene(sd).render(view, abi.default_params(W, H, max_depth=4, spp=4))
    assert np.isfinite(r["hdr"]).all()
    sd2 = scenes.cornell_textured(0)
    for k, w in enumerate(sd2.material_words):
        w = np.array(w, np.uint32)
        if len(w) >= 8:
            w[4:8] = abi.ST_NO_IMAGE
        sd2.material_words[k] = w
    r2 = OracleScene(sd2).render(view, abi.default_params(W, H, max_depth=4, spp=4))
    assert not np.array_equal(r["hdr"], r2["hdr"])


def test_refused_images():
    """Emission images and normal maps (alpha masks are taken: their tests
    are the alpha-mask ones)."""
    for word in (5, 9):
        sd = scenes.cornell_box(0)
        sd.images.append(np.ones((4, 4, 4), np.float32))
        w = scenes.material_record()
        w[word] = len(sd.images) - 1
        vi, n = sd.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([0, 1, 2], np.uint32))
        sd.add_mesh_instance(vi, n, sd.add_material(w))
        with pytest.raises(RuntimeError):
            OracleScene(sd)
