"""File formats: NanoVDB files against the reference's own NanoVDB IO
(compiled from the reference tree into oracle/_ref when present), Mitsuba
.vol conversion, OBJ meshes, glTF scenes, Radiance .hdr images and OpenEXR
images (against the reference's own tinyexr, compiled into oracle/_ref when
present; round trips and the renders they give)."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle.oracle import OracleScene, lib
from stratum2_amd import abi, assets, nanovdb, scenes

import gltf_files

REF_IO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref_nanovdb_io.so")


def _ref_io():
    if not os.path.exists(REF_IO):
        pytest.skip("oracle/_ref/libref_nanovdb_io.so not built (reference tree absent)")
    L = C.CDLL(REF_IO)
    L.ref_nvdb_read.argtypes = [C.c_char_p, C.c_ulonglong, C.c_void_p, C.c_ulonglong]
    L.ref_nvdb_read.restype = C.c_ulonglong
    L.ref_nvdb_write.argtypes = [C.c_char_p, C.c_void_p, C.c_ulonglong, C.c_int]
    return L


def _grid(seed=2):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.1, 1.0, (20, 33, 17)).astype(np.float32)
    d[rng.uniform(size=d.shape) < 0.5] = 0
    return nanovdb.write_float_grid(d, voxel_size=0.5, name="density")


@pytest.mark.parametrize("codec", [assets.CODEC_NONE, assets.CODEC_ZIP])
def test_nvdb_files_match_reference_io(tmp_path, codec):
    R = _ref_io()
    g = _grid()
    ours = str(tmp_path / "ours.nvdb")
    theirs = str(tmp_path / "theirs.nvdb")
    assets.write_nvdb(ours, [g], codec)
    assert R.ref_nvdb_write(theirs.encode(), g.ctypes.data, g.nbytes, codec) == 0
    assert open(ours, "rb").read() == open(theirs, "rb").read()  # identical files
    out = np.zeros(g.nbytes, np.uint8)
    assert R.ref_nvdb_read(ours.encode(), 0, out.ctypes.data, out.nbytes) == g.nbytes
    assert np.array_equal(out, g)
    assert np.array_equal(assets.read_nvdb(theirs), g)


def test_nvdb_multi_grid_round_trip(tmp_path):
    grids = [_grid(1), _grid(2), _grid(3)]
    p = str(tmp_path / "three.nvdb")
    assets.write_nvdb(p, grids, assets.CODEC_ZIP)
    for i, g in enumerate(grids):
        assert np.array_equal(assets.read_nvdb(p, i), g)
    with pytest.raises(IndexError):
        assets.read_nvdb(p, 3)


def test_vol_to_grid_voxel_order(tmp_path):
    rng = np.random.default_rng(4)
    d = rng.uniform(0, 1, (9, 6, 13)).astype(np.float32)
    d[d < 0.3] = 0
    p = str(tmp_path / "v.vol")
    assets.write_vol(p, d, (0, 0, 0), (1, 2, 3))
    back, pmin, pmax = assets.read_vol(p)
    assert np.array_equal(back, d) and pmax == (1, 2, 3)
    raw = np.fromfile(p, "<f4", offset=48)
    assert raw[(2 * 6 + 3) * 9 + 7] == d[7, 3, 2]  # data[(k * yres + j) * xres + i]
    g = assets.vol_to_grid(p)
    L = lib()
    for i, j, k in [(0, 0, 0), (7, 3, 2), (8, 5, 12), (4, 4, 4)]:
        assert L.or_nvdb_read(g.ctypes.data, i, j, k) == d[i, j, k]


def test_obj_scene_renders_like_the_generated_scene(tmp_path):
    """The Cornell icosphere written to OBJ and read back renders identically
    (same vertices, same triangles, same material) through the oracle."""
    v, f = scenes.icosphere(2)
    v = (v * 0.5 + np.array([0, 1, 0])).astype(np.float32)
    mat = scenes.material_record(base_color=(0.6, 0.5, 0.4), specular=0.0, emission=(0, 0, 0))
    a = scenes.SceneData()
    vi, n = a.add_mesh(v, f.reshape(-1).astype(np.uint32))
    a.add_mesh_instance(vi, n, a.add_material(mat))
    a.set_environment(emission=(1.0, 0.9, 0.8))
    p = str(tmp_path / "ico.obj")
    assets.write_obj(p, v, f)
    with open(str(tmp_path / "ico.mtl"), "w") as fm:
        fm.write("newmtl m\nKd 0.6 0.5 0.4\n")
    txt = open(p).read()
    open(p, "w").write("mtllib ico.mtl\nusemtl m\n" + txt)
    b = assets.read_obj(p)
    b.set_environment(emission=(1.0, 0.9, 0.8))
    view = scenes.make_view(32, 32, position=(0, 1, 2.0))
    params = abi.default_params(32, 32, max_depth=3, spp=2)
    ra = OracleScene(a).render(view, params)
    rb = OracleScene(b).render(view, params)
    assert np.array_equal(ra["hdr"], rb["hdr"])


def test_hdr_round_trip_and_rle(tmp_path):
    rng = np.random.default_rng(5)
    img = np.ones((7, 40, 4), np.float32)
    img[..., :3] = rng.lognormal(0, 2, (7, 40, 3))
    p = str(tmp_path / "a.hdr")
    assets.write_hdr(p, img)
    back = assets.read_hdr(p)
    # 8-bit mantissas share the pixel's exponent: error below max(rgb) / 128
    err = np.abs(back[..., :3] - img[..., :3]).max(-1)
    assert np.all(err <= img[..., :3].max(-1) / 128)
    # the same pixels in run-length scanlines decode identically
    rgbe = assets.float_to_rgbe(img)
    q = str(tmp_path / "rle.hdr")
    with open(q, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 7 +X 40\n")
        for y in range(7):
            f.write(bytes([2, 2, 0, 40]))
            for c in range(4):
                row = rgbe[y, :, c]
                x = 0
                while x < 40:
                    run = 1
                    while x + run < 40 and run < 127 and row[x + run] == row[x]:
                        run += 1
                    if run >= 3:
                        f.write(bytes([128 + run, row[x]]))
                        x += run
                    else:
                        n = min(128, 40 - x)
                        f.write(bytes([n]) + bytes(row[x:x + n]))
                        x += n
    assert np.array_equal(assets.read_hdr(q), back)


# ------------------------------------------------------------------ glTF
def _quat_matrix(axis, angle):
    """Rodrigues' rotation formula."""
    k = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def test_gltf_transforms_materials_and_vertices(tmp_path):
    doc, data, src = gltf_files.gltf_content()
    p = str(tmp_path / "s.gltf")
    gltf_files.write_gltf(p, doc, data)
    sd = assets.read_gltf(p)
    # node 0 -> quad, node 1 -> two icosphere primitives, node 2 -> quad
    assert len(sd.instances) == 4
    T0 = np.zeros((3, 4))
    T0[:, :3] = _quat_matrix([0, 0, 1], np.pi / 4) @ np.diag([2.0, 1.0, 0.5])
    T0[:, 3] = [1, 2, 3]
    C = np.array([[0, 0, 2, 0.5], [0, 2, 0, -1], [-2, 0, 0, 3]], np.float64)   # column-major matrix, transposed
    T1 = np.zeros((3, 4))
    T1[:, :3] = T0[:, :3] @ C[:, :3]
    T1[:, 3] = T0[:, :3] @ C[:, 3] + T0[:, 3]
    np.testing.assert_allclose(sd.transforms[0], T0, atol=2e-7)
    np.testing.assert_allclose(sd.transforms[1], T1, atol=2e-6)
    np.testing.assert_allclose(sd.transforms[2], T1, atol=2e-6)
    assert np.array_equal(sd.transforms[3], np.eye(3, 4, dtype=np.float32))
    # materials: factors through Scene::makeMetallicRoughnessMaterial and the extensions
    lum = lambda c: 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]  # noqa: E731
    m0 = scenes.material_record(base_color=(0.9, 0.5, 0.25), specular=lum([0.2, 0.6, 1.0]), emission=(4.0, 2.0, 0.0),
                                metallic=0.25, roughness=0.6, clearcoat=0.4, transmission=0.65, eta=1.33)
    m1 = scenes.material_record(base_color=(0.3, 0.3, 0.8), specular=0.35, metallic=1.0, roughness=1.0)
    m2 = scenes.material_record(base_color=(1, 1, 1), specular=1.0, metallic=1.0, roughness=1.0)
    assert [np.array_equal(a, b) for a, b in zip(sd.material_words, [m0, m1, m2])] == [True] * 3
    addr = [(a >> 4) for a, _ in sd.instances]
    assert addr == [0, 64, 128, 0]
    # vertex data round trip: the quad keeps normals and texture coordinates, indices widen to u32
    qp, qn, quv, qi = src["quad"]
    vi0, n0 = sd.instances[0][1] & 0xFFFF, sd.instances[0][1] >> 16
    assert n0 == 2
    v = sd.vertex_infos[vi0]
    ib, pb = v.packed_buffer_indices[0] & 0xFFFF, v.packed_buffer_indices[0] >> 16
    nb, tb = v.packed_buffer_indices[1] & 0xFFFF, v.packed_buffer_indices[1] >> 16
    assert np.array_equal(np.frombuffer(sd.buffers[ib], np.uint32), qi.astype(np.uint32))
    assert np.array_equal(np.frombuffer(sd.buffers[pb], np.float32).reshape(-1, 3), qp)
    assert np.array_equal(np.frombuffer(sd.buffers[nb], np.float32).reshape(-1, 3), qn)
    assert np.array_equal(np.frombuffer(sd.buffers[tb], np.float32).reshape(-1, 2), quv)
    assert v.packed_strides == (4 | 12 << 8 | 12 << 16 | 8 << 24)
    vi1 = sd.instances[1][1] & 0xFFFF
    assert sd.vertex_infos[vi1].packed_buffer_indices[1] == (scenes.NO_BUFFER | scenes.NO_BUFFER << 16)


def test_gltf_and_glb_read_the_same_scene(tmp_path):
    doc, data, _ = gltf_files.gltf_content(3)
    a, b = str(tmp_path / "s.gltf"), str(tmp_path / "s.glb")
    gltf_files.write_gltf(a, doc, data)
    gltf_files.write_glb(b, doc, data)
    # external .bin buffer as a third container
    import json
    d = json.loads(json.dumps(doc))
    d["buffers"] = [{"byteLength": len(data), "uri": "s.bin"}]
    open(str(tmp_path / "s.bin"), "wb").write(data)
    c = str(tmp_path / "ext.gltf")
    open(c, "w").write(json.dumps(d))
    A, B, Cs = assets.read_gltf(a), assets.read_gltf(b), assets.read_gltf(c)
    for X in (B, Cs):
        assert X.instances == A.instances
        assert X.buffers == A.buffers
        assert all(np.array_equal(x, y) for x, y in zip(X.transforms, A.transforms))
        assert all(np.array_equal(x, y) for x, y in zip(X.material_words, A.material_words))


def test_gltf_scene_renders_like_the_generated_scene(tmp_path):
    """A glTF holding the Cornell icosphere (translated by a node) renders
    bit-identically to the same scene built directly."""
    v, f = scenes.icosphere(2)
    v = (v * 0.5).astype(np.float32)
    i = f.reshape(-1).astype(np.uint32)
    mat = scenes.material_record(base_color=(0.6, 0.5, 0.4), specular=1.0, metallic=0.0, roughness=0.3)
    a = scenes.SceneData()
    vi, n = a.add_mesh(v, i)
    a.add_mesh_instance(vi, n, a.add_material(mat), np.array([[1, 0, 0, 0], [0, 1, 0, 1], [0, 0, 1, 0]], np.float32))
    a.set_environment(emission=(1.0, 0.9, 0.8))
    doc = {"asset": {"version": "2.0"},
           "buffers": [{"byteLength": v.nbytes + i.nbytes, "uri": "ico.bin"}],
           "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": v.nbytes},
                           {"buffer": 0, "byteOffset": v.nbytes, "byteLength": i.nbytes}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "type": "VEC3", "count": len(v)},
                         {"bufferView": 1, "componentType": 5125, "type": "SCALAR", "count": len(i)}],
           "materials": [{"pbrMetallicRoughness": {"baseColorFactor": [0.6, 0.5, 0.4, 1], "metallicFactor": 0.0,
                                                   "roughnessFactor": 0.3}}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1, "material": 0}]}],
           "nodes": [{"translation": [0, 1, 0], "mesh": 0}]}
    import json
    open(str(tmp_path / "ico.bin"), "wb").write(v.tobytes() + i.tobytes())
    p = str(tmp_path / "ico.gltf")
    open(p, "w").write(json.dumps(doc))
    b = assets.read_gltf(p)
    b.set_environment(emission=(1.0, 0.9, 0.8))
    view = scenes.make_view(32, 32, position=(0, 1, 2.0))
    params = abi.default_params(32, 32, max_depth=3, spp=2)
    assert np.array_equal(OracleScene(a).render(view, params)["hdr"], OracleScene(b).render(view, params)["hdr"])


def test_gltf_rejects_what_the_renderer_cannot_take(tmp_path):
    import json
    doc, data, _ = gltf_files.gltf_content()
    cases = []
    d = json.loads(json.dumps(doc))
    d["meshes"][0]["primitives"][0]["mode"] = 5   # triangle strip
    cases.append(d)
    d = json.loads(json.dumps(doc))
    d["extensions"] = {"KHR_lights_punctual": {"lights": [{"type": "point", "extras": {"radius": 0.1}}]}}
    d["nodes"][2]["extensions"] = {"KHR_lights_punctual": {"light": 0}}
    cases.append(d)
    d = json.loads(json.dumps(doc))
    d["accessors"][0]["count"] = 1 << 20          # reads past the buffer
    cases.append(d)
    for k, d in enumerate(cases):
        p = str(tmp_path / f"bad{k}.gltf")
        gltf_files.write_gltf(p, d, data)
        with pytest.raises(ValueError):
            assets.read_gltf(p)


# ------------------------------------------------------------------- EXR
REF_EXR = os.path.join(os.path.dirname(REF_IO), "libref_tinyexr.so")


def _ref_exr():
    if not os.path.exists(REF_EXR):
        pytest.skip("oracle/_ref/libref_tinyexr.so not built (reference tree absent)")
    L = C.CDLL(REF_EXR)
    L.ref_exr_load.argtypes = [C.c_char_p, C.c_void_p, C.c_ulonglong, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.ref_exr_save.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_int, C.c_int]
    return L


def _ref_load(L, path, cap=1 << 22):
    out = np.zeros(cap, np.float32)
    w, h = C.c_int(), C.c_int()
    assert L.ref_exr_load(path.encode(), out.ctypes.data, cap, C.byref(w), C.byref(h)) == 0
    return out[:4 * w.value * h.value].reshape(h.value, w.value, 4)


@pytest.mark.parametrize("names", ["ABGR", "BGR", "Y", "ABGRZ"])
@pytest.mark.parametrize("comp", [assets.EXR_NONE, assets.EXR_RLE, assets.EXR_ZIPS, assets.EXR_ZIP])
@pytest.mark.parametrize("half", [False, True])
def test_exr_files_written_by_the_reference_loader_library(tmp_path, names, comp, half):
    """Files written by tinyexr (the reference's own EXR library) with each
    compression read the same through assets.read_exr and through LoadEXR."""
    L = _ref_exr()
    rng = np.random.default_rng(len(names) * 10 + comp)
    H, W = 45, 70   # several 16-line ZIP blocks, the last one short
    planes = rng.gamma(0.5, 3.0, (len(names), H, W)).astype(np.float32)
    planes[:, 10:20, 5:40] = planes[:, 10, 5][:, None, None]   # runs for RLE
    p = str(tmp_path / f"t_{names}_{comp}_{half}.exr")
    assert L.ref_exr_save(p.encode(), planes.ctypes.data, W, H, names.encode(), comp, 1 if half else 2) == 0
    mine, ref = assets.read_exr(p), _ref_load(L, p)
    assert mine.shape == (H, W, 4)
    np.testing.assert_array_equal(mine, ref)


def test_exr_written_here_loads_in_the_reference_library(tmp_path):
    L = _ref_exr()
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 50, (33, 41, 4)).astype(np.float32)
    for comp in (assets.EXR_NONE, assets.EXR_ZIP):
        for half in (False, True):
            p = str(tmp_path / f"w{comp}{half}.exr")
            assets.write_exr(p, img, compression=comp, half=half)
            want = img.astype(np.float16).astype(np.float32) if half else img
            np.testing.assert_array_equal(_ref_load(L, p), want)
            np.testing.assert_array_equal(assets.read_exr(p), want)


def test_exr_round_trip_rgb_and_rejections(tmp_path):
    rng = np.random.default_rng(4)
    rgb = rng.uniform(0, 5, (20, 17, 3)).astype(np.float32)
    p = str(tmp_path / "rgb.exr")
    assets.write_exr(p, rgb)
    r = assets.read_exr(p)
    np.testing.assert_array_equal(r[..., :3], rgb)
    assert np.all(r[..., 3] == 1.0)
    data = bytearray(open(p, "rb").read())
    data[5] |= 0x02          # version flags: tiled
    open(str(tmp_path / "tiled.exr"), "wb").write(bytes(data))
    with pytest.raises(ValueError):
        assets.read_exr(str(tmp_path / "tiled.exr"))
    with pytest.raises(ValueError):
        assets.write_exr(str(tmp_path / "x.exr"), rgb, compression=assets.EXR_PIZ)


def test_exr_environment_renders_like_the_array(tmp_path):
    """An environment map read back from .exr renders like the array it was
    written from (full float precision)."""
    env = scenes.sky_environment(64, 32)
    p = str(tmp_path / "sky.exr")
    assets.write_exr(p, env)
    img = assets.read_exr(p)
    np.testing.assert_array_equal(img[..., :3], env[..., :3])
