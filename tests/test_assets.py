"""File formats: NanoVDB files against the reference's own NanoVDB IO
(compiled from the reference tree into oracle/_ref when present), Mitsuba
.vol conversion, OBJ meshes and Radiance .hdr images (round trips and the
renders they give)."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle.oracle import OracleScene, lib
from stratum2_amd import abi, assets, nanovdb, scenes

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
