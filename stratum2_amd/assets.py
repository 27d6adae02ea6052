"""On-disk formats for scenes rendered through the C ABI.

* NanoVDB files (.nvdb): segments of a 16-byte header, per-grid 176-byte
  metadata + name, then the grid buffers, uncompressed or zlib ("ZIP")
  compressed -- the layout of extern/nanovdb/util/IO.h (Header, MetaData,
  Internal::read/write), which src/App/loaders/load_volumes.cpp:100-109 reads.
* Mitsuba volumes (.vol, version 3, float32, one channel), converted to a
  NanoVDB float grid as load_volumes.cpp:53-98 does (voxel (i, j, k) takes
  data[(k * yres + j) * xres + i]; values equal to the background 0 are
  inactive).
* Wavefront OBJ meshes -> SceneData (one mesh instance per object/group,
  polygons fanned into triangles, optional .mtl diffuse colours).
* Radiance RGBE images (.hdr) for environment maps, read with both scanline
  encodings (flat and run-length), written flat.
* OpenEXR images (.exr) for environment maps, read as src/Core/Image.cpp:57-72
  loads them (tinyexr's LoadEXR: RGBA float32), for single-part scanline files
  with NONE, RLE, ZIPS or ZIP compression and HALF or FLOAT channels; written
  uncompressed or ZIP.
* glTF 2.0 scenes (.gltf with external or data-URI buffers, .glb) -> SceneData,
  as src/App/loaders/load_gltf.cpp maps them: metallic-roughness materials,
  triangle primitives, node TRS / matrix transforms down the node hierarchy.
"""
import base64
import json
import os
import struct
import zlib

import numpy as np

from . import nanovdb, scenes

NVDB_MAGIC = 0x304244566F6E614E  # "NanoVDB0"
NVDB_VERSION = (32 << 21) | (3 << 10) | 3
CODEC_NONE, CODEC_ZIP = 0, 1
_META = struct.Struct("<4Q2I6d6i3dI4I3IHHI")
assert _META.size == 176


# ------------------------------------------------------------------ NanoVDB
def _string_hash(name):
    """stringHash of IO.h: h = h * 67 + c + (h >> 56), 64-bit."""
    h = 0
    for c in name.encode():
        overflow = h >> 56
        h = (h * 67 + c + overflow) & 0xFFFFFFFFFFFFFFFF
    return h


def grid_name(grid):
    raw = bytes(grid[40:40 + 256])
    return raw.split(b"\0", 1)[0].decode(errors="replace")


def _grid_meta(grid, file_size, codec):
    g = bytes(grid)
    name = grid_name(grid)
    tree = 672
    node_count = struct.unpack_from("<3I", g, tree + 32)
    tile_count = struct.unpack_from("<3I", g, tree + 44)
    voxels = struct.unpack_from("<Q", g, tree + 56)[0]
    root = tree + struct.unpack_from("<Q", g, tree + 24)[0]
    ibox = struct.unpack_from("<6i", g, root)
    wbox = struct.unpack_from("<6d", g, 560)
    vsize = struct.unpack_from("<3d", g, 608)
    gclass, gtype = struct.unpack_from("<2I", g, 632)
    meta = _META.pack(len(g), file_size, _string_hash(name), voxels, gtype, gclass, *wbox, *ibox, *vsize,
                      len(name.encode()) + 1, *node_count, 1, *tile_count, codec, 0, NVDB_VERSION)
    return meta + name.encode() + b"\0"


def write_nvdb(path, grids, codec=CODEC_NONE):
    """Write one segment holding the given grids (uint8 arrays)."""
    blobs = []
    for g in grids:
        raw = np.ascontiguousarray(g, np.uint8).tobytes()
        if codec == CODEC_ZIP:
            c = zlib.compress(raw)
            blobs.append(struct.pack("<Q", len(c)) + c)
        elif codec == CODEC_NONE:
            blobs.append(raw)
        else:
            raise ValueError("codec must be CODEC_NONE or CODEC_ZIP")
    with open(path, "wb") as f:
        f.write(struct.pack("<QIHH", NVDB_MAGIC, NVDB_VERSION, len(grids), codec))
        for g, b in zip(grids, blobs):
            f.write(_grid_meta(np.asarray(g, np.uint8), len(b), codec))
        for b in blobs:
            f.write(b)


def read_nvdb(path, n=0):
    """Grid n of a .nvdb file as a uint8 array (NONE and ZIP codecs)."""
    with open(path, "rb") as f:
        data = f.read()
    pos, counter = 0, 0
    while pos < len(data):
        magic, version, count, codec = struct.unpack_from("<QIHH", data, pos)
        if magic != NVDB_MAGIC:
            raise ValueError("not a NanoVDB file (bad magic)")
        if version >> 21 != 32:
            raise ValueError(f"NanoVDB major version {version >> 21} is not supported")
        pos += 16
        metas = []
        for _ in range(count):
            m = _META.unpack_from(data, pos)
            pos += _META.size + m[21]  # name size
            metas.append(m)
        for m in metas:
            grid_size, file_size = m[0], m[1]
            if counter == n:
                if codec == CODEC_NONE:
                    raw = data[pos:pos + grid_size]
                elif codec == CODEC_ZIP:
                    size = struct.unpack_from("<Q", data, pos)[0]
                    raw = zlib.decompress(data[pos + 8:pos + 8 + size])
                else:
                    raise ValueError("BLOSC-compressed NanoVDB files are not supported")
                if len(raw) != grid_size:
                    raise ValueError("truncated NanoVDB grid")
                return np.frombuffer(raw, np.uint8).copy()
            pos += file_size
            counter += 1
    raise IndexError(f"grid {n} not found in {path}")


# ------------------------------------------------------------- Mitsuba .vol
def read_vol(path):
    """Mitsuba grid volume, version 3, float32, one channel: returns
    (dense[i, j, k], pmin, pmax)."""
    with open(path, "rb") as f:
        head = f.read(48)
        if head[:3] != b"VOL" or head[3] != 3:
            raise ValueError("not a version-3 Mitsuba volume")
        vtype, xres, yres, zres, channels = struct.unpack_from("<5i", head, 4)
        if vtype != 1 or channels != 1:
            raise ValueError("only float32 single-channel volumes are supported")
        pmin = struct.unpack_from("<3f", head, 24)
        pmax = struct.unpack_from("<3f", head, 36)
        data = np.frombuffer(f.read(4 * xres * yres * zres), "<f4")
    return data.reshape(zres, yres, xres).transpose(2, 1, 0).copy(), pmin, pmax


def write_vol(path, dense, pmin=(0, 0, 0), pmax=(1, 1, 1)):
    d = np.asarray(dense, np.float32)
    xres, yres, zres = d.shape
    with open(path, "wb") as f:
        f.write(b"VOL" + bytes([3]) + struct.pack("<5i", 1, xres, yres, zres, 1) + struct.pack("<6f", *pmin, *pmax))
        f.write(d.transpose(2, 1, 0).astype("<f4").tobytes())


def vol_to_grid(path, name=None):
    """load_volumes.cpp:53-98: a fog-volume float grid with voxel size 1 and
    background 0 over the file's index box."""
    dense, _, _ = read_vol(path)
    return nanovdb.write_float_grid(dense, voxel_size=1.0, background=0.0,
                                    name=name or os.path.splitext(os.path.basename(path))[0])


# ---------------------------------------------------------------------- OBJ
def _read_mtl(path):
    mats, cur = {}, None
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "newmtl":
                cur = t[1]
                mats[cur] = {}
            elif cur and t[0] in ("Kd", "Ke"):
                mats[cur][t[0]] = tuple(float(x) for x in t[1:4])
    return mats


def read_obj(path, sd=None, transform=None):
    """Adds the OBJ's meshes to a SceneData (new one by default): one mesh
    instance per object/group (and material), polygons fanned to triangles,
    Kd as base colour and Ke as emission from the .mtl files it names."""
    sd = sd or scenes.SceneData()
    base = os.path.dirname(os.path.abspath(path))
    verts, mtl = [], {}
    groups = {}  # (group, material) -> list of triangles (vertex indices)
    g, m = "default", None
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            if t[0] == "v":
                verts.append([float(x) for x in t[1:4]])
            elif t[0] in ("o", "g"):
                g = " ".join(t[1:]) or "default"
            elif t[0] == "usemtl":
                m = t[1] if len(t) > 1 else None
            elif t[0] == "mtllib":
                for name in t[1:]:
                    p = os.path.join(base, name)
                    if os.path.exists(p):
                        mtl.update(_read_mtl(p))
            elif t[0] == "f":
                idx = []
                for tok in t[1:]:
                    i = int(tok.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(verts) + i)
                for k in range(1, len(idx) - 1):
                    groups.setdefault((g, m), []).append((idx[0], idx[k], idx[k + 1]))
    V = np.asarray(verts, np.float32)
    for (gname, mname), tris in groups.items():
        tri = np.asarray(tris, np.int64)
        used, remap = np.unique(tri, return_inverse=True)
        for start in range(0, len(tri), 0xFFFF):  # 16-bit primitive ids per instance
            part = remap.reshape(-1, 3)[start:start + 0xFFFF]
            vi, nt = sd.add_mesh(V[used], part.reshape(-1).astype(np.uint32))
            props = mtl.get(mname, {})
            kd = props.get("Kd", (0.8, 0.8, 0.8))
            ke = props.get("Ke", (0.0, 0.0, 0.0))
            mat = sd.add_material(scenes.material_record(base_color=[min(max(x, 0.0), 1.0) for x in kd], specular=0.0,
                                                         emission=ke))
            sd.add_mesh_instance(vi, nt, mat, transform)
    return sd


def write_obj(path, positions, indices):
    with open(path, "w") as f:
        for v in np.asarray(positions, np.float32):
            f.write(f"v {float(v[0])!r} {float(v[1])!r} {float(v[2])!r}\n")
        for t in np.asarray(indices).reshape(-1, 3):
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


# ---------------------------------------------------------------- RGBE .hdr
def _rgbe_to_float(rgbe):
    """RGBE -> float: mantissa * 2^(e - 136), black for e = 0."""
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136), np.float32(0.0)).astype(np.float32)
    out = np.empty(rgbe.shape[:-1] + (4,), np.float32)
    for c in range(3):
        out[..., c] = rgbe[..., c].astype(np.float32) * scale
    out[..., 3] = 1
    return out


def read_hdr(path):
    """Radiance RGBE image as float32 (H, W, 4) (alpha 1)."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    fmt_ok = False
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end].strip()
        pos = end + 1
        if line.startswith(b"FORMAT=") and b"32-bit_rle_rgbe" in line:
            fmt_ok = True
        if line == b"":
            break
    if not fmt_ok:
        raise ValueError("not an RGBE Radiance image")
    end = data.index(b"\n", pos)
    res = data[pos:end].split()
    pos = end + 1
    if res[0] != b"-Y" or res[2] != b"+X":
        raise ValueError("only -Y H +X W Radiance images are supported")
    H, W = int(res[1]), int(res[3])
    img = np.zeros((H, W, 4), np.uint8)
    buf = np.frombuffer(data, np.uint8)
    for y in range(H):
        if 8 <= W < 32768 and buf[pos] == 2 and buf[pos + 1] == 2 and (int(buf[pos + 2]) << 8 | int(buf[pos + 3])) == W:
            pos += 4
            for c in range(4):
                x = 0
                while x < W:
                    n = int(buf[pos])
                    pos += 1
                    if n > 128:
                        n -= 128
                        img[y, x:x + n, c] = buf[pos]
                        pos += 1
                    else:
                        img[y, x:x + n, c] = buf[pos:pos + n]
                        pos += n
                    x += n
        else:
            img[y] = buf[pos:pos + 4 * W].reshape(W, 4)
            pos += 4 * W
    return _rgbe_to_float(img)


def float_to_rgbe(rgb):
    rgb = np.asarray(rgb, np.float64)[..., :3]
    m = rgb.max(-1)
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    nz = m > 1e-32
    mant, ex = np.frexp(m[nz])
    scale = mant * 256.0 / m[nz]
    for c in range(3):
        out[..., c][nz] = np.floor(rgb[..., c][nz] * scale).astype(np.uint8)
    out[..., 3][nz] = (ex + 128).astype(np.uint8)
    return out


def write_hdr(path, img):
    img = np.asarray(img, np.float32)
    H, W = img.shape[:2]
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {H} +X {W}\n".encode())
        f.write(float_to_rgbe(img).tobytes())


# --------------------------------------------------------------------- glTF
_GLTF_COMPONENTS = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_GLTF_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_GLB_JSON, _GLB_BIN = 0x4E4F534A, 0x004E4942


def _gltf_document(path):
    """The JSON document and its buffers (.glb container or .gltf text)."""
    with open(path, "rb") as f:
        data = f.read()
    binary = None
    if data[:4] == b"glTF":  # 12-byte header, then chunks of {length, type, bytes}
        _, length = struct.unpack_from("<II", data, 4)
        off, doc = 12, None
        while off + 8 <= min(length, len(data)):
            clen, ctype = struct.unpack_from("<II", data, off)
            chunk = data[off + 8:off + 8 + clen]
            off += 8 + clen
            if ctype == _GLB_JSON:
                doc = json.loads(chunk.decode("utf-8"))
            elif ctype == _GLB_BIN and binary is None:
                binary = chunk
        if doc is None:
            raise ValueError(f"{path}: no JSON chunk")
    else:
        doc = json.loads(data.decode("utf-8"))
    base = os.path.dirname(os.path.abspath(path))
    buffers = []
    for b in doc.get("buffers", []):
        uri = b.get("uri")
        if uri is None:
            if binary is None:
                raise ValueError(f"{path}: buffer without a URI outside a .glb")
            buffers.append(binary)
        elif uri.startswith("data:"):
            buffers.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            with open(os.path.join(base, uri), "rb") as f:
                buffers.append(f.read())
    return doc, buffers


def _gltf_accessor(doc, buffers, index):
    """Accessor `index` as a (count, width) array (strided views copied)."""
    a = doc["accessors"][index]
    if "sparse" in a:
        raise ValueError("sparse glTF accessors are not supported")
    dt = np.dtype(_GLTF_COMPONENTS[a["componentType"]])
    width, count = _GLTF_WIDTH[a["type"]], a["count"]
    if "bufferView" not in a:
        return np.zeros((count, width), dt)
    bv = doc["bufferViews"][a["bufferView"]]
    stride = bv.get("byteStride", dt.itemsize * width)
    start = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
    buf = buffers[bv["buffer"]]
    if count and start + stride * (count - 1) + dt.itemsize * width > len(buf):
        raise ValueError(f"glTF accessor {index} reaches past its buffer")
    arr = np.ndarray((count, width), dt, buf, start, (stride, dt.itemsize)).copy()
    if a.get("normalized") and dt.kind in "iu":
        arr = np.maximum(arr.astype(np.float32) / np.float32(np.iinfo(dt).max), np.float32(-1))
    return arr


def _lum(c):
    c = np.asarray(c, np.float32)
    return float(c[0] * np.float32(0.2126) + c[1] * np.float32(0.7152) + c[2] * np.float32(0.0722))


def _gltf_material(m):
    """Scene::makeMetallicRoughnessMaterial (Scene.cpp:195-209) on the factors of
    a glTF material with the extensions load_gltf.cpp:92-128 reads (textures are
    not sampled by the renderer and are left out)."""
    pbr = m.get("pbrMetallicRoughness", {})
    ext = m.get("extensions", {})
    emission = np.asarray(m.get("emissiveFactor", [0, 0, 0]), np.float32)
    if "KHR_materials_emissive_strength" in ext:
        emission = emission * np.float32(ext["KHR_materials_emissive_strength"].get("emissiveStrength", 1.0))
    eta = ext["KHR_materials_ior"].get("ior", 1.5) if "KHR_materials_ior" in ext else 1.5
    t = ext["KHR_materials_transmission"].get("transmissionFactor", 0.0) if "KHR_materials_transmission" in ext else 0.0
    specular = 1.0
    if "KHR_materials_specular" in ext:
        sp = ext["KHR_materials_specular"]
        if "specularColorFactor" in sp:
            specular = _lum(sp["specularColorFactor"])
        elif "specularFactor" in sp:
            specular = sp["specularFactor"]
    clearcoat = ext["KHR_materials_clearcoat"].get("clearcoatFactor", 0.0) if "KHR_materials_clearcoat" in ext else 0.0
    return scenes.material_record(base_color=pbr.get("baseColorFactor", [1, 1, 1, 1])[:3], specular=specular,
                                  emission=emission, sheen=0.0, metallic=pbr.get("metallicFactor", 1.0),
                                  roughness=pbr.get("roughnessFactor", 1.0), anisotropic=0.0, subsurface=0.0,
                                  clearcoat=clearcoat, clearcoat_gloss=0.0, transmission=_lum([t, t, t]), eta=eta)


def gltf_node_transform(node):
    """TransformData of a glTF node (load_gltf.cpp:267-277): translation,
    rotation quaternion normalised as quatf.h:64-67 does and scale, composed as
    the host TransformData(t, r, s) constructor does (transform.h:61-64:
    Eigen's Quaternion::toRotationMatrix times Scaling(s)), or the top 3x4 of
    the column-major matrix; the identity when it has neither."""
    if any(k in node for k in ("translation", "rotation", "scale")):
        t = np.asarray(node.get("translation", [0, 0, 0]), np.float32)
        s = np.asarray(node.get("scale", [1, 1, 1]), np.float32)
        q = np.asarray(node.get("rotation", [0, 0, 0, 1]), np.float32)
        q = q * (np.float32(1) / np.sqrt(np.float32(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]))
        x, y, z, w = q
        tx, ty, tz = np.float32(2) * x, np.float32(2) * y, np.float32(2) * z
        twx, twy, twz = tx * w, ty * w, tz * w
        txx, txy, txz = tx * x, ty * x, tz * x
        tyy, tyz, tzz = ty * y, tz * y, tz * z
        one = np.float32(1)
        R = np.array([[one - (tyy + tzz), txy - twz, txz + twy],
                      [txy + twz, one - (txx + tzz), tyz - twx],
                      [txz - twy, tyz + twx, one - (txx + tyy)]], np.float32)
        M = np.zeros((3, 4), np.float32)
        M[:, :3] = R * s[None, :]
        M[:, 3] = t
        return M
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).reshape(4, 4).T[:3, :].astype(np.float32)
    return np.eye(3, 4, dtype=np.float32)


def _compose(a, b):
    """a o b for 3x4 affine transforms."""
    m = np.zeros((3, 4), np.float32)
    m[:, :3] = a[:, :3] @ b[:, :3]
    m[:, 3] = a[:, :3] @ b[:, 3] + a[:, 3]
    return m


def read_gltf(path, sd=None):
    """Adds a glTF 2.0 scene (.gltf or .glb) to a SceneData (new by default),
    following src/App/loaders/load_gltf.cpp: every node is placed, with its
    transform composed with its parents' (:297-299), and each triangle
    primitive of its mesh becomes one mesh instance (positions, normals and
    the first texture coordinates, indices widened to 32 bits).  Primitives
    without a material get glTF's default material (base colour 1, metallic 1,
    roughness 1).  Other primitive modes, meshes beyond 65535 triangles (the
    16-bit primitive ids of InstanceData) and KHR_lights_punctual sphere lights
    (procedural spheres, which the renderer does not take) raise ValueError."""
    sd = sd or scenes.SceneData()
    doc, buffers = _gltf_document(path)
    mats = [sd.add_material(_gltf_material(m)) for m in doc.get("materials", [])]
    default_mat = None
    meshes = {}  # (mesh, primitive) -> (vertex info, triangle count, material address)
    for mi, mesh in enumerate(doc.get("meshes", [])):
        for pi, prim in enumerate(mesh.get("primitives", [])):
            if prim.get("mode", 4) != 4:
                raise ValueError(f"glTF mesh {mi} primitive {pi}: only triangle lists are supported")
            attrs = prim["attributes"]
            pos = _gltf_accessor(doc, buffers, attrs["POSITION"]).astype(np.float32)
            nrm = _gltf_accessor(doc, buffers, attrs["NORMAL"]).astype(np.float32) if "NORMAL" in attrs else None
            uv = _gltf_accessor(doc, buffers, attrs["TEXCOORD_0"]).astype(np.float32) if "TEXCOORD_0" in attrs else None
            if "indices" in prim:
                idx = _gltf_accessor(doc, buffers, prim["indices"]).reshape(-1).astype(np.uint32)
            else:
                idx = np.arange(len(pos), dtype=np.uint32)
            if len(idx) % 3 or (len(idx) and int(idx.max()) >= len(pos)):
                raise ValueError(f"glTF mesh {mi} primitive {pi}: malformed indices")
            if len(idx) // 3 > 0xFFFF:
                raise ValueError(f"glTF mesh {mi} primitive {pi}: more than 65535 triangles")
            if "material" in prim:
                mat = mats[prim["material"]]
            else:
                if default_mat is None:
                    default_mat = sd.add_material(_gltf_material({}))
                mat = default_mat
            vi, nt = sd.add_mesh(pos, idx, normals=nrm, texcoords=uv)
            meshes[(mi, pi)] = (vi, nt, mat)
    nodes = doc.get("nodes", [])
    parent = {}
    for i, n in enumerate(nodes):
        for c in n.get("children", []):
            parent[c] = i
    world = {}

    def world_of(i):
        if i not in world:
            local = gltf_node_transform(nodes[i])
            world[i] = _compose(world_of(parent[i]), local) if i in parent else local
        return world[i]

    for i, n in enumerate(nodes):
        light = n.get("extensions", {}).get("KHR_lights_punctual")
        if light is not None and "light" in light:
            l = doc.get("extensions", {}).get("KHR_lights_punctual", {}).get("lights", [])[light["light"]]
            if l.get("type") == "point" and "radius" in l.get("extras", {}):
                raise ValueError(f"glTF node {i}: sphere lights (KHR_lights_punctual point with a radius) are not supported")
        if "mesh" in n and n["mesh"] < len(doc.get("meshes", [])):
            for pi in range(len(doc["meshes"][n["mesh"]].get("primitives", []))):
                vi, nt, mat = meshes[(n["mesh"], pi)]
                sd.add_mesh_instance(vi, nt, mat, world_of(i))
    return sd


# ---------------------------------------------------------------------- EXR
EXR_NONE, EXR_RLE, EXR_ZIPS, EXR_ZIP, EXR_PIZ = 0, 1, 2, 3, 4
_EXR_LINES = {EXR_NONE: 1, EXR_RLE: 1, EXR_ZIPS: 1, EXR_ZIP: 16}
_EXR_HALF, _EXR_FLOAT, _EXR_UINT = 1, 2, 0
_EXR_DTYPE = {_EXR_UINT: np.dtype("<u4"), _EXR_HALF: np.dtype("<f2"), _EXR_FLOAT: np.dtype("<f4")}


def _exr_unpredict(buf):
    """Undo the byte predictor and the even/odd split of the RLE and ZIP
    codecs (OpenEXR's ImfZip / ImfRle reconstruction)."""
    t = np.frombuffer(buf, np.uint8)
    t = (np.cumsum(t.astype(np.int64)) - 128 * np.arange(len(t))) % 256  # t[i] = t[i-1] + t[i] - 128
    t = t.astype(np.uint8)
    half = (len(t) + 1) // 2
    out = np.empty(len(t), np.uint8)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _exr_rle_decode(data, expect):
    out = bytearray()
    i = 0
    while i < len(data) and len(out) < expect:
        n = struct.unpack_from("b", data, i)[0]
        i += 1
        if n < 0:
            out += data[i:i - n]
            i -= n
        else:
            out += data[i:i + 1] * (n + 1)
            i += 1
    if len(out) != expect:
        raise ValueError("EXR: RLE block decodes to the wrong size")
    return bytes(out)


def read_exr(path):
    """OpenEXR image as float32 (H, W, 4), as tinyexr's LoadEXR returns it
    (extern/tiny_exr.h:5850-6080): channel names taken after their last '.',
    R, G and B looked up among the first four channels in file order (A, or 1
    when absent), a single channel copied into all four; HALF read as float,
    UINT bits reinterpreted as float.  Tiled, multi-part, deep and PIZ, PXR24
    or B44 files raise ValueError."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < 8 or struct.unpack_from("<I", data, 0)[0] != 20000630:
        raise ValueError(f"{path}: not an OpenEXR file")
    version = struct.unpack_from("<I", data, 4)[0]
    if version & 0x200:
        raise ValueError(f"{path}: tiled EXR images are not supported")
    if version & 0x1800:
        raise ValueError(f"{path}: multi-part and deep EXR images are not supported")
    pos, attrs = 8, {}
    while True:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        pos = end + 1
        if not name:
            break
        end = data.index(b"\0", pos)
        typ = data[pos:end].decode()
        size = struct.unpack_from("<i", data, end + 1)[0]
        attrs[name] = (typ, data[end + 5:end + 5 + size])
        pos = end + 5 + size
    channels = []
    raw = attrs["channels"][1]
    p = 0
    while raw[p] != 0:
        end = raw.index(b"\0", p)
        cname = raw[p:end].decode()
        ptype, _, xs, ys = struct.unpack_from("<iB3xii", raw, end + 1)
        if (xs, ys) != (1, 1):
            raise ValueError(f"{path}: subsampled channels are not supported")
        channels.append((cname, ptype))
        p = end + 17
    comp = attrs["compression"][1][0]
    if comp not in _EXR_LINES:
        raise ValueError(f"{path}: EXR compression {comp} is not supported (NONE, RLE, ZIPS, ZIP are)")
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    lines = _EXR_LINES[comp]
    nblocks = (H + lines - 1) // lines
    offsets = struct.unpack_from(f"<{nblocks}Q", data, pos)
    planes = [np.zeros((H, W), _EXR_DTYPE[t]) for _, t in channels]
    row_bytes = sum(W * _EXR_DTYPE[t].itemsize for _, t in channels)
    for off in offsets:
        y, size = struct.unpack_from("<ii", data, off)
        ny = min(lines, y1 - y + 1)
        expect = ny * row_bytes
        blk = data[off + 8:off + 8 + size]
        if size < expect:
            if comp in (EXR_ZIPS, EXR_ZIP):
                blk = _exr_unpredict(zlib.decompress(blk))
            elif comp == EXR_RLE:
                blk = _exr_unpredict(_exr_rle_decode(blk, expect))
        if len(blk) != expect:
            raise ValueError(f"{path}: EXR block at y={y} has {len(blk)} bytes, expected {expect}")
        q = 0
        for r in range(ny):
            for c, (_, t) in enumerate(channels):
                n = W * _EXR_DTYPE[t].itemsize
                planes[c][y - y0 + r] = np.frombuffer(blk, _EXR_DTYPE[t], W, q)
                q += n
    fl = [pl.view(np.float32) if t == _EXR_UINT else pl.astype(np.float32) for pl, (_, t) in zip(planes, channels)]
    short = [n.rsplit(".", 1)[-1] for n, _ in channels]
    out = np.empty((H, W, 4), np.float32)
    if len(channels) == 1:
        out[...] = fl[0][..., None]
        return out
    idx = {n: i for i, n in enumerate(short[:4])}
    for k in "RGB":
        if k not in idx:
            raise ValueError(f"{path}: {k} channel not found")
    out[..., 0], out[..., 1], out[..., 2] = fl[idx["R"]], fl[idx["G"]], fl[idx["B"]]
    out[..., 3] = fl[idx["A"]] if "A" in idx else 1.0
    return out


def _exr_attr(name, typ, value):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(value)) + value


def write_exr(path, img, compression=EXR_ZIP, half=False):
    """An (H, W, 3|4) image as a single-part scanline OpenEXR file with
    channels A (if present), B, G, R in HALF or FLOAT and NONE or ZIP
    compression."""
    img = np.asarray(img, np.float32)
    H, W, C = img.shape
    if compression not in (EXR_NONE, EXR_ZIP):
        raise ValueError("write_exr writes NONE or ZIP")
    names = (["A"] if C == 4 else []) + ["B", "G", "R"]
    src = {"R": 0, "G": 1, "B": 2, "A": 3}
    ptype = _EXR_HALF if half else _EXR_FLOAT
    dt = _EXR_DTYPE[ptype]
    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1) for n in names) + b"\0"
    hdr = (struct.pack("<II", 20000630, 2)
           + _exr_attr("channels", "chlist", chl)
           + _exr_attr("compression", "compression", bytes([compression]))
           + _exr_attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, W - 1, H - 1))
           + _exr_attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, W - 1, H - 1))
           + _exr_attr("lineOrder", "lineOrder", b"\0")
           + _exr_attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
           + _exr_attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0))
           + _exr_attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    lines = _EXR_LINES[compression]
    blocks = []
    for y in range(0, H, lines):
        rows = img[y:y + lines]
        raw = b"".join(np.ascontiguousarray(rows[r, :, src[n]]).astype(dt).tobytes()
                       for r in range(rows.shape[0]) for n in names)
        if compression == EXR_ZIP:
            b = np.frombuffer(raw, np.uint8)
            inter = np.concatenate([b[0::2], b[1::2]]).astype(np.int64)
            pred = inter.copy()
            pred[1:] = (inter[1:] - inter[:-1] + 128) % 256
            z = zlib.compress(pred.astype(np.uint8).tobytes(), 6)
            if len(z) < len(raw):
                raw = z
        blocks.append(struct.pack("<ii", y, len(raw)) + raw)
    table_start = len(hdr)
    offs, at = [], table_start + 8 * len(blocks)
    for b in blocks:
        offs.append(at)
        at += len(b)
    with open(path, "wb") as f:
        f.write(hdr + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(blocks))
