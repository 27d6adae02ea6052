"""glTF 2.0 test files: a scene exercising what src/App/loaders/load_gltf.cpp
reads (u16 and u32 indices, normals, texture coordinates, the material
extensions, TRS and matrix nodes in a hierarchy) and writers for the .gltf
(base64 data URI) and .glb containers."""
import base64
import json
import struct

import numpy as np

from stratum2_amd import scenes

def gltf_content(seed=0):
    """Two meshes (u16-indexed quad with normals and texture coordinates;
    u32-indexed icosphere without either), three materials exercising the
    extensions load_gltf.cpp reads, and a two-level node hierarchy (TRS parent,
    column-major matrix child, plain second root)."""
    rng = np.random.default_rng(seed)
    quad_p = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    quad_n = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    quad_uv = rng.uniform(size=(4, 2)).astype(np.float32)
    quad_i = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    ico_p, ico_f = scenes.icosphere(1)
    ico_p = ico_p.astype(np.float32)
    ico_i = ico_f.reshape(-1).astype(np.uint32)
    blobs, views, accessors = [], [], []

    def add(arr, ctype, typ, target=None):
        off = sum(len(b) for b in blobs)
        raw = np.ascontiguousarray(arr).tobytes()
        raw += b"\0" * (-len(raw) % 4)
        blobs.append(raw)
        views.append({"buffer": 0, "byteOffset": off, "byteLength": arr.nbytes})
        accessors.append({"bufferView": len(views) - 1, "componentType": ctype, "type": typ,
                          "count": int(arr.shape[0])})
        return len(accessors) - 1

    a_qp, a_qn = add(quad_p, 5126, "VEC3"), add(quad_n, 5126, "VEC3")
    a_quv, a_qi = add(quad_uv, 5126, "VEC2"), add(quad_i, 5123, "SCALAR")
    a_ip, a_ii = add(ico_p, 5126, "VEC3"), add(ico_i, 5125, "SCALAR")
    materials = [
        {"pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.5, 0.25, 1.0], "metallicFactor": 0.25,
                                  "roughnessFactor": 0.6},
         "emissiveFactor": [1.0, 0.5, 0.0],
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0},
                        "KHR_materials_ior": {"ior": 1.33},
                        "KHR_materials_transmission": {"transmissionFactor": 0.65},
                        "KHR_materials_clearcoat": {"clearcoatFactor": 0.4},
                        "KHR_materials_specular": {"specularColorFactor": [0.2, 0.6, 1.0]}}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.3, 0.3, 0.8, 1.0]},
         "extensions": {"KHR_materials_specular": {"specularFactor": 0.35}}},
        {},
    ]
    child_matrix = [0.0, 0.0, -2.0, 0.0,  0.0, 2.0, 0.0, 0.0,  2.0, 0.0, 0.0, 0.0,  0.5, -1.0, 3.0, 1.0]
    doc = {
        "asset": {"version": "2.0"},
        "meshes": [{"primitives": [{"attributes": {"POSITION": a_qp, "NORMAL": a_qn, "TEXCOORD_0": a_quv},
                                    "indices": a_qi, "material": 0}]},
                   {"primitives": [{"attributes": {"POSITION": a_ip}, "indices": a_ii, "material": 1},
                                   {"attributes": {"POSITION": a_ip}, "indices": a_ii, "material": 2}]}],
        "materials": materials,
        "nodes": [{"translation": [1.0, 2.0, 3.0], "rotation": [0.0, 0.0, 0.3826834, 0.9238795],
                   "scale": [2.0, 1.0, 0.5], "mesh": 0, "children": [1]},
                  {"matrix": child_matrix, "mesh": 1},
                  {"mesh": 0}],
        "scenes": [{"nodes": [0, 2]}],
        "accessors": accessors,
        "bufferViews": views,
    }
    data = b"".join(blobs)
    return doc, data, dict(quad=(quad_p, quad_n, quad_uv, quad_i), ico=(ico_p, ico_i))


def write_gltf(path, doc, data):
    d = json.loads(json.dumps(doc))
    d["buffers"] = [{"byteLength": len(data), "uri": "data:application/octet-stream;base64,"
                     + base64.b64encode(data).decode()}]
    open(path, "w").write(json.dumps(d))


def write_glb(path, doc, data):
    d = json.loads(json.dumps(doc))
    d["buffers"] = [{"byteLength": len(data)}]
    js = json.dumps(d).encode()
    js += b" " * (-len(js) % 4)
    data = data + b"\0" * (-len(data) % 4)
    body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(data), 0x004E4942) + data
    open(path, "wb").write(struct.pack("<4sII", b"glTF", 2, 12 + len(body)) + body)
