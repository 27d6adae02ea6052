"""Seeded synthetic scenes in the reference's scene-buffer formats.

SceneBuilder packs meshes, materials and instances exactly like
Scene::updateFrameData does (src/App/Scene.cpp:428-791): 64-byte material
records (Material::store, src/App/Material.hpp:22-40), MeshVertexInfo records
(src/Shaders/compat/scene.h:89-132), InstanceData words (scene.h:30-80) and a
16-byte environment record (EnvironmentMap::store, src/App/Scene.hpp:42-45).
The workloads (Cornell box with a displaced icosphere; 2000 displaced spheres
with 4096 emissive triangles under a procedural sky) are generated from seeds.
"""
import ctypes as C
import math

import numpy as np

from . import abi

NO_BUFFER = 0xFFFF  # buffer index >= gVertexBufferCount (2048) means "absent"


def unorm8(v):
    """(uint)floor(saturate(v)*255 + 0.5f), src/Shaders/compat/material_data.h:51-64."""
    v = np.float32(min(max(float(v), 0.0), 1.0))
    return int(math.floor(np.float32(v * np.float32(255)) + np.float32(0.5)))


def pack4(a, b, c, d):
    return unorm8(a) | unorm8(b) << 8 | unorm8(c) << 16 | unorm8(d) << 24


def material_record(base_color=(0.8, 0.8, 0.8), specular=0.5, emission=(0, 0, 0), sheen=0.0,
                    metallic=0.0, roughness=0.5, anisotropic=0.0, subsurface=0.0, clearcoat=0.0,
                    clearcoat_gloss=0.0, transmission=0.0, eta=1.5):
    """64-byte Material::store record (Material.hpp:22-40) as 16 u32 words."""
    e = np.asarray(emission, np.float32)
    scale = np.float32(e.max())
    en = e / scale if scale > 0 else np.zeros(3, np.float32)
    w = np.zeros(16, np.uint32)
    w[0] = pack4(*base_color, specular)
    w[1] = pack4(en[0], en[1], en[2], sheen)
    w[2] = pack4(metallic, roughness, anisotropic, subsurface)
    w[3] = pack4(clearcoat, clearcoat_gloss, transmission, eta - 1.0)
    w[4:8] = abi.ST_NO_IMAGE
    w[8] = np.float32(scale).view(np.uint32)
    w[9] = abi.ST_NO_IMAGE          # bump image
    w[10] = np.float32(0).view(np.uint32)
    w[11] = 0
    w[12] = abi.ST_NO_IMAGE         # alpha mask
    w[13] = np.float32(0.5).view(np.uint32)
    return w


def medium_record(density=(1.0, 1.0, 1.0), albedo=(1.0, 1.0, 1.0), anisotropy=0.0,
                  density_volume=0xFFFFFFFF, albedo_volume=0xFFFFFFFF):
    """28-byte Medium::store record (src/App/Material.hpp:49-66) as 7 u32 words:
    density scale f32x3, albedo RGBA8 (alpha 1), density/albedo volume
    indices (0xFFFFFFFF = none), anisotropy g."""
    w = np.zeros(7, np.uint32)
    w[0:3] = np.asarray(density, np.float32).view(np.uint32)
    w[3] = pack4(*albedo, 1.0)
    w[4] = density_volume
    w[5] = albedo_volume
    w[6] = np.float32(anisotropy).view(np.uint32)
    return w


class SceneData:
    """Host copy of one scene's buffers, convertible to st_scene_desc."""

    def __init__(self):
        self.instances = []      # (type_material_address, data)
        self.transforms = []     # 3x4
        self.material_words = []
        self.vertex_infos = []
        self.buffers = []        # bytes objects
        self.images = []         # float32 HxWx4
        self.env_address = 0xFFFFFFFF
        self.volumes = []
        self.volume_info = []
        self._keep = []

    # -- building -------------------------------------------------------
    def add_material(self, words):
        addr = 4 * sum(len(w) for w in self.material_words)
        self.material_words.append(np.asarray(words, np.uint32))
        return addr

    def add_buffer(self, arr):
        self.buffers.append(np.ascontiguousarray(arr).tobytes())
        return len(self.buffers) - 1

    def add_mesh(self, positions, indices, normals=None, texcoords=None):
        """One mesh in its own buffers (u32 indices, f32x3 positions, optional
        f32x3 normals and f32x2 texture coordinates)."""
        positions = np.ascontiguousarray(positions, np.float32)
        indices = np.ascontiguousarray(indices, np.uint32)
        ib = self.add_buffer(indices)
        pb = self.add_buffer(positions)
        nb = self.add_buffer(np.ascontiguousarray(normals, np.float32)) if normals is not None else NO_BUFFER
        tb = self.add_buffer(np.ascontiguousarray(texcoords, np.float32)) if texcoords is not None else NO_BUFFER
        vi = abi.st_mesh_vertex_info()
        vi.packed_buffer_indices[0] = ib | pb << 16
        vi.packed_buffer_indices[1] = nb | tb << 16
        vi.packed_strides = (4 | 12 << 8 | (12 if normals is not None else 0) << 16
                             | (8 if texcoords is not None else 0) << 24)
        vi.packed_offsets[0] = 0
        vi.packed_offsets[1] = 0
        vi.packed_offsets[2] = 0
        vi.packed_offsets[3] = 0
        self.vertex_infos.append(vi)
        return len(self.vertex_infos) - 1, len(indices) // 3

    def add_mesh_shared(self, ib, pb, index_offset_bytes, position_offset_bytes, tri_count):
        vi = abi.st_mesh_vertex_info()
        vi.packed_buffer_indices[0] = ib | pb << 16
        vi.packed_buffer_indices[1] = NO_BUFFER | NO_BUFFER << 16
        vi.packed_strides = 4 | 12 << 8
        vi.packed_offsets[0] = index_offset_bytes
        vi.packed_offsets[1] = position_offset_bytes
        self.vertex_infos.append(vi)
        return len(self.vertex_infos) - 1, tri_count

    def add_mesh_instance(self, vertex_info, tri_count, material_address, transform=None):
        if tri_count > 0xFFFF or vertex_info > 0xFFFF:
            raise ValueError("16-bit primitive/vertex-info limit (scene.h:46-58)")
        self.instances.append((0 | (material_address << 4), vertex_info | tri_count << 16))
        self.transforms.append(np.eye(3, 4, dtype=np.float32) if transform is None else np.asarray(transform, np.float32))
        return len(self.instances) - 1

    def add_sphere_instance(self, material_address, radius, transform=None):
        """SphereInstanceData (compat/scene.h:60-69): a sphere of `radius` about
        the instance's object-space origin.  The reference's scene build keeps
        only the translation of a sphere's node transform (Scene.cpp:642-645)."""
        self.instances.append((1 | (material_address << 4), int(np.float32(radius).view(np.uint32))))
        self.transforms.append(np.eye(3, 4, dtype=np.float32) if transform is None else np.asarray(transform, np.float32))
        return len(self.instances) - 1

    def add_volume(self, grid):
        """A NanoVDB float grid (bytes, as stratum2_amd.nanovdb writes it)."""
        self.volumes.append(np.ascontiguousarray(grid, np.uint8))
        return len(self.volumes) - 1

    def add_volume_instance(self, medium_address, volume_index=0xFFFFFFFF, transform=None):
        """VolumeInstanceData (compat/scene.h:71-80) plus its VolumeInfo
        (object-space bounds: the grid's world bbox, or the [-1, 1] cube
        without a grid; Scene.cpp:672-699)."""
        idx = len(self.instances)
        self.instances.append((2 | (medium_address << 4), volume_index))
        self.transforms.append(np.eye(3, 4, dtype=np.float32) if transform is None else np.asarray(transform, np.float32))
        if volume_index == 0xFFFFFFFF:
            mn, mx = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
        else:
            bb = np.frombuffer(self.volumes[volume_index][560:608].tobytes(), "<f8")
            mn, mx = tuple(np.float32(bb[:3])), tuple(np.float32(bb[3:]))
        info = abi.st_volume_info()
        info.min[:] = mn
        info.max[:] = mx
        info.instance_index = idx
        self.volume_info.append(info)
        return idx

    def set_environment(self, emission=(1, 1, 1), image=None):
        img = abi.ST_NO_IMAGE
        if image is not None:
            self.images.append(np.ascontiguousarray(image, np.float32))
            img = len(self.images) - 1
        e = np.asarray(emission, np.float32).view(np.uint32)
        self.env_address = self.add_material(np.array([e[0], e[1], e[2], img], np.uint32))

    # -- export ---------------------------------------------------------
    def desc(self):
        """st_scene_desc pointing into arrays kept alive by this object."""
        keep = []
        n = len(self.instances)
        inst = (abi.st_instance * max(n, 1))()
        for i, (a, b) in enumerate(self.instances):
            inst[i].type_material_address, inst[i].data = a, b
        xf = np.ascontiguousarray(np.stack(self.transforms) if n else np.zeros((1, 3, 4)), np.float32)
        mats = np.ascontiguousarray(np.concatenate(self.material_words) if self.material_words else np.zeros(1), np.uint32)
        vis = (abi.st_mesh_vertex_info * max(len(self.vertex_infos), 1))(*self.vertex_infos)
        bufs = (abi.st_buffer_view * max(len(self.buffers), 1))()
        for i, b in enumerate(self.buffers):
            cb = C.create_string_buffer(b, len(b))
            keep.append(cb)
            bufs[i].data = C.cast(cb, C.c_void_p)
            bufs[i].size_bytes = len(b)
        imgs = (abi.st_image_desc * max(len(self.images), 1))()
        for i, im in enumerate(self.images):
            imgs[i].rgba = im.ctypes.data
            imgs[i].height, imgs[i].width = im.shape[0], im.shape[1]
        vols = (abi.st_buffer_view * max(len(self.volumes), 1))()
        for i, v in enumerate(self.volumes):
            vols[i].data = v.ctypes.data
            vols[i].size_bytes = v.nbytes
        vinf = (abi.st_volume_info * max(len(self.volume_info), 1))(*self.volume_info)
        d = abi.st_scene_desc()
        d.instance_count = n
        d.instances = C.cast(inst, C.c_void_p)
        d.transforms = xf.ctypes.data
        d.material_word_count = mats.size if self.material_words else 0
        d.material_data = mats.ctypes.data
        d.mesh_vertex_info_count = len(self.vertex_infos)
        d.mesh_vertex_info = C.cast(vis, C.c_void_p)
        d.vertex_buffer_count = len(self.buffers)
        d.vertex_buffers = C.cast(bufs, C.c_void_p)
        d.image_count = len(self.images)
        d.images = C.cast(imgs, C.c_void_p)
        d.environment_material_address = self.env_address
        d.volume_count = len(self.volumes)
        d.volumes = C.cast(vols, C.c_void_p)
        d.volume_info_count = len(self.volume_info)
        d.volume_info = C.cast(vinf, C.c_void_p)
        keep += [inst, xf, mats, vis, bufs, imgs, vols, vinf]
        self._keep = keep
        return d

    @property
    def triangle_count(self):
        return sum(b >> 16 for a, b in self.instances if (a & 0xF) == 0)

    def bounds(self):
        """World-space box (lo, hi) holding every mesh and volume instance: each
        instance's object-space box (its position buffer from the mesh's
        offset on; a volume's bounds) with its 8 corners transformed."""
        lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
        boxes = {}
        vol_box = {int(v.instance_index): (np.array(v.min[:]), np.array(v.max[:])) for v in self.volume_info}
        for i, ((a, b), xf) in enumerate(zip(self.instances, self.transforms)):
            kind = a & 0xF
            if kind == 0:
                vi = self.vertex_infos[b & 0xFFFF]
                key = (vi.packed_buffer_indices[0] >> 16, vi.packed_offsets[1])
                if key not in boxes:
                    pts = np.frombuffer(self.buffers[key[0]], np.float32)[key[1] // 4:]
                    pts = pts[:len(pts) // 3 * 3].reshape(-1, 3)
                    boxes[key] = (pts.min(0), pts.max(0))
                bmin, bmax = boxes[key]
            elif kind == 2 and i in vol_box:
                bmin, bmax = vol_box[i]
            else:
                continue
            c = np.array([[bmax[0] if k & 1 else bmin[0], bmax[1] if k & 2 else bmin[1], bmax[2] if k & 4 else bmin[2]]
                          for k in range(8)], np.float64)
            w = c @ np.asarray(xf, np.float64)[:, :3].T + np.asarray(xf, np.float64)[:, 3]
            lo, hi = np.minimum(lo, w.min(0)), np.maximum(hi, w.max(0))
        return tuple(lo.astype(np.float32)), tuple(hi.astype(np.float32))


# ---------------------------------------------------------------- camera
def quat_to_transform(pos, q):
    """TransformData(t, quatf, 1), src/Shaders/compat/transform.h:73-104 (q = x, y, z, w)."""
    x, y, z, w = [np.float32(v) for v in q]
    sqw, sqx, sqy, sqz = w * w, x * x, y * y, z * z
    invs = np.float32(1) / (sqx + sqy + sqz + sqw)
    m = np.zeros((3, 4), np.float32)
    m[0, 0] = (sqx - sqy - sqz + sqw) * invs
    m[1, 1] = (-sqx + sqy - sqz + sqw) * invs
    m[2, 2] = (-sqx - sqy + sqz + sqw) * invs
    t1, t2 = x * z, y * w
    m[2, 0] = 2 * (t1 + t2) * invs
    m[0, 2] = 2 * (t1 - t2) * invs
    t1, t2 = x * y, z * w
    m[1, 0] = 2 * (t1 - t2) * invs
    m[0, 1] = 2 * (t1 + t2) * invs
    t1, t2 = z * y, x * w
    m[1, 2] = 2 * (t1 + t2) * invs
    m[2, 1] = 2 * (t1 - t2) * invs
    m[:, 3] = pos
    return m


def make_view(width, height, position=(0, 1.5, 0), orientation=(0, 0, 0, 1), fovy_deg=70.0):
    """Camera of src/main.cpp:104-116 with the aspect fix of main.cpp:219-224
    (ProjectionData::makePerspective, transform.h:181-190, near = -0.001)."""
    v = abi.st_view()
    fovy = np.float32(fovy_deg) * np.float32(math.pi) / np.float32(180)
    aspect = np.float32(height) / np.float32(width)
    s1 = np.float32(1) / np.float32(math.tan(float(fovy) / 2))
    v.projection.scale[1] = s1
    v.projection.scale[0] = aspect * s1
    v.projection.offset[0] = v.projection.offset[1] = 0
    v.projection.near_plane = -0.001
    v.projection.far_plane = -10000.0
    v.projection.sensor_area = 0
    v.projection.vertical_fov = fovy
    v.image_min[0] = v.image_min[1] = 0
    v.image_max[0], v.image_max[1] = width, height
    m = quat_to_transform(position, orientation)
    for r in range(3):
        for c in range(4):
            v.camera_to_world.m[r][c] = float(m[r, c])
    return v


def make_ortho_view(width, height, size=(2.0, 2.0), position=(0, 1.5, 0), orientation=(0, 0, 0, 1), znear=-0.001,
                    zfar=-10000.0):
    """ProjectionData::makeOrthographic (compat/transform.h:172-180): scale
    2/size, vertical_fov -1 marks the orthographic branch of backProject."""
    v = make_view(width, height, position, orientation)
    v.projection.scale[0] = np.float32(2) / np.float32(size[0])
    v.projection.scale[1] = np.float32(2) / np.float32(size[1])
    v.projection.near_plane = znear
    v.projection.far_plane = zfar
    v.projection.vertical_fov = -1.0
    return v


def axis_angle_quat(axis, deg):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    h = math.radians(deg) / 2
    return (a[0] * math.sin(h), a[1] * math.sin(h), a[2] * math.sin(h), math.cos(h))


def look_at_quat(position, target):
    """Orientation (x, y, z, w) turning the camera's forward axis (-z) towards
    target: a yaw about y after a pitch about x (no roll)."""
    d = np.asarray(target, np.float64) - np.asarray(position, np.float64)
    d /= np.linalg.norm(d)
    # quat_to_transform (TransformData) rotates by the conjugate of the
    # usual convention, hence the negated angles
    yaw, pitch = -math.atan2(-d[0], -d[2]), -math.asin(max(-1.0, min(1.0, d[1])))
    qy = (0.0, math.sin(yaw / 2), 0.0, math.cos(yaw / 2))
    qp = (math.sin(pitch / 2), 0.0, 0.0, math.cos(pitch / 2))
    x1, y1, z1, w1 = qp
    x2, y2, z2, w2 = qy
    return (w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
            w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2)


def cornell_cameras(width, height, n=16):
    """n cameras on a 4-wide grid in front of the Cornell box's open side,
    each looking at (0, 1, -0.2)."""
    views = []
    for i in range(n):
        x = -0.5 + (i % 4) / 3.0
        y = 0.7 + (i // 4 % 4) / 4.0
        pos = (x, y, 2.3 + 0.1 * (i // 16))
        views.append(make_view(width, height, position=pos, orientation=look_at_quat(pos, (0.0, 1.0, -0.2))))
    return views


# ------------------------------------------------------------ geometry
def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); normal = cross(v0-v2, v1-v2) side."""
    return np.array([p0, p1, p2, p3], np.float32), np.array([0, 1, 2, 0, 2, 3], np.uint32)


def icosphere(subdiv):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces, np.int64)


def seeded_noise(rng, pts, octaves=4):
    """Sum of seeded random sinusoids, normalised to [-1, 1]."""
    acc = np.zeros(len(pts))
    norm = 0.0
    for o in range(octaves):
        freq = 2.0 ** o * 2.0
        amp = 0.5 ** o
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        ph = rng.uniform(0, 2 * math.pi)
        acc += amp * np.sin(freq * pts @ d * math.pi + ph)
        norm += amp
    return acc / norm


def _orient_outward(verts, faces, center):
    """Make cross(v0-v2, v1-v2) point away from center (the reference's geometric normal)."""
    v0, v1, v2 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    n = np.cross(v0 - v2, v1 - v2)
    c = (v0 + v1 + v2) / 3 - center
    flip = (n * c).sum(1) < 0
    faces = faces.copy()
    faces[flip] = faces[flip][:, [0, 2, 1]]
    return faces


def _tri_normal_ok(verts, faces, want):
    v0, v1, v2 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    n = np.cross(v0 - v2, v1 - v2)
    return faces if (n @ np.asarray(want)).sum() > 0 else faces[:, [0, 2, 1]]


def _random_material(rng, kind=None, min_roughness=0.0):
    """Config material: albedo U(0.2,0.9)^3, roughness U(0.1,0.8) (raised to
    min_roughness); 'lambert' is Disney {metallic 0, specular 0}, 'ggx' is
    Disney {metallic 1}."""
    albedo = rng.uniform(0.2, 0.9, 3)
    rough = max(rng.uniform(0.1, 0.8), min_roughness)
    if kind is None:
        kind = "lambert" if rng.uniform() < 0.5 else "ggx"
    if kind == "lambert":
        return material_record(base_color=albedo, specular=0.0, metallic=0.0, roughness=rough)
    return material_record(base_color=albedo, specular=0.0, metallic=1.0, roughness=rough)


def cornell_box(seed=0, light_radiance=15.0, min_roughness=0.0):
    """Seeded Cornell box: 12 axis-aligned quads (5 walls, ceiling emitter,
    6-sided block) + an icosphere (subdivision 4, 5120 triangles) displaced by
    seeded noise of amplitude 0.05.  Camera at (0, 1, 2.4) looking down -z.
    min_roughness 0.3 keeps every metallic material out of the Disney BSDF's
    singular class (max(specular, metallic) * (1 - roughness) > 0.75,
    disney.hlsli:387), whose vertices make no NEE."""
    rng = np.random.default_rng(seed)
    s = SceneData()
    center = np.array([0, 1, 0], np.float64)
    walls = [  # (quad corners), all facing the box interior
        [(-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)],    # floor
        [(-1, 2, -1), (-1, 2, 1), (1, 2, 1), (1, 2, -1)],    # ceiling
        [(-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)],  # back
        [(-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)],  # left
        [(1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1)],      # right
    ]
    for w in walls:
        p, i = quad(*w)
        f = i.reshape(-1, 3)
        c = p.mean(0)
        f = _tri_normal_ok(p, f, center - c)
        vi, n = s.add_mesh(p, f.reshape(-1))
        s.add_mesh_instance(vi, n, s.add_material(_random_material(rng, min_roughness=min_roughness)))
    # ceiling emitter, facing down
    p, i = quad((-0.3, 1.98, -0.3), (0.3, 1.98, -0.3), (0.3, 1.98, 0.3), (-0.3, 1.98, 0.3))
    f = _tri_normal_ok(p, i.reshape(-1, 3), (0, -1, 0))
    vi, n = s.add_mesh(p, f.reshape(-1))
    s.add_mesh_instance(vi, n, s.add_material(material_record(base_color=(0.8, 0.8, 0.8), specular=0, roughness=1.0,
                                                                emission=(light_radiance,) * 3)))
    # block: 6 quads, outward
    lo, hi = np.array([-0.75, 0.002, -0.6]), np.array([-0.15, 0.7, 0.0])
    bc = (lo + hi) / 2
    corners = lambda a, b, c, d: [tuple(lo + (hi - lo) * np.array(k)) for k in (a, b, c, d)]
    faces6 = [corners((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)), corners((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)),
              corners((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), corners((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)),
              corners((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), corners((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))]
    bp, bi = [], []
    for q in faces6:
        p, i = quad(*q)
        f = _tri_normal_ok(p, i.reshape(-1, 3), p.mean(0) - bc)
        bi.append(f + len(bp) * 4)
        bp.append(p)
    vi, n = s.add_mesh(np.concatenate(bp), np.concatenate(bi).reshape(-1))
    s.add_mesh_instance(vi, n, s.add_material(_random_material(rng, min_roughness=min_roughness)))
    # displaced icosphere
    v, f = icosphere(4)
    v = v * (1 + 0.05 / 0.4 * seeded_noise(rng, v))[:, None]
    v = v * 0.4 + np.array([0.45, 0.45, -0.2])
    f = _orient_outward(v, f, np.array([0.45, 0.45, -0.2]))
    vi, n = s.add_mesh(v.astype(np.float32), f.reshape(-1).astype(np.uint32))
    s.add_mesh_instance(vi, n, s.add_material(_random_material(rng, min_roughness=min_roughness)))
    return s


def cornell_box_smooth(seed=0, subdiv=1, light_radiance=15.0):
    """The seeded Cornell box with its sphere replaced by a coarse icosphere
    (subdivision subdiv) carrying per-vertex normals, so that shading normals
    differ from the geometric normals by up to ~20 degrees."""
    s = cornell_box(seed, light_radiance)
    v, f = icosphere(subdiv)
    n = v / np.linalg.norm(v, axis=1, keepdims=True)
    ctr = np.array([0.45, 0.45, -0.2])
    v = v * 0.4 + ctr
    f = _orient_outward(v, f, ctr)
    vi, nt = s.add_mesh(v.astype(np.float32), f.reshape(-1).astype(np.uint32), normals=n.astype(np.float32))
    a, _ = s.instances[7]
    s.instances[7] = (a, vi | (nt << 16))
    return s


def translation(x, y, z):
    t = np.eye(3, 4, dtype=np.float32)
    t[:, 3] = (x, y, z)
    return t


def cornell_spheres(seed=0, light_radiance=15.0, sphere_light=True):
    """The seeded Cornell box with three procedural sphere instances: a
    diffuse one, a rough anisotropic metal one, and (sphere_light) a small
    emissive one of radiance 8 under the ceiling."""
    s = cornell_box(seed, light_radiance)
    s.add_sphere_instance(s.add_material(material_record(base_color=(0.7, 0.3, 0.2), specular=0.3, roughness=0.6)), 0.25,
                          translation(0.45, 1.3, 0.3))
    s.add_sphere_instance(s.add_material(material_record(base_color=(0.9, 0.8, 0.5), metallic=1.0, roughness=0.45,
                                                         anisotropic=0.6)), 0.2, translation(-0.5, 0.95, 0.25))
    if sphere_light:
        s.add_sphere_instance(s.add_material(material_record(base_color=(0.8, 0.8, 0.8), specular=0.0, roughness=1.0,
                                                             emission=(8.0, 6.0, 4.0))), 0.12, translation(0.0, 1.6, 0.1))
    return s


def checker_image(width, height, cells, c0, c1, rng=None):
    """RGBA32F checkerboard of cells x cells squares alternating c0 and c1,
    with per-texel noise of 0.05 when rng is given."""
    y, x = np.mgrid[0:height, 0:width]
    sel = ((x * cells // width) + (y * cells // height)) % 2
    img = np.where(sel[..., None] == 0, np.asarray(c0, np.float32), np.asarray(c1, np.float32)).astype(np.float32)
    if rng is not None:
        img = img + rng.uniform(-0.05, 0.05, img.shape).astype(np.float32)
    return np.clip(img, 0, 1).astype(np.float32)


def cornell_textured(seed=0):
    """The seeded Cornell box with two textured quads: a floor tile whose base
    colour is a 64x64 checker (texture coordinates repeating twice) and a
    back panel whose roughness/metallic word comes from a 16x32 image
    (material.hlsli:22-45 modulation), plus a textured sphere."""
    rng = np.random.default_rng(seed + 100)
    s = cornell_box(seed)
    base = len(s.images)
    s.images.append(checker_image(64, 64, 8, (0.9, 0.2, 0.1, 1), (0.1, 0.8, 0.3, 1), rng))
    s.images.append(checker_image(16, 32, 4, (1.0, 0.1, 1.0, 1.0), (0.3, 0.9, 0.5, 0.4), rng))
    p, i = quad((-0.9, 0.003, -0.9), (0.9, 0.003, -0.9), (0.9, 0.003, 0.9), (-0.9, 0.003, 0.9))
    f = _tri_normal_ok(p, i.reshape(-1, 3), (0, 1, 0))
    uv = np.array([[0, 0], [2, 0], [2, 2], [0, 2]], np.float32)
    vi, n = s.add_mesh(p, f.reshape(-1), texcoords=uv)
    w = material_record(base_color=(1, 1, 1), specular=0.4, roughness=0.5)
    w[4] = base
    s.add_mesh_instance(vi, n, s.add_material(w))
    p, i = quad((-0.8, 0.2, -0.99), (0.8, 0.2, -0.99), (0.8, 1.6, -0.99), (-0.8, 1.6, -0.99))
    f = _tri_normal_ok(p, i.reshape(-1, 3), (0, 0, 1))
    vi, n = s.add_mesh(p, f.reshape(-1), texcoords=np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32))
    w = material_record(base_color=(0.8, 0.7, 0.6), metallic=1.0, roughness=0.6, anisotropic=0.5, clearcoat=0.5)
    w[6] = base + 1
    w[7] = base + 1
    s.add_mesh_instance(vi, n, s.add_material(w))
    w = material_record(base_color=(1, 1, 1), specular=0.5, roughness=0.4)
    w[4] = base
    s.add_sphere_instance(s.add_material(w), 0.22, translation(0.4, 1.2, 0.3))
    return s


def cornell_alpha(seed=0, cutoff=0.5):
    """The seeded Cornell box with two alpha-masked quads (material word 12 the
    mask image, word 13 the cutoff, material.hlsli:72-76): a panel facing the
    camera (texture coordinates 0..1 over x in [-0.5, 0.5], y in [0.4, 1.4] at
    z = 0) and a grille under the ceiling light (x, z in [-0.4, 0.4] at y =
    1.7, facing down) that casts masked shadows.  The mask is a 32x32 image
    whose channel 0 is a 4x4 checker of 1 and 0 (cell (0, 0) solid) with a
    soft ramp across the grille's cells."""
    s = cornell_box(seed)
    base = len(s.images)
    y, x = np.mgrid[0:32, 0:32]
    mask = np.zeros((32, 32, 4), np.float32)
    mask[..., 0] = (((x // 8) + (y // 8)) % 2 == 0).astype(np.float32)
    mask[..., 3] = 1
    s.images.append(mask)
    ramp = np.zeros((32, 32, 4), np.float32)
    ramp[..., 0] = (x + 0.5) / 32 * (((x // 8) + (y // 8)) % 2 == 0)
    ramp[..., 3] = 1
    s.images.append(ramp)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    for k, (corners, nrm, mat) in enumerate((
            (((-0.5, 0.4, 0.0), (0.5, 0.4, 0.0), (0.5, 1.4, 0.0), (-0.5, 1.4, 0.0)), (0, 0, 1),
             material_record(base_color=(0.9, 0.5, 0.2), specular=0.5, roughness=0.4)),
            (((-0.4, 1.7, -0.4), (0.4, 1.7, -0.4), (0.4, 1.7, 0.4), (-0.4, 1.7, 0.4)), (0, -1, 0),
             material_record(base_color=(0.7, 0.7, 0.7), specular=0.3, roughness=0.8)))):
        p, i = quad(*corners)
        f = _tri_normal_ok(p, i.reshape(-1, 3), nrm)
        vi, n = s.add_mesh(p, f.reshape(-1), texcoords=uv)
        mat[12] = base + k
        mat[13] = int(np.float32(cutoff).view(np.uint32))
        s.add_mesh_instance(vi, n, s.add_material(mat))
    return s


def cornell_view(width=256, height=256):
    return make_view(width, height, position=(0.0, 1.0, 2.4))


def sky_environment(width=1024, height=512, sun_elevation_deg=35.0, sun_azimuth_deg=60.0, sun_peak=1e4,
                    sun_sigma_deg=1.0):
    """Procedural RGBA32F HDR sky: gradient + a Gaussian sun lobe of peak sun_peak.
    Pixel (x, y) is the direction of sphericalUvToCartesian(((x+.5)/W, (y+.5)/H))
    (src/Shaders/compat/common.h:145-150)."""
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    uu, vv = np.meshgrid(u, v)
    phi = (uu * 2 - 1) * math.pi
    th = vv * math.pi
    d = np.stack([np.sin(th) * np.cos(phi), np.cos(th), np.sin(th) * np.sin(phi)], -1)
    up = np.clip(d[..., 1], 0, 1)
    zen = np.array([0.25, 0.45, 0.9])
    hor = np.array([0.9, 0.85, 0.75])
    gnd = np.array([0.12, 0.1, 0.08])
    sky = hor[None, None] * (1 - up[..., None] ** 0.5) + zen[None, None] * up[..., None] ** 0.5
    col = np.where(d[..., 1:2] >= 0, sky, gnd[None, None] * (0.5 + 0.5 * (1 + d[..., 1:2])))
    el, az = math.radians(sun_elevation_deg), math.radians(sun_azimuth_deg)
    sd = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)])
    ang = np.arccos(np.clip(d @ sd, -1, 1))
    sun = sun_peak * np.exp(-0.5 * (ang / math.radians(sun_sigma_deg)) ** 2)
    col = col + sun[..., None] * np.array([1.0, 0.95, 0.85])[None, None]
    img = np.ones((height, width, 4), np.float32)
    img[..., :3] = col
    return img


def uv_sphere(segments=25, bands=11):
    """Closed UV sphere with pole caps: 2*segments + 2*segments*(bands-2) triangles
    (500 for 25 x 11)."""
    lat = np.linspace(0, math.pi, bands + 1)[1:-1]
    lon = np.arange(segments) * 2 * math.pi / segments
    ring = np.stack([np.outer(np.sin(lat), np.cos(lon)), np.repeat(np.cos(lat)[:, None], segments, 1),
                     np.outer(np.sin(lat), np.sin(lon))], -1).reshape(-1, 3)
    verts = np.concatenate([[[0, 1, 0]], ring, [[0, -1, 0]]])
    nr = bands - 1
    faces = []
    for j in range(segments):
        faces.append((0, 1 + (j + 1) % segments, 1 + j))
    for r in range(nr - 1):
        for j in range(segments):
            a, b = 1 + r * segments + j, 1 + r * segments + (j + 1) % segments
            c, d = a + segments, b + segments
            faces += [(a, b, d), (a, d, c)]
    last = len(verts) - 1
    base = 1 + (nr - 1) * segments
    for j in range(segments):
        faces.append((last, base + j, base + (j + 1) % segments))
    return verts, np.array(faces, np.int64)


def spheres_field(seed=1, n_spheres=2000, n_emitters=256, tris_per_emitter=16, env=True):
    """1M triangles: n_spheres seeded displaced UV spheres (500 triangles each)
    on a ground plane, n_emitters emissive instances of tris_per_emitter
    triangles each (radiance lognormal(mu=1, sigma=1)), procedural sky."""
    rng = np.random.default_rng(seed)
    s = SceneData()
    base_v, base_f = uv_sphere()
    assert len(base_f) == 500
    nx = int(math.ceil(math.sqrt(n_spheres * 1.25)))
    nz = int(math.ceil(n_spheres / nx))
    spacing = 1.6
    allp, alli, metas = [], [], []
    voff, ioff = 0, 0
    for k in range(n_spheres):
        gx, gz = k % nx, k // nx
        r = rng.uniform(0.35, 0.6)
        c = np.array([(gx - nx / 2) * spacing + rng.uniform(-0.3, 0.3), r,
                      (gz - nz / 2) * spacing + rng.uniform(-0.3, 0.3)])
        v = base_v * (1 + 0.15 * seeded_noise(rng, base_v, 3))[:, None] * r
        f = _orient_outward(v, base_f, np.zeros(3))
        allp.append(v.astype(np.float32))
        alli.append(f.astype(np.uint32).reshape(-1))
        metas.append((c, voff * 12, ioff * 4, len(f)))
        voff += len(v)
        ioff += f.size
    pb = s.add_buffer(np.concatenate(allp))
    ib = s.add_buffer(np.concatenate(alli))
    for c, po, io, nt in metas:
        vi, n = s.add_mesh_shared(ib, pb, io, po, nt)
        xf = np.eye(3, 4, dtype=np.float32)
        xf[:, 3] = c
        s.add_mesh_instance(vi, n, s.add_material(_random_material(rng)), xf)
    ext = max(nx, nz) * spacing
    p, i = quad((-ext, 0, -ext), (ext, 0, -ext), (ext, 0, ext), (-ext, 0, ext))
    f = _tri_normal_ok(p, i.reshape(-1, 3), (0, 1, 0))
    vi, n = s.add_mesh(p, f.reshape(-1))
    s.add_mesh_instance(vi, n, s.add_material(material_record(base_color=(0.5, 0.5, 0.5), specular=0, roughness=0.9)))
    # emitters: flat 4x2 grids of quads (16 triangles) facing down, above the field
    gx = np.linspace(-0.4, 0.4, 5)
    gz = np.linspace(-0.2, 0.2, 3)
    gp = np.array([(x, 0, z) for z in gz for x in gx], np.float32)
    gf = []
    for zz in range(2):
        for xx in range(4):
            a = zz * 5 + xx
            gf += [(a, a + 1, a + 6), (a, a + 6, a + 5)]
    gf = _tri_normal_ok(gp, np.array(gf), (0, -1, 0))
    assert len(gf) == tris_per_emitter
    evi, en = s.add_mesh(gp, gf.reshape(-1))
    for k in range(n_emitters):
        L = float(np.exp(rng.normal(1.0, 1.0)))
        xf = np.eye(3, 4, dtype=np.float32)
        xf[:, 3] = (rng.uniform(-ext * 0.8, ext * 0.8), rng.uniform(2.5, 4.0), rng.uniform(-ext * 0.8, ext * 0.8))
        s.add_mesh_instance(evi, en, s.add_material(material_record(base_color=(0.8, 0.8, 0.8), specular=0,
                                                                      roughness=1.0, emission=(L, L, L))), xf)
    if env:
        s.set_environment((1, 1, 1), sky_environment())
    return s


def spheres_view(width=1920, height=1080):
    return make_view(width, height, position=(0.0, 7.0, 40.0), orientation=axis_angle_quat((1, 0, 0), -12.0))


def unit_cube_transform(bbox_min, bbox_max):
    """createTransform (src/App/loaders/load_volumes.cpp:18-27): centre the
    volume's world bbox at the origin and scale its largest extent to 1."""
    mn = np.asarray(bbox_min, np.float64)
    mx = np.asarray(bbox_max, np.float64)
    c = (mx + mn) / 2
    s = np.float32(1 / np.float32((mx - mn).max()))
    t = np.eye(3, 4, dtype=np.float32) * s
    t[:, 3] = (-c.astype(np.float32)) * s
    return t


def cloud(seed=0, n=256, sigma_scale=20.0, albedo=0.9, anisotropy=0.3, env=True, env_size=(1024, 512)):
    """Heterogeneous medium workload: an n^3 NanoVDB float grid of seeded
    4-octave fBm thresholded at 0.4, density scale sigma_scale, albedo and
    Henyey-Greenstein g as given, placed in a unit cube at the origin, lit by
    the procedural sky."""
    from . import nanovdb
    sd = SceneData()
    dense = nanovdb.fbm_density(n=n, seed=seed, octaves=4, threshold=0.4)
    grid = nanovdb.write_float_grid(dense, voxel_size=1.0)
    vol = sd.add_volume(grid)
    bb = np.frombuffer(grid[560:608].tobytes(), "<f8")
    med = sd.add_material(medium_record(density=(sigma_scale,) * 3, albedo=(albedo,) * 3, anisotropy=anisotropy,
                                        density_volume=vol))
    sd.add_volume_instance(med, vol, unit_cube_transform(bb[:3], bb[3:]))
    if env:
        sd.set_environment(image=sky_environment(*env_size))
    return sd


def cornell_cloud(seed=0, n=64, sigma_scale=6.0, albedo=0.8, anisotropy=0.2, fbm_threshold=0.3, density=None,
                  light_radiance=15.0):
    """The seeded Cornell box with an n^3 density grid (seeded fBm, or
    `density`, indexed [i, j, k]) filling the box [-0.6, 0.95, -0.7] -
    [0.6, 1.75, 0.3] above the block and the sphere, for differentiable
    rendering: the grid stores every leaf (write_float_grid(all_leaves=True)),
    so each voxel's value address does not depend on the values.  Returns
    (SceneData, value byte offsets of the n^3 voxels)."""
    from . import nanovdb
    sd = cornell_box(seed, light_radiance)
    if density is None:
        density = nanovdb.fbm_density(n=n, seed=seed, octaves=4, threshold=fbm_threshold)
    grid, offs = nanovdb.write_float_grid(np.asarray(density, np.float32), voxel_size=1.0, all_leaves=True,
                                          return_offsets=True)
    add_medium_box(sd, (sigma_scale,) * 3, (albedo,) * 3, anisotropy, (-0.6, 0.95, -0.7), (0.6, 1.75, 0.3), grid=grid)
    return sd, offs


def cloud_view(width=1024, height=1024):
    return make_view(width, height, position=(0.0, 0.0, 1.6))


# ------------------------------------------------------------ media helpers
def layered_grid(values_y):
    """NanoVDB grid of 4 x len(values_y) x 4 voxels, constant in x and z."""
    from . import nanovdb
    d = np.zeros((4, len(values_y), 4), np.float32)
    for j, v in enumerate(values_y):
        d[:, j, :] = v
    return nanovdb.write_float_grid(d, voxel_size=1.0)


def add_medium_box(sd, sigma, albedo, g, box_min, box_max, grid=None):
    """A medium filling the world box [box_min, box_max]: the [-1, 1] cube
    (no grid) or the grid's index box mapped onto it."""
    box_min, box_max = np.asarray(box_min, np.float64), np.asarray(box_max, np.float64)
    if grid is None:
        lo, hi, vol = np.full(3, -1.0), np.full(3, 1.0), 0xFFFFFFFF
    else:
        vol = sd.add_volume(grid)
        bb = np.frombuffer(grid[560:608].tobytes(), "<f8")
        lo, hi = bb[:3], bb[3:]
    s = (box_max - box_min) / (hi - lo)
    t = np.zeros((3, 4), np.float32)
    t[:, :3] = np.diag(s)
    t[:, 3] = box_min - lo * s
    med = sd.add_material(medium_record(density=sigma, albedo=albedo, anisotropy=g, density_volume=vol))
    return sd.add_volume_instance(med, vol, t)


def emissive_inward_cube(sd, half, emission):
    """Closed cube [-half, half]^3 of 12 triangles whose geometric normals
    face inwards, with a black emissive material."""
    c = np.array([[x, y, z] for z in (-half, half) for y in (-half, half) for x in (-half, half)], np.float32)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    idx = []
    for a, b, cc, dd in faces:
        for tri in ((a, b, cc), (a, cc, dd)):
            v = c[list(tri)]
            n = np.cross(v[0] - v[2], v[1] - v[2])
            if n @ (-v.mean(0)) < 0:  # make the geometric normal point into the box
                tri = (tri[0], tri[2], tri[1])
            idx.extend(tri)
    vi, nt = sd.add_mesh(c, np.array(idx, np.uint32))
    # black walls that emit: every path that reaches a wall ends there
    mat = sd.add_material(material_record(base_color=(0, 0, 0), emission=emission))
    sd.add_mesh_instance(vi, nt, mat)
