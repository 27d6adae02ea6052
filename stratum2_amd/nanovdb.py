"""Writer of NanoVDB 32.3 float grids (the byte layout read by
extern/nanovdb/PNanoVDB.h and by the gfx950 reader in csrc/st_volume.h).

Layout written, in this order and 32-byte aligned: GridData (672 B),
TreeData (64 B), RootData (64 B + 32 B per tile), upper internal nodes
(32^3 children of 4096^3 voxels, 270400 B), lower internal nodes (16^3
children of 128^3 voxels, 33856 B), leaf nodes (8^3 voxels, 2144 B).
Offsets, masks and tables follow the PNANOVDB_*_OFF_* constants of
PNanoVDB.h:702-1111 and the float row of pnanovdb_grid_type_constants
(PNanoVDB.h:1157-1160).  Voxels equal to the background are inactive and
leaves without active voxels are not stored (their parent holds a tile).
"""
import struct

import numpy as np

GRID_SIZE, TREE_SIZE, ROOT_SIZE, TILE_SIZE = 672, 64, 64, 32
UPPER_SIZE, LOWER_SIZE, LEAF_SIZE = 270400, 33856, 2144
MAGIC = 0x304244566F6E614E  # "NanoVDB0"
VERSION = (32 << 21) | (3 << 10) | 3
GRID_TYPE_FLOAT, GRID_CLASS_FOG = 1, 2


def _mask_bytes(bits):
    return np.packbits(bits.astype(np.uint8), bitorder="little").tobytes()


def write_float_grid(dense, voxel_size=1.0, background=0.0, name="density", all_leaves=False, return_offsets=False):
    """dense: float32 array indexed [i, j, k] with 0 <= i, j, k < 4096.
    Returns the grid as a uint8 numpy array.  all_leaves: store every leaf
    that covers the dense array, active or not, so that the tree layout (and
    each voxel's value address) does not depend on the values.
    return_offsets: also return an int64 array of dense's shape holding the
    byte offset of each voxel's value in the grid (-1 where no leaf holds
    it)."""
    dense = np.asarray(dense, np.float32)
    nx, ny, nz = dense.shape
    assert max(nx, ny, nz) <= 4096
    px, py, pz = [(-n) % 128 for n in dense.shape]
    d = np.pad(dense, ((0, px), (0, py), (0, pz)), constant_values=background)
    active = d != background
    LX, LY, LZ = [s // 8 for s in d.shape]
    leaf_active = active.reshape(LX, 8, LY, 8, LZ, 8).any(axis=(1, 3, 5))
    if all_leaves:
        leaf_active[:(nx + 7) // 8, :(ny + 7) // 8, :(nz + 7) // 8] = True
    # lower nodes: 128^3 voxels = 16^3 leaves
    WX, WY, WZ = LX // 16, LY // 16, LZ // 16
    lower_has = leaf_active.reshape(WX, 16, WY, 16, WZ, 16).any(axis=(1, 3, 5))
    lowers = [tuple(c) for c in np.argwhere(lower_has)]  # lower node coords (units of 128)
    leaves = []  # (lower idx, leaf coord in leaves)
    for li, (a, b, c) in enumerate(lowers):
        sub = leaf_active[16 * a:16 * a + 16, 16 * b:16 * b + 16, 16 * c:16 * c + 16]
        for (x, y, z) in np.argwhere(sub):
            leaves.append((li, (16 * a + x, 16 * b + y, 16 * c + z)))
    n_upper = 1 if lowers else 0
    n_tiles = 1 if lowers else 0
    root_bytes = ROOT_SIZE + TILE_SIZE * n_tiles
    off_tree = GRID_SIZE
    off_root = off_tree + TREE_SIZE
    off_upper = off_root + root_bytes
    off_lower = off_upper + UPPER_SIZE * n_upper
    off_leaf = off_lower + LOWER_SIZE * len(lowers)
    total = off_leaf + LEAF_SIZE * len(leaves)
    buf = bytearray(total)

    def put(fmt, off, *v):
        struct.pack_into("<" + fmt, buf, off, *v)

    vmax = float(d[active].max()) if active.any() else background
    vmin = float(d[active].min()) if active.any() else background
    # ---- leaves
    leaf_index = {}
    for n, (li, (x, y, z)) in enumerate(leaves):
        o = off_leaf + n * LEAF_SIZE
        leaf_index[(x, y, z)] = o
        vals = d[8 * x:8 * x + 8, 8 * y:8 * y + 8, 8 * z:8 * z + 8]
        act = active[8 * x:8 * x + 8, 8 * y:8 * y + 8, 8 * z:8 * z + 8]
        put("3i", o, 8 * x, 8 * y, 8 * z)
        put("I", o + 12, 7 | 7 << 8 | 7 << 16)  # bbox dif (full) and flags
        buf[o + 16:o + 80] = _mask_bytes(act.reshape(-1))  # offset n = i*64 + j*8 + k
        av = vals[act]
        put("4f", o + 80, float(av.min()), float(av.max()), float(av.mean()), float(av.std()))
        buf[o + 96:o + 96 + 2048] = vals.reshape(-1).astype("<f4").tobytes()
    # ---- lower nodes
    for n, (a, b, c) in enumerate(lowers):
        o = off_lower + n * LOWER_SIZE
        put("3i", o, 128 * a, 128 * b, 128 * c)
        put("3i", o + 12, 128 * a + 127, 128 * b + 127, 128 * c + 127)
        child = np.zeros(4096, bool)
        vmask = np.zeros(4096, bool)
        table = np.zeros(4096, np.int64)
        for x in range(16):
            for y in range(16):
                for z in range(16):
                    key = (16 * a + x, 16 * b + y, 16 * c + z)
                    t = x * 256 + y * 16 + z
                    if key in leaf_index:
                        child[t] = True
                        table[t] = leaf_index[key] - o
        buf[o + 32:o + 544] = _mask_bytes(vmask)
        buf[o + 544:o + 1056] = _mask_bytes(child)
        put("4f", o + 1056, vmin, vmax, 0.0, 0.0)
        tbl = np.zeros(4096, dtype=[("v", "<i8")])
        tbl["v"] = table
        # tile entries hold the background value in their first 4 bytes
        raw = bytearray(tbl.tobytes())
        bg = struct.pack("<f", background)
        for t in np.where(~child)[0]:
            raw[8 * t:8 * t + 4] = bg
        buf[o + 1088:o + 1088 + 8 * 4096] = raw
    # ---- upper node
    if n_upper:
        o = off_upper
        put("3i", o, 0, 0, 0)
        put("3i", o + 12, 4095, 4095, 4095)
        child = np.zeros(32768, bool)
        raw = bytearray(8 * 32768)
        bg = struct.pack("<f", background)
        lower_index = {l: off_lower + n * LOWER_SIZE for n, l in enumerate(lowers)}
        for t in range(32768):
            x, y, z = t >> 10, (t >> 5) & 31, t & 31
            if (x, y, z) in lower_index:
                child[t] = True
                struct.pack_into("<q", raw, 8 * t, lower_index[(x, y, z)] - o)
            else:
                raw[8 * t:8 * t + 4] = bg
        buf[o + 32:o + 4128] = _mask_bytes(np.zeros(32768, bool))
        buf[o + 4128:o + 8224] = _mask_bytes(child)
        put("4f", o + 8224, vmin, vmax, 0.0, 0.0)
        buf[o + 8256:o + 8256 + 8 * 32768] = raw
    # ---- root
    o = off_root
    bbmin = np.argwhere(active).min(0) if active.any() else np.zeros(3, int)
    bbmax = np.argwhere(active).max(0) if active.any() else np.zeros(3, int)
    put("3i", o, *[int(v) for v in bbmin])
    put("3i", o + 12, *[int(v) for v in bbmax])
    put("I", o + 24, n_tiles)
    put("5f", o + 28, background, vmin, vmax, 0.0, 0.0)
    if n_tiles:
        put("Q", o + 64, 0)                    # key of the tile holding ijk in [0, 4096)
        put("q", o + 64 + 8, off_upper - off_root)
        put("I", o + 64 + 16, 1)
        put("f", o + 64 + 20, background)
    # ---- tree
    o = off_tree
    put("4Q", o, off_leaf - off_tree if leaves else 0, off_lower - off_tree if lowers else 0,
        off_upper - off_tree if n_upper else 0, off_root - off_tree)
    put("3I", o + 32, len(leaves), len(lowers), n_upper)
    put("3I", o + 44, 0, 0, 0)
    put("Q", o + 56, int(active.sum()))
    # ---- grid
    o = 0
    put("Q", o, MAGIC)
    put("Q", o + 8, 0)
    put("I", o + 16, VERSION)
    put("I", o + 20, 0)
    put("I", o + 24, 0)
    put("I", o + 28, 1)
    put("Q", o + 32, total)
    nb = name.encode()[:255]
    buf[o + 40:o + 40 + len(nb)] = nb
    vs = float(voxel_size)
    m = o + 296
    put("9f", m, vs, 0, 0, 0, vs, 0, 0, 0, vs)
    put("9f", m + 36, 1 / vs, 0, 0, 0, 1 / vs, 0, 0, 0, 1 / vs)
    put("3f", m + 72, 0, 0, 0)
    put("f", m + 84, 1.0)
    put("9d", m + 88, vs, 0, 0, 0, vs, 0, 0, 0, vs)
    put("9d", m + 160, 1 / vs, 0, 0, 0, 1 / vs, 0, 0, 0, 1 / vs)
    put("3d", m + 232, 0, 0, 0)
    put("d", m + 256, 1.0)
    put("6d", o + 560, *(bbmin * vs).tolist(), *((bbmax + 1) * vs).tolist())
    put("3d", o + 608, vs, vs, vs)
    put("I", o + 632, GRID_CLASS_FOG)
    put("I", o + 636, GRID_TYPE_FLOAT)
    put("q", o + 640, 0)
    put("I", o + 648, 0)
    grid = np.frombuffer(bytes(buf), np.uint8).copy()
    if not return_offsets:
        return grid
    offs = np.full(d.shape, -1, np.int64)
    local = 96 + 4 * np.arange(512, dtype=np.int64).reshape(8, 8, 8)  # n = i*64 + j*8 + k
    for (x, y, z), o in leaf_index.items():
        offs[8 * x:8 * x + 8, 8 * y:8 * y + 8, 8 * z:8 * z + 8] = o + local
    return grid, offs[:nx, :ny, :nz].copy()


ROOT_OFFSET = GRID_SIZE + TREE_SIZE  # where write_float_grid puts the root node
ROOT_MAX_OFFSET = ROOT_OFFSET + 36   # the root's maximum value: the tracking majorant


def fbm_density(n=256, seed=0, octaves=4, threshold=0.4):
    """Seeded 4-octave fBm of value noise in [0, 1] on an n^3 grid, values
    below `threshold` set to 0 (inactive)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    acc = np.zeros((n, n, n), np.float32)
    amp, tot = 1.0, 0.0
    for o in range(octaves):
        f = 4 * 2 ** o
        lat = rng.uniform(0, 1, (f + 1, f + 1, f + 1)).astype(np.float32)
        g = x * f
        i0 = np.minimum(g.astype(int), f - 1)
        t = (g - i0).astype(np.float32)
        t = t * t * (3 - 2 * t)
        # trilinear interpolation of the lattice
        c = np.zeros((n, n, n), np.float32)
        for dx in (0, 1):
            wx = (t if dx else 1 - t)[:, None, None]
            for dy in (0, 1):
                wy = (t if dy else 1 - t)[None, :, None]
                for dz in (0, 1):
                    wz = (t if dz else 1 - t)[None, None, :]
                    c += wx * wy * wz * lat[np.ix_(i0 + dx, i0 + dy, i0 + dz)]
        acc += amp * c
        tot += amp
        amp *= 0.5
    acc /= tot
    lo, hi = acc.min(), acc.max()
    acc = (acc - lo) / (hi - lo)
    acc[acc < threshold] = 0
    return acc.astype(np.float32)
