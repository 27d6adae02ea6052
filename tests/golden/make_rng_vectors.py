"""Writes tests/golden/rng.json: pcg4d / pcg / xxhash32 known answers computed
from the formulas of src/Shaders/common/rng.hlsli:5-35 in Python integers."""
import json
import os

M = 0xFFFFFFFF


def pcg4d(v):
    x, y, z, w = [(a * 1664525 + 1013904223) & M for a in v]
    x = (x + y * w) & M; y = (y + z * x) & M; z = (z + x * y) & M; w = (w + y * z) & M
    x ^= x >> 16; y ^= y >> 16; z ^= z >> 16; w ^= w >> 16
    x = (x + y * w) & M; y = (y + z * x) & M; z = (z + x * y) & M; w = (w + y * z) & M
    return [x, y, z, w]


def pcg(v):
    s = (v * 747796405 + 2891336453) & M
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & M
    return (w >> 22) ^ w


def xxhash32(p):
    h = (p + 374761393) & M
    h = (668265263 * (((h << 17) | (h >> 15)) & M)) & M
    h = (2246822519 * (h ^ (h >> 15))) & M
    h = (3266489917 * (h ^ (h >> 13))) & M
    return h ^ (h >> 16)


ins = [[0, 0, 0, 1], [0, 0, 0, 2], [1, 2, 3, 4], [255, 127, 0, 1], [1919, 1079, 63, 9], [M, M, M, M]]
out = {"pcg4d": [{"in": i, "out": pcg4d(i)} for i in ins],
       "pcg": [{"in": v, "out": pcg(v)} for v in [0, 1, 12345, M]],
       "xxhash32": [{"in": v, "out": xxhash32(v)} for v in [0, 1, 12345, M]]}
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng.json"), "w"), indent=1)
