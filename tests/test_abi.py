"""The C-ABI library loads and exports every entry point include/stratum2.h declares."""
import ctypes as C
import os
import re

from stratum2_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "stratum2.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(st_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    assert "st_render" in syms and "st_scene_create" in syms
    assert set(syms) == set(abi.EXPORTS)


def test_library_exports_every_symbol():
    lib = abi.load_library()
    for s in declared_symbols():
        assert hasattr(lib, s), s
        assert C.cast(getattr(lib, s), C.c_void_p).value


def test_struct_layouts_match_reference_sizes():
    # InstanceData 8 B, TransformData 48 B, MeshVertexInfo 32 B, VolumeInfo 32 B,
    # ProjectionData 32 B, ViewData 48 B (+ camera TransformData)
    assert C.sizeof(abi.st_instance) == 8
    assert C.sizeof(abi.st_transform) == 48
    assert C.sizeof(abi.st_mesh_vertex_info) == 32
    assert C.sizeof(abi.st_volume_info) == 32
    assert C.sizeof(abi.st_projection) == 32
    assert C.sizeof(abi.st_view) == 48 + 48


def test_errors_without_gpu_are_reported_not_crashing():
    lib = abi.load_library()
    assert lib.st_render(None, None, None, None) != 0
    assert lib.st_last_error()
