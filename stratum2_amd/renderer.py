"""TestRenderer-shaped wrapper over the C ABI (src/App/TestRenderer.hpp:14-20).

Device memory comes from torch tensors on the HIP device; every compute call
goes through libstratum2.so.  There is no CPU path: a missing library raises.
"""
import ctypes as C

from . import abi


def _torch():
    import torch
    return torch


class Context:
    def __init__(self, device=0):
        self.lib = abi.load_library()
        self.h = C.c_void_p()
        abi.check(self.lib, self.lib.st_context_create(device, C.byref(self.h)), "st_context_create")
        self.device = device

    def close(self):
        if self.h:
            self.lib.st_context_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timings(self):
        out = (C.c_double * 12)()
        abi.check(self.lib, self.lib.st_last_timings(self.h, out), "st_last_timings")
        keys = ["trace_ms", "shade_ms", "shadow_ms", "trace_launches", "shade_launches", "shadow_launches",
                "closest_nodes", "closest_tris", "shadow_nodes", "shadow_tris", "closest_rays", "shadow_rays"]
        return dict(zip(keys, list(out)))


class Scene:
    """Device copy of a SceneData (Scene::updateFrameData + BVH build)."""

    def __init__(self, ctx, scene_data):
        self.ctx = ctx
        self.lib = ctx.lib
        self._desc = scene_data.desc()
        self._data = scene_data
        self.h = C.c_void_p()
        abi.check(self.lib, self.lib.st_scene_create(ctx.h, C.byref(self._desc), C.byref(self.h)), "st_scene_create")

    def update(self, transforms):
        """st_scene_update: new (instance_count, 3, 4) transforms, BVH rebuilt on
        the GPU.  Returns the device build time in ms."""
        import numpy as np
        xf = np.ascontiguousarray(transforms, np.float32)
        ms = C.c_double(0)
        abi.check(self.lib, self.lib.st_scene_update(self.h, xf.ctypes.data, C.byref(ms)), "st_scene_update")
        return ms.value

    def stats(self):
        out = (C.c_uint64 * 4)()
        abi.check(self.lib, self.lib.st_scene_stats(self.h, out), "st_scene_stats")
        return dict(nodes=out[0], triangles=out[1], depth=out[2], bytes=out[3])

    def close(self):
        if self.h:
            self.lib.st_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_rays(self, rays, closest=True):
        """rays: (N, 8) float32 device tensor {o.xyz, tmax, d.xyz, 0}."""
        torch = _torch()
        n = rays.shape[0]
        out = torch.zeros((n, 4) if closest else (n,), dtype=torch.int32, device=rays.device)
        abi.check(self.lib, self.lib.st_trace_rays(self.h, rays.data_ptr(), n, int(closest), out.data_ptr()),
                  "st_trace_rays")
        return out

    def trace_primary(self, view, width, height):
        """st_trace_primary: (visibility (H, W, 2) int32, t (H, W) float32) device tensors."""
        torch = _torch()
        dev = torch.device("cuda", self.ctx.device)
        vis = torch.empty((height, width, 2), dtype=torch.int32, device=dev)
        t = torch.empty((height, width), dtype=torch.float32, device=dev)
        abi.check(self.lib, self.lib.st_trace_primary(self.h, C.byref(view), width, height, vis.data_ptr(), t.data_ptr()),
                  "st_trace_primary")
        return vis, t


class TestRenderer:
    """render(view, params) -> dict of device tensors (hdr, visibility, path_length, ray_counts)."""

    def __init__(self, scene):
        self.scene = scene

    def render(self, view, params, want_path_length=False, out=None, want_aux=False):
        torch = _torch()
        dev = torch.device("cuda", self.scene.ctx.device)
        W, H, S = params.width, params.height, params.spp
        res = out or {}
        if "hdr" not in res:
            res["hdr"] = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
            res["visibility"] = torch.empty((H, W, 2), dtype=torch.int32, device=dev)
            res["ray_counts"] = torch.zeros(2, dtype=torch.int64, device=dev)
            res["path_length"] = torch.empty((S, H, W), dtype=torch.int32, device=dev) if want_path_length else None
            if want_aux:
                res["albedo"] = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
                res["depth"] = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
                res["prev_uv"] = torch.empty((H, W, 2), dtype=torch.float32, device=dev)

        def ptr(k):
            return res[k].data_ptr() if res.get(k) is not None else None
        o = abi.st_outputs(ptr("hdr"), ptr("visibility"), ptr("path_length"), ptr("ray_counts"), ptr("albedo"), ptr("depth"),
                           ptr("prev_uv"))
        abi.check(self.scene.lib, self.scene.lib.st_render(self.scene.h, C.byref(view), C.byref(params), C.byref(o)),
                  "st_render")
        return res
