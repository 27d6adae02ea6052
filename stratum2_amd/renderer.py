"""TestRenderer-shaped wrapper over the C ABI (src/App/TestRenderer.hpp:14-20).

Device memory comes from torch tensors on the HIP device; every compute call
goes through libstratum2.so.  There is no CPU path: a missing library raises.
"""
import ctypes as C

from . import abi


def _torch():
    import torch
    return torch


def _enter(device):
    """Ordering contract of the C ABI: a call waits for the work queued on the
    device's null stream (torch's default stream) before it reads its inputs,
    and its outputs are complete when it returns.  Work on any other torch
    stream is finished here first."""
    torch = _torch()
    cur = torch.cuda.current_stream(device)
    if cur != torch.cuda.default_stream(device):
        cur.synchronize()


def _check_tensor(t, name, shape, dtype, device):
    if (tuple(t.shape) != tuple(shape) or t.dtype != dtype or not t.is_contiguous() or t.device.type != "cuda"
            or t.device.index != device):
        raise ValueError(f"{name} must be a contiguous {tuple(shape)} {dtype} tensor on cuda:{device}")


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

    def sync(self):
        """st_sync: wait for the context stream (ST_FLAG_ASYNC renders)."""
        abi.check(self.lib, self.lib.st_sync(self.h), "st_sync")

    def timings(self):
        out = (C.c_double * 12)()
        abi.check(self.lib, self.lib.st_last_timings(self.h, out), "st_last_timings")
        keys = ["trace_ms", "shade_ms", "shadow_ms", "trace_launches", "shade_launches", "shadow_launches",
                "closest_nodes", "closest_tris", "shadow_nodes", "shadow_tris", "closest_rays", "shadow_rays"]
        return dict(zip(keys, list(out)))

    def kernel_times(self):
        """Per-kernel device ms and launches of the last render
        (st_last_kernel_times); exclusive when rendered with ST_FLAG_SERIAL_WAVES."""
        out = (C.c_double * 10)()
        abi.check(self.lib, self.lib.st_last_kernel_times(self.h, out), "st_last_kernel_times")
        names = ["k_trace_closest", "k_shade", "k_trace_shadow", "k_ris", "k_restir_reuse"]
        return {n: (out[i], int(out[5 + i])) for i, n in enumerate(names)}


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

    def set_triangle_albedo(self, rgb):
        """st_scene_set_triangle_albedo: (triangles, 3) float32 device tensor
        (copied), or None to remove the table."""
        if rgb is None:
            abi.check(self.lib, self.lib.st_scene_set_triangle_albedo(self.h, None, 0), "st_scene_set_triangle_albedo")
            return
        torch = _torch()
        n = rgb.shape[0]
        _check_tensor(rgb, "rgb", (n, 3), torch.float32, self.ctx.device)
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_scene_set_triangle_albedo(self.h, rgb.data_ptr(), n), "st_scene_set_triangle_albedo")

    def set_volume_grid(self, volume, grid):
        """st_scene_set_volume_grid: new values of grid `volume` as a uint8
        device tensor of the grid's byte size and tree layout."""
        torch = _torch()
        _check_tensor(grid, "grid", (grid.numel(),), torch.uint8, self.ctx.device)
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_scene_set_volume_grid(self.h, volume, grid.data_ptr(), grid.numel()),
                  "st_scene_set_volume_grid")

    def adjoint(self, view, params, d_hdr, grad_albedo=None, grad_grid=None, grad_volume=0):
        """st_adjoint: adds the gradient of sum(d_hdr * hdr) to grad_albedo
        ((triangles, 3) float32) and grad_grid (float32, one per 4-byte word of
        the grid); d_hdr is (H, W, 4) float32.  All device tensors."""
        torch = _torch()
        W, H = params.width, params.height
        _check_tensor(d_hdr, "d_hdr", (H, W, 4), torch.float32, self.ctx.device)
        for name, t in (("grad_albedo", grad_albedo), ("grad_grid", grad_grid)):
            if t is not None:
                _check_tensor(t, name, tuple(t.shape), torch.float32, self.ctx.device)
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_adjoint(self.h, C.byref(view), C.byref(params), d_hdr.data_ptr(),
                                                grad_albedo.data_ptr() if grad_albedo is not None else None,
                                                grad_grid.data_ptr() if grad_grid is not None else None, grad_volume),
                  "st_adjoint")

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
        _check_tensor(rays, "rays", (n, 8), torch.float32, self.ctx.device)
        out = torch.empty((n, 4) if closest else (n,), dtype=torch.int32, device=rays.device)  # every entry is written
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_trace_rays(self.h, rays.data_ptr(), n, int(closest), out.data_ptr()),
                  "st_trace_rays")
        return out

    def trace_primary(self, view, width, height):
        """st_trace_primary: (visibility (H, W, 2) int32, t (H, W) float32) device tensors."""
        torch = _torch()
        dev = torch.device("cuda", self.ctx.device)
        vis = torch.empty((height, width, 2), dtype=torch.int32, device=dev)
        t = torch.empty((height, width), dtype=torch.float32, device=dev)
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_trace_primary(self.h, C.byref(view), width, height, vis.data_ptr(), t.data_ptr()),
                  "st_trace_primary")
        return vis, t


class TestRenderer:
    """render(view, params) -> dict of device tensors (hdr, visibility, path_length, ray_counts)."""

    def __init__(self, scene):
        self.scene = scene

    def render(self, view, params, want_path_length=False, out=None, want_aux=False):
        res, o = self.outputs(params, want_path_length, out, want_aux)
        _enter(self.scene.ctx.device)
        abi.check(self.scene.lib, self.scene.lib.st_render(self.scene.h, C.byref(view), C.byref(params), C.byref(o)),
                  "st_render")
        return res

    def outputs(self, params, want_path_length=False, out=None, want_aux=False):
        """Output tensors for a render with params (or the caller's, checked) and their st_outputs."""
        torch = _torch()
        dev = torch.device("cuda", self.scene.ctx.device)
        W, H, S = params.width, params.height, params.spp
        layout = {"hdr": ((H, W, 4), torch.float32), "visibility": ((H, W, 2), torch.int32),
                  "ray_counts": ((2,), torch.int64), "path_length": ((S, H, W), torch.int32),
                  "albedo": ((H, W, 4), torch.float32), "depth": ((H, W, 4), torch.float32),
                  "prev_uv": ((H, W, 2), torch.float32)}
        res = out or {}
        if "hdr" not in res:
            res["hdr"] = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
            res["visibility"] = torch.empty((H, W, 2), dtype=torch.int32, device=dev)
            res["ray_counts"] = torch.empty(2, dtype=torch.int64, device=dev)
            res["path_length"] = torch.empty((S, H, W), dtype=torch.int32, device=dev) if want_path_length else None
            if want_aux:
                res["albedo"] = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
                res["depth"] = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
                res["prev_uv"] = torch.empty((H, W, 2), dtype=torch.float32, device=dev)
        else:  # caller-provided outputs: the kernels write them in place, so they must fit exactly
            for k, t in res.items():
                if k not in layout:
                    raise ValueError(f"unknown output '{k}'")
                if t is not None:
                    _check_tensor(t, f"out['{k}']", *layout[k], self.scene.ctx.device)

        def ptr(k):
            return res[k].data_ptr() if res.get(k) is not None else None
        o = abi.st_outputs(ptr("hdr"), ptr("visibility"), ptr("path_length"), ptr("ray_counts"), ptr("albedo"), ptr("depth"),
                           ptr("prev_uv"))
        return res, o


class Batch:
    """Many renders in one graph launch (st_batch_*): requests is a list of
    (scene, view, params); each gets output tensors of its own, the same as
    TestRenderer.render would return for it alone."""

    def __init__(self, ctx, requests, want_path_length=False):
        torch = _torch()
        self.ctx, self.lib = ctx, ctx.lib
        self.outputs = []
        arr = (abi.st_render_request * len(requests))()
        dev = torch.device("cuda", ctx.device)
        for i, (scene, view, p) in enumerate(requests):
            if scene.ctx is not ctx:
                raise ValueError("every request's scene must belong to this context")
            W, H, S = p.width, p.height, p.spp
            o = {"hdr": torch.empty((H, W, 4), dtype=torch.float32, device=dev),
                 "visibility": torch.empty((H, W, 2), dtype=torch.int32, device=dev),
                 "ray_counts": torch.empty(2, dtype=torch.int64, device=dev),
                 "path_length": torch.empty((S, H, W), dtype=torch.int32, device=dev) if want_path_length else None}
            self.outputs.append(o)
            arr[i].scene = scene.h
            arr[i].view = view
            arr[i].params = p
            pl = o["path_length"]
            arr[i].outputs = abi.st_outputs(o["hdr"].data_ptr(), o["visibility"].data_ptr(),
                                            pl.data_ptr() if pl is not None else None, o["ray_counts"].data_ptr(),
                                            None, None, None)
        self._scenes = [r[0] for r in requests]  # kept alive as long as the batch
        self.h = C.c_void_p()
        abi.check(self.lib, self.lib.st_batch_create(ctx.h, arr, len(requests), C.byref(self.h)), "st_batch_create")

    def render(self):
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_batch_render(self.h), "st_batch_render")
        return self.outputs

    def autotune(self, repeats=2):
        """st_batch_autotune: switched off in the library (raises)."""
        _enter(self.ctx.device)
        cfg = abi.st_batch_config()
        abi.check(self.lib, self.lib.st_batch_autotune(self.h, repeats, C.byref(cfg)), "st_batch_autotune")
        return self.config()

    def config(self):
        cfg = abi.st_batch_config()
        abi.check(self.lib, self.lib.st_batch_get_config(self.h, C.byref(cfg)), "st_batch_get_config")
        return {k: getattr(cfg, k) for k, _ in cfg._fields_}

    def set_config(self, **kw):
        cfg = abi.st_batch_config()
        abi.check(self.lib, self.lib.st_batch_get_config(self.h, C.byref(cfg)), "st_batch_get_config")
        for k, v in kw.items():
            setattr(cfg, k, v)
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_batch_set_config(self.h, C.byref(cfg)), "st_batch_set_config")

    def close(self):
        if self.h:
            self.lib.st_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Denoiser:
    """The reference's Denoiser node (src/App/Denoiser.hpp:14-59) over
    st_denoise: temporal accumulation with reprojection, then optional
    variance estimation and edge-aware filter passes.  Settings are the
    reference's members (defaults from Denoiser.hpp and Denoiser.cpp:24-31)."""

    def __init__(self, ctx, width, height, **settings):
        self.ctx, self.lib = ctx, ctx.lib
        self.width, self.height = width, height
        self.params = abi.default_denoise_params(**settings)
        self.h = C.c_void_p()
        abi.check(self.lib, self.lib.st_denoiser_create(ctx.h, width, height, C.byref(self.h)), "st_denoiser_create")

    def close(self):
        if self.h:
            self.lib.st_denoiser_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset_accumulation(self):
        abi.check(self.lib, self.lib.st_denoiser_reset(self.h), "st_denoiser_reset")

    @property
    def accumulated_frames(self):
        return int(self.lib.st_denoiser_accumulated_frames(self.h))

    def timings(self):
        """Per-pass GPU ms and launch counts of the last denoise call."""
        out = (C.c_double * 8)()
        abi.check(self.lib, self.lib.st_denoiser_last_timings(self.h, out), "st_denoiser_last_timings")
        names = ["accumulate", "variance", "atrous", "copy_rgb"]
        return {**{f"{n}_ms": out[i] for i, n in enumerate(names)},
                **{f"{n}_launches": int(out[4 + i]) for i, n in enumerate(names)}}

    @property
    def demodulate_albedo(self):
        return bool(self.params.flags & abi.ST_DENOISE_DEMODULATE_ALBEDO)

    def denoise(self, view, frame, out=None, instance_index_map=None):
        """frame: the dict TestRenderer.render returns with want_aux=True
        (hdr, albedo, prev_uv, visibility, depth).  Returns the output image
        (a new device tensor unless `out` is given)."""
        torch = _torch()
        H, W = self.height, self.width
        for k, c in (("hdr", 4), ("albedo", 4), ("prev_uv", 2), ("visibility", 2), ("depth", 4)):
            t = frame.get(k)
            if t is None:
                raise ValueError(f"denoise needs frame['{k}'] (render with want_aux=True)")
            if tuple(t.shape) != (H, W, c) or not t.is_contiguous() or t.device.type != "cuda":
                raise ValueError(f"frame['{k}'] must be a contiguous ({H}, {W}, {c}) device tensor")
        if out is None:
            out = torch.empty((H, W, 4), dtype=torch.float32, device=frame["hdr"].device)
        m = instance_index_map
        if m is not None and (m.dtype != torch.int32 or m.device.type != "cuda" or not m.is_contiguous()):
            raise ValueError("instance_index_map must be a contiguous int32 device tensor")
        _check_tensor(out, "out", (H, W, 4), torch.float32, self.ctx.device)
        _enter(self.ctx.device)
        inp = abi.st_denoise_inputs(frame["hdr"].data_ptr(), frame["albedo"].data_ptr(), frame["prev_uv"].data_ptr(),
                                    frame["visibility"].data_ptr(), frame["depth"].data_ptr(),
                                    m.data_ptr() if m is not None else None, m.numel() if m is not None else 0)
        abi.check(self.lib, self.lib.st_denoise(self.h, C.byref(view), C.byref(inp), C.byref(self.params),
                                                out.data_ptr()), "st_denoise")
        return out
