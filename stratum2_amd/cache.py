"""The learned radiance cache over the C ABI (st_cache_*, st_render_cached):
queries and targets are torch device tensors; parameters come back as numpy
arrays."""
import ctypes as C

import numpy as np

from . import abi
from .renderer import TestRenderer, _check_tensor, _enter, _torch


class RadianceCache:
    def __init__(self, ctx, bbox_min, bbox_max, **config):
        self.ctx, self.lib = ctx, ctx.lib
        self.config = abi.default_cache_config(**config)
        self.h = C.c_void_p()
        f3 = C.c_float * 3
        abi.check(self.lib, self.lib.st_cache_create(ctx.h, C.byref(self.config), f3(*bbox_min), f3(*bbox_max), C.byref(self.h)),
                  "st_cache_create")
        t, w = C.c_uint32(), C.c_uint32()
        abi.check(self.lib, self.lib.st_cache_param_counts(self.h, C.byref(t), C.byref(w)), "st_cache_param_counts")
        self.table_floats, self.weight_floats = t.value, w.value

    def close(self):
        if self.h:
            self.lib.st_cache_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _queries(self, q):
        torch = _torch()
        _check_tensor(q, "queries", (q.shape[0], 12), torch.float32, self.ctx.device)
        return q.shape[0]

    def infer(self, queries):
        """(n, 12) float32 queries -> (n, 3) radiance."""
        torch = _torch()
        n = self._queries(queries)
        out = torch.empty((n, 3), dtype=torch.float32, device=queries.device)
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_cache_infer(self.h, queries.data_ptr(), n, out.data_ptr()), "st_cache_infer")
        return out

    def train(self, queries, targets):
        """One Adam step; returns the mean relative L2 loss before it."""
        torch = _torch()
        n = self._queries(queries)
        _check_tensor(targets, "targets", (n, 3), torch.float32, self.ctx.device)
        loss = C.c_float()
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_cache_train(self.h, queries.data_ptr(), targets.data_ptr(), n, C.byref(loss)),
                  "st_cache_train")
        return loss.value

    def params(self, grads=False):
        """(table (levels * entries, 2), weights (4, 64, 64)[, their last gradients])."""
        t = np.zeros(self.table_floats, np.float32)
        w = np.zeros(self.weight_floats, np.float32)
        tg = np.zeros_like(t) if grads else None
        wg = np.zeros_like(w) if grads else None
        abi.check(self.lib, self.lib.st_cache_get_params(self.h, t.ctypes.data, w.ctypes.data,
                                                         tg.ctypes.data if grads else None, wg.ctypes.data if grads else None),
                  "st_cache_get_params")
        out = (t.reshape(-1, 2), w.reshape(4, 64, 64))
        return out + ((tg.reshape(-1, 2), wg.reshape(4, 64, 64)) if grads else ())

    def set_params(self, table=None, weights=None):
        t = np.ascontiguousarray(table, np.float32) if table is not None else None
        w = np.ascontiguousarray(weights, np.float32) if weights is not None else None
        abi.check(self.lib, self.lib.st_cache_set_params(self.h, t.ctypes.data if t is not None else None,
                                                         w.ctypes.data if w is not None else None), "st_cache_set_params")

    def render(self, scene, view, params, cache_depth=2, train_stride=16, train_steps=1, frame=0, want_train=False,
               want_path_length=False, out=None):
        """st_render_cached: the path tracer with paths ending in this cache at
        path length cache_depth, one training path per train_stride paths, then
        train_steps Adam steps on the training rows.  Returns the render's
        output tensors plus "loss" (mean loss of the first step, or None) and,
        with want_train, "train_queries" (rows, 12) and "train_targets"
        (rows, 4: target rgb, weight)."""
        torch = _torch()
        if scene.ctx is not self.ctx:
            raise ValueError("the scene and the cache belong to different contexts")
        res, o = TestRenderer(scene).outputs(params, want_path_length, out)
        rows = (params.width * params.height * params.spp + train_stride - 1) // max(1, train_stride)
        dev = torch.device("cuda", self.ctx.device)
        tq = torch.empty((rows, 12), dtype=torch.float32, device=dev) if want_train else None
        tt = torch.empty((rows, 4), dtype=torch.float32, device=dev) if want_train else None
        loss = C.c_float(0)
        cp = abi.st_cache_render_params(cache_depth, train_stride, train_steps, frame, tq.data_ptr() if want_train else None,
                                        tt.data_ptr() if want_train else None, C.pointer(loss))
        _enter(self.ctx.device)
        abi.check(self.lib, self.lib.st_render_cached(scene.h, self.h, C.byref(view), C.byref(params), C.byref(cp), C.byref(o)),
                  "st_render_cached")
        res["loss"] = loss.value if train_steps else None
        if want_train:
            res["train_queries"], res["train_targets"] = tq, tt
        return res

    def last_times(self):
        """(inference + target ms, their launches, training ms, steps, cache queries,
        training rows) of the last render."""
        out = (C.c_double * 6)()
        abi.check(self.lib, self.lib.st_cache_last_times(self.ctx.h, out), "st_cache_last_times")
        return tuple(out)
