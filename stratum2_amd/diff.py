"""Differentiable rendering as a torch.autograd.Function over the C ABI.

The parameters are a per-triangle albedo ((triangles, 3), replacing the base
colour of Lambertian camera-path vertices) and the values of one density
grid (a dense (n, n, n) tensor whose voxels map to fixed value slots of a
NanoVDB grid written with every leaf stored).  forward() uploads them
(st_scene_set_triangle_albedo, st_scene_set_volume_grid) and renders with
st_render; backward() runs st_adjoint, the path-replay adjoint of the same
estimator (fp32 NEE: ST_FLAG_FLOAT_NEE is added to the render flags).  The
reference has no derivatives; this is the differentiable-rendering extension
of its camera-path estimator.
"""
import torch

from . import abi, nanovdb
from .renderer import TestRenderer


class DiffScene:
    """A Scene plus the layout of its differentiable parameters.

    grid_offsets: the (n, n, n) int64 byte offsets of the voxels' values in
    grid `volume` (scenes.cornell_cloud returns them), or None when the
    scene's density is not a parameter.  The majorant written with each
    upload is majorant_scale times the largest value: a majorant above every
    value keeps null collisions possible everywhere, which the gradient of a
    voxel at the maximum needs."""

    def __init__(self, scene, grid_offsets=None, volume=0, majorant_scale=1.1):
        self.scene, self.volume = scene, volume
        dev = torch.device("cuda", scene.ctx.device)
        self.dev = dev
        self.triangles = int(scene.stats()["triangles"])
        self.grid = None
        if grid_offsets is not None:
            host = scene._data.volumes[volume]
            self.grid = torch.from_numpy(host.copy()).to(dev)
            self.words = self.grid.numel() // 4
            self.slots = torch.as_tensor(grid_offsets.reshape(-1) // 4, device=dev)
            self.shape = tuple(grid_offsets.shape)
            self.majorant_scale = float(majorant_scale)

    def upload(self, albedo=None, density=None):
        if albedo is not None:
            self.scene.set_triangle_albedo(albedo.detach().clamp(0, 1).contiguous())
        if density is not None:
            d = density.detach().clamp(min=0).reshape(-1).to(torch.float32)
            f = self.grid.view(torch.float32)
            f[self.slots] = d
            f[nanovdb.ROOT_MAX_OFFSET // 4] = d.max() * self.majorant_scale + 1e-6
            self.scene.set_volume_grid(self.volume, self.grid)


# backward() replays paths of its own: a loss that is not linear in the image
# (an L2 image loss) has an unbiased gradient only when the image factor and
# the derivative factor come from independent samples
DECORRELATE = 0x68E31DA4


class RenderFunction(torch.autograd.Function):
    """hdr (H, W, 3) = render(albedo (T, 3), density (n, n, n)); the gradient
    is estimated with paths independent of the image's (seed xor
    DECORRELATE)."""

    @staticmethod
    def forward(ctx, albedo, density, dscene, view, params):
        dscene.upload(albedo if albedo.numel() else None, density if density.numel() else None)
        p = abi.st_render_params.from_buffer_copy(params)
        p.flags |= abi.ST_FLAG_FLOAT_NEE
        out = TestRenderer(dscene.scene).render(view, p)
        ctx.dscene, ctx.view, ctx.params = dscene, view, p
        ctx.has_albedo, ctx.has_density = albedo.numel() > 0, density.numel() > 0
        return out["hdr"][..., :3].clone()

    @staticmethod
    def backward(ctx, grad_hdr):
        ds, p = ctx.dscene, ctx.params
        d_hdr = torch.zeros((p.height, p.width, 4), dtype=torch.float32, device=ds.dev)
        d_hdr[..., :3] = grad_hdr
        ga = torch.zeros((ds.triangles, 3), dtype=torch.float32, device=ds.dev) if ctx.has_albedo else None
        gg = torch.zeros(ds.words, dtype=torch.float32, device=ds.dev) if ctx.has_density else None
        q = abi.st_render_params.from_buffer_copy(p)
        q.seed = (p.seed ^ DECORRELATE) & 0xFFFFFFFF
        ds.scene.adjoint(ctx.view, q, d_hdr, ga, gg, ds.volume)
        gd = gg[ds.slots].reshape(ds.shape) if gg is not None else None
        return ga, gd, None, None, None


def render(dscene, view, params, albedo=None, density=None):
    """Differentiable render: hdr (H, W, 3) as a function of albedo and density
    (either may be None: an empty tensor stands in)."""
    empty = torch.empty(0, device=dscene.dev)
    return RenderFunction.apply(albedo if albedo is not None else empty, density if density is not None else empty, dscene,
                                view, params)
