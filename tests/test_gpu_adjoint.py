"""st_adjoint (path-replay adjoint) on the GPU against the oracle's or_adjoint
on the same seeded inputs, the per-triangle albedo's forward parity, a
size-independent check at 512^2 x 8 spp (the gradient's directional
derivative along the albedos equals the fixed-seed central difference of
the GPU's own render when roulette is off), and 100 Adam steps of
torch.autograd over the C ABI.

Per path the GPU and the oracle compute the same fp32 values; only the
order of the atomic sums differs, hence the tolerances: 1e-4 of each
gradient plus 1e-5 of the largest."""
import numpy as np
import pytest

from stratum2_amd import abi, scenes

pytestmark = pytest.mark.gpu

FLAGS = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_LAMBERTIAN | abi.ST_FLAG_FLOAT_NEE
NO_RR = 255


def _params(W, H, spp, depth, seed=3, rr=NO_RR):
    return abi.default_params(W, H, max_depth=depth, spp=spp, flags=FLAGS, rr_start_depth=rr, seed=seed)


def _close(got, want):
    tol = 1e-4 * np.abs(want) + 1e-5 * np.abs(want).max()
    bad = np.abs(got - want) > tol
    assert not bad.any(), (int(bad.sum()), got[bad][:5], want[bad][:5])


@pytest.mark.parametrize("rr", [NO_RR, 2])
def test_albedo_gradient_matches_oracle(gpu_ctx, rr):
    import torch
    from oracle.oracle import OracleScene
    from stratum2_amd.renderer import Scene, TestRenderer
    rng = np.random.default_rng(5)
    sd = scenes.cornell_box(0)
    sc, o = Scene(gpu_ctx, sd), OracleScene(sd)
    T = o.triangle_count()
    assert sc.stats()["triangles"] == T
    alb = rng.uniform(0.2, 0.9, (T, 3)).astype(np.float32)
    o.set_triangle_albedo(alb)
    sc.set_triangle_albedo(torch.from_numpy(alb).cuda())
    W = H = 32
    view, p = scenes.cornell_view(W, H), _params(W, H, spp=4, depth=5, rr=rr)
    # the albedo table's forward parity
    got = TestRenderer(sc).render(view, p, want_path_length=True)
    torch.cuda.synchronize()
    ref = o.render(view, p)
    assert np.array_equal(got["path_length"].cpu().numpy(), ref["path_length"])
    h, r = got["hdr"].cpu().numpy(), ref["hdr"]
    assert (np.abs(h - r) <= 1e-4 * np.maximum(np.abs(r), 1e-3)).mean() == 1.0
    dh = rng.uniform(0.5, 1.5, (H, W, 4)).astype(np.float32)
    ga = torch.zeros((T, 3), dtype=torch.float32, device="cuda")
    sc.adjoint(view, p, torch.from_numpy(dh).cuda(), grad_albedo=ga)
    torch.cuda.synchronize()
    want, _ = o.adjoint(view, p, dh)
    _close(ga.cpu().numpy().astype(np.float64), want)
    sc.set_triangle_albedo(None)


def test_density_gradient_matches_oracle(gpu_ctx):
    import torch
    from oracle.oracle import OracleScene
    from stratum2_amd.renderer import Scene
    rng = np.random.default_rng(6)
    sd, offs = scenes.cornell_cloud(seed=1, n=16)
    sc, o = Scene(gpu_ctx, sd), OracleScene(sd)
    W = H = 32
    view, p = scenes.cornell_view(W, H), _params(W, H, spp=4, depth=5, rr=2)
    words = sd.volumes[0].nbytes // 4
    dh = rng.uniform(0.5, 1.5, (H, W, 4)).astype(np.float32)
    gg = torch.zeros(words, dtype=torch.float32, device="cuda")
    sc.adjoint(view, p, torch.from_numpy(dh).cuda(), grad_grid=gg, grad_volume=0)
    torch.cuda.synchronize()
    _, want = o.adjoint(view, p, dh, albedo=False, grid_vol=0, grid_words=words)
    got = gg.cpu().numpy().astype(np.float64)
    assert np.abs(want).max() > 0
    _close(got, want)


def test_directional_derivative_at_full_size(gpu_ctx):
    """512^2 x 8 spp, 64^3 grid: sum_t albedo_t . grad_t = d/ds L(s albedo) at
    s = 1, which with roulette off is the fixed-seed central difference of the
    GPU render (an exact polynomial in s up to fp32 rounding)."""
    import torch
    from stratum2_amd.renderer import Scene, TestRenderer
    rng = np.random.default_rng(7)
    sd, _ = scenes.cornell_cloud(seed=2, n=64)
    sc = Scene(gpu_ctx, sd)
    T = int(sc.stats()["triangles"])
    alb = torch.from_numpy(rng.uniform(0.2, 0.8, (T, 3)).astype(np.float32)).cuda()
    W = H = 512
    view, p = scenes.cornell_view(W, H), _params(W, H, spp=8, depth=6)
    dh = torch.from_numpy(rng.uniform(0.5, 1.5, (H, W, 4)).astype(np.float32)).cuda()

    def loss(s):
        sc.set_triangle_albedo((alb * s).contiguous())
        out = TestRenderer(sc).render(view, p)
        torch.cuda.synchronize()
        return float((out["hdr"].double() * dh.double()).sum())

    eps = 1e-3
    fd = (loss(1 + eps) - loss(1 - eps)) / (2 * eps)
    sc.set_triangle_albedo(alb)
    ga = torch.zeros((T, 3), dtype=torch.float32, device="cuda")
    sc.adjoint(view, p, dh, grad_albedo=ga)
    torch.cuda.synchronize()
    dd = float((ga.double() * alb.double()).sum())
    assert np.isfinite(dd) and abs(dd - fd) <= 2e-3 * abs(fd), (dd, fd)


def test_adam_recovers_albedo_and_density(gpu_ctx):
    """100 Adam steps (lr 1e-2, 8 spp per step, the gradient from paths
    independent of the image's) from albedo 0.5 and density 0.3 towards a
    target rendered at 256 spp with other values: the image error, measured
    at 128 spp with a fixed seed before and after, falls."""
    import torch
    from stratum2_amd import diff
    from stratum2_amd.renderer import Scene
    rng = np.random.default_rng(8)
    n = 16
    target_density = np.clip(rng.uniform(-0.2, 1.0, (n, n, n)), 0, None).astype(np.float32)
    sd, offs = scenes.cornell_cloud(seed=3, n=n, density=np.full((n, n, n), 0.3, np.float32))
    sc = Scene(gpu_ctx, sd)
    ds = diff.DiffScene(sc, offs)
    T = ds.triangles
    W = H = 48
    view = scenes.cornell_view(W, H)
    true_alb = torch.from_numpy(rng.uniform(0.1, 0.9, (T, 3)).astype(np.float32)).cuda()
    with torch.no_grad():
        target = diff.render(ds, view, abi.default_params(W, H, max_depth=4, spp=256, flags=FLAGS, rr_start_depth=3, seed=99),
                             true_alb, torch.from_numpy(target_density).cuda())
    alb = torch.full((T, 3), 0.5, device="cuda", requires_grad=True)
    dens = torch.full((n, n, n), 0.3, device="cuda", requires_grad=True)

    def error():
        with torch.no_grad():
            img = diff.render(ds, view, abi.default_params(W, H, max_depth=4, spp=128, flags=FLAGS, rr_start_depth=3, seed=7),
                              alb, dens)
            return float(((img - target) ** 2).mean())

    before = error()
    opt = torch.optim.Adam([alb, dens], lr=1e-2)
    p = _params(W, H, spp=8, depth=4, rr=3)
    losses = []
    for it in range(100):
        p.seed = 1000 + it
        img = diff.render(ds, view, p, alb, dens)
        loss = ((img - target) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        assert torch.isfinite(alb.grad).all() and torch.isfinite(dens.grad).all()
        assert float(alb.grad.abs().sum()) > 0 and float(dens.grad.abs().sum()) > 0
        opt.step()
        with torch.no_grad():
            alb.clamp_(0, 1)
            dens.clamp_(min=0)
        losses.append(float(loss))
    after = error()
    print("step losses", np.round(losses[::10], 5), "image error at 128 spp:", before, "->", after)
    assert after < 0.7 * before, (before, after)


def test_adjoint_argument_errors(gpu_ctx):
    import torch
    from stratum2_amd.renderer import Scene, TestRenderer
    sc = Scene(gpu_ctx, scenes.cornell_box(0))
    view, p = scenes.cornell_view(8, 8), _params(8, 8, spp=1, depth=3)
    dh = torch.ones((8, 8, 4), dtype=torch.float32, device="cuda")
    T = int(sc.stats()["triangles"])
    ga = torch.zeros((T, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError, match="per-triangle albedo"):
        sc.adjoint(view, p, dh, grad_albedo=ga)
    with pytest.raises(RuntimeError):
        sc.set_triangle_albedo(torch.zeros((T - 1, 3), dtype=torch.float32, device="cuda"))
    sc.set_triangle_albedo(torch.full((T, 3), 0.5, dtype=torch.float32, device="cuda"))
    with pytest.raises(RuntimeError, match="LAMBERTIAN"):
        TestRenderer(sc).render(view, abi.default_params(8, 8, max_depth=3, spp=1))
    with pytest.raises(RuntimeError, match="plain NEE"):
        sc.adjoint(view, abi.default_params(8, 8, max_depth=3, spp=1, flags=FLAGS | abi.ST_FLAG_RESTIR_DI), dh, grad_albedo=ga)
    with pytest.raises(RuntimeError, match="no such volume"):
        sc.adjoint(view, p, dh, grad_grid=torch.zeros(16, dtype=torch.float32, device="cuda"))
