"""The oracle's double-precision build (oracle/_build/libpt_oracle64.so: the
same tracer with its arithmetic type set to double and libm's double
transcendentals, stored formats unchanged).  It is the reference the GPU's
converged images are measured against (tests/converged_rmse.py), so it is
held to the same closed forms as the float build, and it must compute in
double."""
import functools

import numpy as np
import pytest

import test_oracle_media as M
import test_oracle_units as U
from oracle.oracle import OracleScene, lib
from stratum2_amd import abi, scenes


@pytest.fixture
def fp64(monkeypatch):
    for mod in (U, M):
        monkeypatch.setattr(mod, "OracleScene", functools.partial(OracleScene, fp64=True))


def test_transcendentals_are_libm_double():
    """or_math of the double build returns libm's double result rounded to
    float (the float build's Cody-Waite polynomials are off by an ulp or two)."""
    L = lib(fp64=True)
    x = np.linspace(-20, 20, 4001).astype(np.float32)
    y = np.zeros_like(x)
    out = np.zeros_like(x)
    L.or_math(0, x.size, x.ctypes.data, y.ctypes.data, out.ctypes.data)
    assert np.array_equal(out, np.sin(x.astype(np.float64)).astype(np.float32))
    xp = np.linspace(0.01, 50, 4001).astype(np.float32)
    L.or_math(4, xp.size, xp.ctypes.data, y.ctypes.data, out.ctypes.data)
    assert np.array_equal(out, np.log(xp.astype(np.float64)).astype(np.float32))


def test_emission_only_image(fp64):
    U.test_emission_only_image_is_analytic()


def test_furnace_closed_box(fp64):
    U.test_furnace_closed_box()


@pytest.mark.parametrize("use_grid", [False, True])
def test_absorbing_medium_chord_transmittance(fp64, use_grid):
    M.test_absorbing_medium_chord_transmittance(use_grid)


@pytest.mark.parametrize("use_grid", [False, True])
def test_lit_floor_under_absorbing_slab(fp64, use_grid):
    M.test_lit_floor_under_absorbing_slab(use_grid)


@pytest.mark.parametrize("g", [0.0, 0.3])
def test_scattering_medium_furnace(fp64, g):
    M.test_scattering_medium_furnace(g)


def test_same_paths_as_the_float_build_nearly_everywhere():
    """On the same PCG streams the two builds take the same paths except
    where a rounding difference moves a sampled direction across a packing or
    comparison boundary: almost every path has the same length, and the
    images agree to a few float ulps where they do."""
    sd = scenes.cornell_box(0)
    W = H = 32
    view = scenes.cornell_view(W, H)
    p = abi.default_params(W, H, max_depth=4, spp=8)
    a = OracleScene(sd).render(view, p)
    b = OracleScene(sd, fp64=True).render(view, p)
    same = (a["path_length"] == b["path_length"]).mean()
    assert same > 0.97, same
    assert not np.array_equal(a["hdr"], b["hdr"])
    assert abs(a["hdr"][..., :3].mean() / b["hdr"][..., :3].mean() - 1) < 0.01
