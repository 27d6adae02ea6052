"""Alpha-masked triangles in the oracle (the CANDIDATE_NON_OPAQUE_TRIANGLE
branch of intersection.hlsli:125-147, getMaterialAlphaMask
material.hlsli:72-76), pinned by what the mask geometry fixes:

- rays through the interior of a checker mask's cells hit the panel exactly
  where channel 0 is 1 and pass through to the wall behind where it is 0,
  with the cell found from the texture coordinate (u along x, v along y);
- a linear ramp sampled bilinearly at level 0 reproduces itself inside a
  cell, so a ray hits the grille exactly where u reaches the cutoff (a
  transposed coordinate, a mip level other than 0 or a flipped comparison
  moves that boundary);
- the any-hit query applies the same test;
- masks on sphere materials are refused.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from stratum2_amd import scenes


def _rays(o, d, tmax=np.inf):
    r = np.zeros((len(o), 8), np.float32)
    r[:, :3], r[:, 3], r[:, 4:7] = o, tmax, d
    return r


def _interior_points(rng, n, cells=4, margin=0.08):
    """(u, v) points inside checker cells, at least margin (in cell units)
    from a cell edge, with their cell's checker parity."""
    c = rng.integers(0, cells, (n, 2))
    f = rng.uniform(margin, 1 - margin, (n, 2))
    uv = (c + f) / cells
    return uv, (c.sum(1) % 2 == 0)


def _masked_vs_bounds(cutoff, o, d, tmax=np.inf, closest=True):
    """The hits with the mask at cutoff, and the two bounds a mask moves
    between: every masked triangle committed (cutoff -1) and none (cutoff 2,
    above any mask value)."""
    r = _rays(o, d, tmax)
    return [OracleScene(scenes.cornell_alpha(0, cutoff=c)).trace_rays(r, closest=closest) for c in (cutoff, -1.0, 2.0)]


def test_panel_hits_follow_the_checker_mask():
    """Where channel 0 is 1 the hit is the one with the panel present, where
    it is 0 the one with the panel absent (the wall or a block behind), to
    the bit, for both the closest and the any-hit query."""
    panel = len(scenes.cornell_alpha(0).instances) - 2
    uv, solid = _interior_points(np.random.default_rng(3), 4000)
    o = np.stack([uv[:, 0] - 0.5, uv[:, 1] + 0.4, np.full(len(uv), 0.6)], 1)
    d = np.tile([0.0, 0.0, -1.0], (len(uv), 1))
    for cutoff in (0.5, 0.999):
        h, full, none = _masked_vs_bounds(cutoff, o, d)
        assert ((full[:, 1] & 0xFFFF) == panel).mean() > 0.7  # the panel is in front nearly everywhere
        assert not np.array_equal(full, none)
        assert np.array_equal(h, np.where(solid[:, None], full, none))
        a, afull, anone = _masked_vs_bounds(cutoff, o, d, tmax=0.6 + 1e-3, closest=False)
        assert afull.all() or afull.mean() > 0.7
        assert np.array_equal(a, np.where(solid, afull, anone))


@pytest.mark.parametrize("cutoff", [0.3, 0.5, 0.8])
def test_ramp_mask_cuts_at_the_cutoff(cutoff):
    """Channel 0 of the grille's mask is u inside its solid cells (a linear
    ramp reproduced by the bilinear filter one texel clear of the cell
    edges): the grille is hit exactly where u reaches the cutoff."""
    grille = len(scenes.cornell_alpha(0).instances) - 1
    rng = np.random.default_rng(int(cutoff * 10))
    uv, solid = _interior_points(rng, 6000, margin=0.13)
    uv = uv[solid]
    uv = uv[np.abs(uv[:, 0] - cutoff) > 1e-4]
    o = np.stack([uv[:, 0] * 0.8 - 0.4, np.full(len(uv), 1.0), uv[:, 1] * 0.8 - 0.4], 1)
    d = np.tile([0.0, 1.0, 0.0], (len(uv), 1))
    h, full, none = _masked_vs_bounds(cutoff, o, d)
    assert np.all((full[:, 1] & 0xFFFF) == grille)
    want = uv[:, 0] >= cutoff
    assert np.array_equal(h, np.where(want[:, None], full, none))
    assert 0.1 < want.mean() < 0.9


def test_mask_on_a_sphere_is_refused():
    sd = scenes.cornell_spheres(0)
    sd.images.append(np.ones((4, 4, 4), np.float32))
    w = scenes.material_record()
    w[12] = len(sd.images) - 1
    sd.add_sphere_instance(sd.add_material(w), 0.1, scenes.translation(0, 1, 0))
    with pytest.raises(RuntimeError, match="alpha"):
        OracleScene(sd)
