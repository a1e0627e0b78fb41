"""Pins of the oracle's surface-tension curvature (surf_smooth.comp + surf_curv.comp, the
curvature of the smoothed liquid fraction from its node normals), on the CPU, against
geometry:

  * a disc of liquid of radius R: over the interface band (cells within 1.5 of the rim) the
    mean of curvature x distance to the centre is 1 (the curvature of a circle, 1 / r) within
    5 %, for two radii, positive for a convex blob;
  * a straight surface (a half-plane of liquid, at 0 and 30 degrees): no curvature in the band.
A dropped half, a sign error or a stencil slip fails one of them.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim


def curvature_of(phi):
    H, W = phi.shape
    grid = (W, H)
    params, field, parts, rows = cases.dam_break(grid, frac=(0.1, 0.1))
    g, mats = N.derive_grid_params(params, W, H, rows)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    cell = np.zeros((H, W, 4), np.float32)
    cell[..., 0] = phi
    cell[..., 1] = 1.0
    cell[..., 3] = phi - 0.5
    o.put_field("CELL", cell.reshape(-1, 4))
    comp = np.zeros((H, W, 4), np.float32)
    comp[..., 0] = np.where(phi > 0, 1.0, 0.0)
    comp[..., 2] = -1.0
    comp[..., 3] = phi
    o.put_field("COMP", comp.reshape(-1, 4))
    o.stage("SURF")
    return o.field("SURF", np.float32, (H, W, 4))[..., 2].astype(np.float64)


@pytest.mark.parametrize("R", [12.0, 20.0])
def test_disc_has_the_curvature_of_a_circle(R):
    W = H = 64
    y, x = np.mgrid[0:H, 0:W].astype(np.float64) + 0.5
    r = np.hypot(x - 32.0, y - 32.0)
    k = curvature_of(np.clip(R + 0.5 - r, 0.0, 1.0).astype(np.float32))
    band = (k != 0.0) & (np.abs(r - R) < 1.5)
    assert band.sum() > 100
    assert np.mean(k[band] * r[band]) == pytest.approx(1.0, abs=0.05)


@pytest.mark.parametrize("deg", [0.0, 30.0])
def test_straight_surface_has_none(deg):
    W = H = 64
    y, x = np.mgrid[0:H, 0:W].astype(np.float64) + 0.5
    a = np.radians(deg)
    d = (y - 30.0) * np.cos(a) - (x - 32.0) * np.sin(a)  # signed distance above the surface
    k = curvature_of(np.clip(0.5 - d, 0.0, 1.0).astype(np.float32))
    inner = (np.abs(d) < 1.5) & (x > 8) & (x < W - 8) & (y > 8) & (y < H - 8)
    assert inner.sum() > 40
    assert np.abs(k[inner]).max() < 0.02
