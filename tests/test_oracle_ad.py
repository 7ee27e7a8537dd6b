"""Pins of the oracle's forward-mode (dual number) derivatives of the spherical
polygon area and of the clipped intersection area (oracle/cr_oracle_ad.cpp),
which the GPU vertex-gradient tests compare against.

Checks independent of the dual-number code:
  * central finite differences (h = 1e-6) of the oracle's own FP64 area and
    clip functions, which are themselves pinned to 50-digit references
    (test_oracle_golden.py);
  * invariants of the exact derivative: the gradient of a function of
    normalised vertices is tangential (v . g_v = 0), and areas are invariant
    under rotations, so sum_v v x g_v = 0;
  * the octant triangle: moving one vertex along a great circle away from the
    opposite edge changes the area at the closed-form rate.
"""
import numpy as np
import pytest

import __graft_entry__

orc = __graft_entry__.load_oracle()


def _unit(a):
    a = np.asarray(a, dtype=np.float64)
    return a / np.linalg.norm(a, axis=-1, keepdims=True)


def _cell(lon0, lat0, dlon, dlat, n_side=1, jitter=0.0, rng=None):
    """A lon-lat-like cell with n_side points per side, CCW seen from outside."""
    lons = np.concatenate([np.linspace(lon0, lon0 + dlon, n_side + 1)[:-1],
                           np.full(n_side, lon0 + dlon),
                           np.linspace(lon0 + dlon, lon0, n_side + 1)[:-1],
                           np.full(n_side, lon0)])
    lats = np.concatenate([np.full(n_side, lat0),
                           np.linspace(lat0, lat0 + dlat, n_side + 1)[:-1],
                           np.full(n_side, lat0 + dlat),
                           np.linspace(lat0 + dlat, lat0, n_side + 1)[:-1]])
    if jitter and rng is not None:
        lons = lons + jitter * dlon * (rng.random(lons.shape) - 0.5)
        lats = lats + jitter * dlat * (rng.random(lats.shape) - 0.5)
    lo, la = np.radians(lons), np.radians(lats)
    return np.stack([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)], -1)


def _fd(f, pts, h=1e-6):
    g = np.zeros_like(pts)
    for i in range(pts.shape[0]):
        for k in range(3):
            p, m = pts.copy(), pts.copy()
            p[i, k] += h
            m[i, k] -= h
            g[i, k] = (f(p) - f(m)) / (2 * h)
    return g


def _pairs():
    rng = np.random.default_rng(21)
    out = []
    for t in range(12):
        lat = rng.uniform(-70, 70)
        lon = rng.uniform(0, 360)
        d = rng.uniform(1.0, 4.0)
        subj = _cell(lon, lat, d, d * rng.uniform(0.6, 1.2), n_side=1 + (t % 4), jitter=0.05, rng=rng)
        clip = _cell(lon + rng.uniform(0.2, 0.7) * d, lat + rng.uniform(0.2, 0.7) * d, d * 0.9, d * 0.8)
        out.append((subj, clip))
    return out


@pytest.mark.parametrize("n", [3, 4, 7, 16])
def test_polygon_area_gradient_vs_finite_differences(n):
    rng = np.random.default_rng(n)
    ring = _cell(rng.uniform(0, 360), rng.uniform(-60, 60), 3.0, 2.0, n_side=max(1, n // 4), jitter=0.1,
                 rng=rng)[:max(n, 3)]
    a, g = orc.polygon_area_grad(ring)
    assert a == pytest.approx(orc.polygon_area(ring), rel=1e-13)
    fd = _fd(lambda p: orc.polygon_area(_unit(p)), ring)
    assert np.max(np.abs(g - fd)) <= 1e-6 * np.max(np.abs(fd))


@pytest.mark.parametrize("i", range(12))
def test_clip_area_gradient_vs_finite_differences(i):
    subj, clip = _pairs()[i]
    a, gs, gc = orc.clip_area_grad(subj, clip)
    assert a > 0
    assert a == pytest.approx(orc.clip_area(subj, clip), rel=1e-12)
    fs = _fd(lambda p: orc.clip_area(_unit(p), clip), subj)
    fc = _fd(lambda p: orc.clip_area(subj, _unit(p)), clip)
    scale = max(np.max(np.abs(fs)), np.max(np.abs(fc)))
    assert np.max(np.abs(gs - fs)) <= 1e-6 * scale
    assert np.max(np.abs(gc - fc)) <= 1e-6 * scale


@pytest.mark.parametrize("i", range(12))
def test_clip_area_gradient_invariants(i):
    subj, clip = _pairs()[i]
    _, gs, gc = orc.clip_area_grad(subj, clip)
    scale = np.abs(gs).sum() + np.abs(gc).sum()
    # tangential
    assert np.max(np.abs(np.sum(gs * subj, -1))) <= 1e-12 * scale
    assert np.max(np.abs(np.sum(gc * clip, -1))) <= 1e-12 * scale
    # rotation invariance: sum of v x dA/dv over every vertex vanishes
    torque = np.cross(subj, gs).sum(0) + np.cross(clip, gc).sum(0)
    assert np.max(np.abs(torque)) <= 1e-12 * scale


def test_octant_vertex_rate():
    """Octant (x, y, z): moving z along the meridian toward x by angle t keeps
    the triangle's x-y edge and shrinks the area; the excess of (x, y, p(t))
    with p = (sin t, 0, cos t) is pi/2 - t, so dA/dt = -1."""
    tri = np.eye(3)
    _, g = orc.polygon_area_grad(tri)
    dp_dt = np.array([1.0, 0.0, 0.0])  # d/dt (sin t, 0, cos t) at t = 0
    assert g[2] @ dp_dt == pytest.approx(-1.0, rel=1e-12)
