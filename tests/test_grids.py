"""Grid generators of libcrhip.so (host code; no GPU needed).

They replace Healpix.jl (boundariesRing, nest2ring, ring2nest), ClimaCore's
cubed sphere and UnitSphereFromGeographic, none of which ship with the
reference; parity to those packages is unpinned, so these tests pin the
generators with closed forms and topology:

* HEALPix pixel centres vs the ring-scheme closed forms of Gorski et al. 2005
  (restated below in numpy), nest<->ring bijection, Euler's formula for the
  boundary polygons, total area 4 pi, pixel area -> pi / (3 nside^2).
* Cubed sphere: Euler's formula with bitwise-shared face edges, area 4 pi.
* Lon-lat: exact poles and exact 0/360 longitude wrap.
"""
import math

import numpy as np
import pytest


def pix2ang_ring(nside, p):
    """Pixel-centre (z, phi) of ring-scheme pixel p (Gorski et al. 2005, sec. 4)."""
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    p = np.asarray(p, dtype=np.int64)
    z = np.empty(p.shape)
    phi = np.empty(p.shape)
    north = p < ncap
    south = p >= npix - ncap
    eq = ~north & ~south
    ph = (p[north] + 1) / 2.0
    i = np.floor(np.sqrt(ph - np.sqrt(np.floor(ph)))).astype(np.int64) + 1
    j = p[north] + 1 - 2 * i * (i - 1)
    z[north] = 1 - i * i / (3.0 * nside * nside)
    phi[north] = math.pi / (2 * i) * (j - 0.5)
    pp = p[eq] - ncap
    i = pp // (4 * nside) + nside
    j = pp % (4 * nside) + 1
    s = (i - nside + 1) % 2
    z[eq] = 4.0 / 3.0 - 2.0 * i / (3.0 * nside)
    phi[eq] = math.pi / (2 * nside) * (j - s / 2.0)
    ps = npix - p[south]
    ph = ps / 2.0
    i = np.floor(np.sqrt(ph - np.sqrt(np.floor(ph)))).astype(np.int64) + 1
    j = 4 * i + 1 - (ps - 2 * i * (i - 1))
    z[south] = -1 + i * i / (3.0 * nside * nside)
    phi[south] = math.pi / (2 * i) * (j - 0.5)
    return z, phi


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16])
def test_healpix_nest_ring_bijection(cr, nside):
    n = 12 * nside * nside
    ring = np.array([cr.healpix_nest2ring(nside, p) for p in range(n)])
    assert sorted(ring.tolist()) == list(range(n))
    back = np.array([cr.healpix_ring2nest(nside, r) for r in ring])
    np.testing.assert_array_equal(back, np.arange(n))


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64])
def test_healpix_pixel_centres_match_ring_closed_form(cr, nside):
    # the centre of nested pixel p is the north corner of its child 4p at 2*nside
    fine = cr.healpix_cells(2 * nside, "nested", 1)
    n = 12 * nside * nside
    ring = np.array([cr.healpix_nest2ring(nside, p) for p in range(n)])
    centres = fine[4 * np.arange(n), 0, :]
    z, phi = pix2ang_ring(nside, ring)
    sth = np.sqrt((1 - z) * (1 + z))
    ref = np.stack([sth * np.cos(phi), sth * np.sin(phi), z], axis=1)
    np.testing.assert_allclose(centres, ref, atol=2e-15)


def test_healpix_base_pixel_centres(cr):
    fine = cr.healpix_cells(2, "nested", 1)
    z = fine[4 * np.arange(12), 0, 2]
    np.testing.assert_allclose(z, [2 / 3] * 4 + [0] * 4 + [-2 / 3] * 4, atol=1e-15)


def test_healpix_ring_order_is_permuted_nested(cr):
    nside = 8
    nested = cr.healpix_cells(nside, "nested", 2)
    ring = cr.healpix_cells(nside, "ring", 2)
    perm = np.array([cr.healpix_ring2nest(nside, r) for r in range(12 * nside * nside)])
    np.testing.assert_array_equal(ring, nested[perm])


@pytest.mark.parametrize("step", [1, 2, 4])
def test_healpix_boundaries_tile_the_sphere(cr, orc, step):
    nside = 8
    cells = cr.healpix_cells(nside, "nested", step)
    F = cells.shape[0]
    a = orc.cell_areas(cells)
    assert np.all(a > 0)  # counter-clockwise
    assert a.sum() == pytest.approx(4 * math.pi, rel=1e-14)
    # Euler: V - E + F = 2 with E = 2*step*F boundary segments
    keys = np.round(cells.reshape(-1, 3) * 1e12).astype(np.int64)
    V = np.unique(keys, axis=0).shape[0]
    assert V == (2 * step - 1) * F + 2


def test_healpix_pixel_area_converges_to_equal_area(cr, orc):
    nside = 16
    exact = math.pi / (3 * nside * nside)
    errs = []
    for step in (1, 4, 16):
        a = orc.cell_areas(cr.healpix_cells(nside, "nested", step))
        errs.append(np.max(np.abs(a - exact)) / exact)
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 2e-3


def test_healpix_poles_are_exact(cr):
    cells = cr.healpix_cells(4, "nested", 4)
    north = cells[:, 0, :]
    for f in range(4):
        p = f * 16 + 15  # ix = iy = nside-1 on the north faces
        np.testing.assert_array_equal(north[p], [0.0, 0.0, 1.0])


@pytest.mark.parametrize("ne", [1, 3, 8])
def test_cubed_sphere_edges_are_shared_bitwise(cr, orc, ne):
    P = cr.cubed_sphere_points(ne)
    pts = P.reshape(-1, 3)
    V = np.unique(pts, axis=0).shape[0]  # exact (bitwise) uniqueness
    assert V == 6 * ne * ne + 2  # Euler for a closed quad mesh
    np.testing.assert_allclose(np.linalg.norm(pts, axis=1), 1.0, atol=2e-16)
    cells = cr.points_to_cells(P)
    a = orc.cell_areas(cells)
    assert np.all(a > 0)
    assert a.sum() == pytest.approx(4 * math.pi, rel=1e-14)


def test_cubed_sphere_equiangular_spacing(cr):
    ne = 12
    P = cr.cubed_sphere_points(ne)
    # along a face's centre line, consecutive vertices are equally spaced in angle
    line = P[0, :, ne // 2, :]
    ang = np.arccos(np.clip(np.sum(line[1:] * line[:-1], axis=1), -1, 1))
    np.testing.assert_allclose(ang, math.pi / 2 / ne, rtol=1e-12)


def test_lonlat_poles_and_wrap(cr):
    P = cr.lonlat_points(90, 45)
    assert P.shape == (91, 46, 3)
    np.testing.assert_array_equal(np.abs(P[:, 0, :]), np.tile([0.0, 0.0, 1.0], (91, 1)))
    assert np.all(P[:, 0, 2] == -1.0) and np.all(P[:, -1, 2] == 1.0)
    np.testing.assert_array_equal(P[90], P[0])
    # degree-exact points
    Q = cr.lonlat_points(4, 2)  # 90-degree spacing
    np.testing.assert_array_equal(Q[1, 1], [0.0, 1.0, 0.0])
    np.testing.assert_array_equal(Q[2, 1], [-1.0, 0.0, 0.0])


def test_lonlat_cells_follow_cellbasedgrid_order(cr):
    """getcell(grid, i, j) = P[i,j], P[i+1,j], P[i+1,j+1], P[i,j+1]; linear index
    i + j*ni (src/trees/grids.jl:71-84, src/trees/interfaces.jl:236-243)."""
    P = cr.lonlat_points(6, 3)
    C = cr.points_to_cells(P)
    i, j = 4, 2
    np.testing.assert_array_equal(C[i + j * 6], [P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]])


def test_generator_argument_errors(cr):
    with pytest.raises(cr.ArgumentError):
        cr.lonlat_points(10, 10, lat0=-95.0)
    with pytest.raises(cr.ArgumentError):
        cr.healpix_cells(3, "nested", 1)  # nested needs a power of two
    with pytest.raises(cr.ArgumentError):
        cr.healpix_cells(4, "bogus", 1)
