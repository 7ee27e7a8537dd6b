"""Generic grid adapter (host code; no GPU needed): ragged polygon lists, the
Trees-interface flattener, RingGrids full grids and lon/lat conversion.

The adapters replace Trees.treeify for grids given by Trees.ncells /
Trees.getcell (docs/src/interfaces/trees.md:34-44) and the RingGrids
extension (ext/ConservativeRegriddingRingGridsExt.jl:22-50).  RingGrids itself
is not in the image, so its latitude sets are pinned by their definitions
(Legendre roots, equidistant colatitudes) and the grids by the invariants the
reference tests use for them (test/usecases/fullclenshaw.jl: constant fields
preserved to 1e-10, area-weighted means to rtol 1e-10), checked here with the
oracle build.
"""
import math

import numpy as np
import pytest

from polygrids import geodesic_dual, rotation


def test_polygons_pad_and_close(cr):
    v = cr.lonlat_to_unit([0, 90, 0, 90, 45], [0, 0, 45, 45, 60])
    off = np.array([0, 4, 8])
    idx = np.array([0, 1, 4, 0, 0, 1, 3, 2])  # a closed triangle, then a quad
    assert cr.polygons_width(v, off, idx) == 4
    cells, vidx = cr.polygons_to_cells(v, off, idx)
    assert cells.shape == (2, 4, 3)
    np.testing.assert_array_equal(vidx, [[0, 1, 4, 4], [0, 1, 3, 2]])
    np.testing.assert_array_equal(cells[0], v[[0, 1, 4, 4]])
    np.testing.assert_array_equal(cells[1], v[[0, 1, 3, 2]])
    wide, wv = cr.polygons_to_cells(v, off, idx, nv=6)
    np.testing.assert_array_equal(wv[1], [0, 1, 3, 2, 2, 2])


@pytest.mark.parametrize("off,idx,nv", [
    ([0, 3], [0, 1, 7], None),        # vertex out of range
    ([0, 3], [0, 1, -1], None),       # negative vertex
    ([0, 2], [0, 1], None),           # two vertices
    ([0, 3], [0, 1, 0], None),        # closed two-gon
    ([1, 3], [0, 1, 2], None),        # offsets must start at 0
    ([0, 3], [0, 1, 2], 2),           # table narrower than the widest cell
])
def test_polygons_reject(cr, off, idx, nv):
    v = cr.lonlat_to_unit([0, 90, 45], [0, 0, 60])
    off = np.array(off)
    if off[-1] != len(idx):
        with pytest.raises(cr.ArgumentError):
            cr.polygons_to_cells(v, off, idx, nv=nv)
        return
    with pytest.raises(cr.ArgumentError):
        cr.polygons_to_cells(v, off, idx, nv=nv)


def test_polygons_reject_decreasing_offsets(cr):
    v = cr.lonlat_to_unit([0, 90, 45, 10], [0, 0, 60, 10])
    with pytest.raises(cr.ArgumentError):
        cr.polygons_width(v, np.array([0, 4, 3, 7]), np.array([0, 1, 2, 3, 0, 1, 2]))


def test_lonlat_to_unit_exact(cr):
    p = cr.lonlat_to_unit([0.0, 123.0, 90.0, 180.0, -90.0], [90.0, -90.0, 0.0, 0.0, 0.0])
    np.testing.assert_array_equal(p[0], [0, 0, 1])
    np.testing.assert_array_equal(p[1], [0, 0, -1])
    np.testing.assert_array_equal(p[2], [0, 1, 0])
    np.testing.assert_array_equal(p[3], [-1, 0, 0])
    np.testing.assert_array_equal(p[4], [0, -1, 0])
    # the lon-lat generator uses the same conversion
    P = cr.lonlat_points(8, 4)
    lon, lat = np.meshgrid(np.arange(9) * 45.0, -90.0 + np.arange(5) * 45.0, indexing="ij")
    np.testing.assert_array_equal(cr.lonlat_to_unit(lon, lat), P)


def test_flatten_tree_shares_and_closes(cr):
    P = cr.lonlat_points(6, 3)
    cells = cr.points_to_cells(P)

    class Tree:  # Trees.ncells / Trees.getcell, closed rings like GI.LinearRing
        def ncells(self):
            return cells.shape[0]

        def getcell(self, i):
            return np.vstack([cells[i], cells[i][:1]])

    v, off, idx = cr.flatten_tree(Tree())
    assert v.shape[0] == 6 * 2 + 2  # lon 360 is lon 0 bitwise; each pole is one point
    np.testing.assert_array_equal(np.diff(off), 4)
    np.testing.assert_array_equal(cr.polygons_to_cells(v, off, idx)[0], cells)
    v2, off2, idx2 = cr.flatten_tree(Tree(), share_vertices=False)
    assert v2.shape[0] == 4 * cells.shape[0]

    class LonLatTree:  # rings of lon/lat degrees, clockwise
        ncells = 2

        def getcell(self, i):
            return [(0 + 10 * i, 0), (0 + 10 * i, 10), (10 + 10 * i, 10), (10 + 10 * i, 0), (0 + 10 * i, 0)]

    v3, off3, idx3 = cr.flatten_tree(LonLatTree())
    assert v3.shape[0] == 6 and list(off3) == [0, 4, 8]


def test_ring_latitudes(cr):
    g = cr.ring_latitudes("gaussian", 12)
    assert g.size == 24 and np.all(np.diff(g) < 0)
    np.testing.assert_allclose(g, -g[::-1], atol=1e-13)
    leg = np.polynomial.legendre.Legendre.basis(24)
    assert np.max(np.abs(leg(np.sin(np.radians(g))))) < 1e-12
    c = cr.ring_latitudes("clenshaw", 12)
    assert c.size == 23 and c[11] == 0.0
    np.testing.assert_allclose(np.diff(c), -180.0 / 24, rtol=1e-13)
    assert cr.ring_longitudes(12).size == 48
    with pytest.raises(cr.ArgumentError):
        cr.ring_latitudes("octahedral", 4)


def test_ring_full_polygons_layout(cr, orc):
    latd, lond = cr.ring_latitudes("clenshaw", 6), cr.ring_longitudes(6)
    v, off, idx = cr.ring_full_polygons(latd, lond)
    nlat, nlon = latd.size, lond.size
    cells, _ = cr.polygons_to_cells(v, off, idx)
    assert cells.shape == (nlat * nlon, 4, 3)
    # ring 0 touches the north pole, the last ring the south pole
    assert np.all(cells[:nlon, 2:, 2] == 1.0) and np.all(cells[-nlon:, :2, 2] == -1.0)
    # cell k = ring*nlon + i sits between the ring's latitude edges, east of lon edge i
    lat = np.degrees(np.arcsin(np.clip(cells[..., 2], -1, 1)))
    ring = np.arange(nlat * nlon) // nlon
    edges = np.concatenate([[90.0], 0.5 * (latd[:-1] + latd[1:]), [-90.0]])
    np.testing.assert_allclose(lat.max(axis=1), edges[ring], atol=1e-12)
    np.testing.assert_allclose(lat.min(axis=1), edges[ring + 1], atol=1e-12)
    a = orc.cell_areas(cells)
    assert np.all(a > 0)
    assert abs(a.sum() - 4 * math.pi) < 1e-12 * 4 * math.pi


def test_geodesic_dual_grid(cr, orc):
    v, off, idx = geodesic_dual(3)
    lens = np.diff(off)
    assert lens.size == 642 and np.sum(lens == 5) == 12 and np.all((lens == 5) | (lens == 6))
    cells, _ = cr.polygons_to_cells(v, off, idx)
    a = orc.cell_areas(cells)
    assert np.all(a > 0)
    assert abs(a.sum() - 4 * math.pi) < 1e-12 * 4 * math.pi


@pytest.mark.parametrize("clip_is_dst", [True, False])
def test_oracle_regrid_ragged_invariants(cr, orc, clip_is_dst):
    """Full sphere to full sphere with ragged cells: row sums = destination
    areas, column sums = source areas (sweat.jl:113-131, rtol sqrt(eps)) and
    a constant field is preserved."""
    v, off, idx = geodesic_dual(3, rotate=rotation(4))
    hexes, _ = cr.polygons_to_cells(v, off, idx)
    ll = cr.points_to_cells(cr.lonlat_points(36, 18))
    dst, src = (hexes, ll) if clip_is_dst else (ll, hexes)
    ref = orc.build(dst, src, clip_is_dst=True)
    rows = np.add.reduceat(ref.val, ref.rowptr[:-1]) if ref.nnz else np.zeros(0)
    np.testing.assert_allclose(rows, ref.a_dst, rtol=1.5e-8)
    cols = np.bincount(ref.col, weights=ref.val, minlength=ref.n_src)
    np.testing.assert_allclose(cols, ref.a_src, rtol=1.5e-8)
    np.testing.assert_allclose(ref.apply(np.ones(ref.n_src)), 1.0, atol=1e-10)


def test_oracle_zero_area_cells_in_no_pair(cr, orc):
    """A zero-area cell (a padding ghost) intersects nothing, also as the clip
    polygon, where its zero-length edges would pass the whole subject."""
    ll = cr.points_to_cells(cr.lonlat_points(12, 6))
    dst = cr.points_to_cells(cr.lonlat_points(8, 4))
    dst[5] = dst[5][0]  # all four vertices at one point
    for clip_is_dst in (True, False):
        ref = orc.build(dst, ll, clip_is_dst=clip_is_dst)
        assert ref.a_dst[5] == 0.0
        assert ref.rowptr[6] == ref.rowptr[5]
        src = ll.copy()
        src[7] = src[7][2]
        ref = orc.build(dst, src, clip_is_dst=clip_is_dst)
        assert ref.a_src[7] == 0.0 and not np.any(ref.col == 7)


def test_folded_grid_partition(cr, orc):
    """The synthetic RightCenterFolded grid tiles the sphere once: primaries +
    rest sum to 4 pi, ghosts have zero area, and regridding it to a lon-lat
    grid reproduces the cell areas as row and column sums."""
    from polygrids import folded_cells, folded_points
    pts = folded_points(cr, 24, 10, 60.0)
    cells, ghost = folded_cells(cr, pts)
    a = orc.cell_areas(cells)
    assert np.all(a[ghost] == 0.0) and np.all(np.delete(a, ghost) > 0)
    assert abs(a.sum() - 4 * math.pi) < 1e-12 * 4 * math.pi
    ll = cr.points_to_cells(cr.lonlat_points(30, 15))
    ref = orc.build(ll, cells)
    cols = np.bincount(ref.col, weights=ref.val, minlength=ref.n_src)
    np.testing.assert_allclose(cols, ref.a_src, rtol=1.5e-8, atol=0)
    rows = np.add.reduceat(ref.val, ref.rowptr[:-1])
    np.testing.assert_allclose(rows, ref.a_dst, rtol=1.5e-8)
