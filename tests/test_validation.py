"""Host-side input checks and manifold resolution (no GPU needed).

* cr_validate_cells / cr_validate_points: non-unit or non-finite vertices are
  rejected (CR_INVALID_ARGUMENT), antipodal edges and non-convex cells of a
  convex grid are degenerate (CR_DEGENERATE_GEOMETRY), cells collapsed to one
  point are legal and counted; every grid the configurations use passes.
* The grid constructors run the same checks before touching the device
  (point matrices: their points; their cells on the device after the upload).
* Regridder(dst, src) manifold resolution of src/regridder/regridder.jl:105-123.
"""
import ctypes

import numpy as np
import pytest

from polygrids import folded_points, geodesic_dual


def _validate(cr, cells, convex=True):
    L = cr._lib
    c = np.ascontiguousarray(cells, dtype=np.float64)
    n = ctypes.c_int64(-1)
    st = L.lib.cr_validate_cells(L.dptr(c), c.shape[0], c.shape[1], 1 if convex else 0, ctypes.byref(n))
    return st, n.value, L.last_error()


def _lonlat_cell(cr, lon0, lat0, dlon, dlat):
    lon = np.array([lon0, lon0 + dlon, lon0 + dlon, lon0])
    lat = np.array([lat0, lat0, lat0 + dlat, lat0 + dlat])
    return cr.lonlat_to_unit(lon, lat)


def test_good_cells_pass_and_collapsed_cells_are_counted(cr):
    good = np.stack([_lonlat_cell(cr, 10.0 * k, 5.0, 4.0, 4.0) for k in range(5)])
    st, n, _ = _validate(cr, good)
    assert st == 0 and n == 0
    # a cell with every vertex at one point (a zero-area ghost) is legal
    coincident = np.repeat(good[2:3, :1], 4, axis=1)
    st, n, _ = _validate(cr, np.concatenate([good, coincident]))
    assert st == 0 and n == 1


def test_antipodal_edge_is_degenerate(cr):
    L = cr._lib
    cells = np.stack([_lonlat_cell(cr, 0.0, 0.0, 4.0, 4.0)] * 3)
    a = cr.lonlat_to_unit(np.array([0.0, 180.0, 90.0]), np.array([0.0, 0.0, 45.0]))
    cells[1, :3] = a
    cells[1, 3] = a[2]
    st, _, msg = _validate(cr, cells, convex=False)
    assert st == L.CR_DEGENERATE_GEOMETRY and "cell 1" in msg and "antipodal" in msg


def test_nonconvex_cell_in_a_convex_grid_is_degenerate(cr):
    L = cr._lib
    # an arrowhead: the fourth vertex is pushed inside the triangle of the others
    lon = np.array([0.0, 10.0, 5.0, 0.0])
    lat = np.array([0.0, 0.0, 10.0, 0.0])
    tri = cr.lonlat_to_unit(lon, lat)
    dart = tri.copy()
    dart[3] = cr.lonlat_to_unit(np.array([5.0]), np.array([3.0]))[0]
    cells = np.stack([_lonlat_cell(cr, 20.0, 0.0, 4.0, 4.0), dart])
    st, _, msg = _validate(cr, cells, convex=True)
    assert st == L.CR_DEGENERATE_GEOMETRY and "cell 1" in msg and "convex" in msg
    # the same table is fine when it is not declared convex (it is then the subject of the clip)
    st, _, _ = _validate(cr, cells, convex=False)
    assert st == 0


@pytest.mark.parametrize("bad", ["scaled", "nan", "inf"])
def test_non_unit_or_nonfinite_vertex_is_rejected(cr, bad):
    L = cr._lib
    cells = np.stack([_lonlat_cell(cr, 0.0, 0.0, 4.0, 4.0)] * 4)
    if bad == "scaled":
        cells[2, 1] *= 1.001
    else:
        cells[2, 1, 0] = np.nan if bad == "nan" else np.inf
    st, _, msg = _validate(cr, cells)
    assert st == L.CR_INVALID_ARGUMENT and "cell 2" in msg and "unit vector" in msg


def test_grid_constructors_check_before_the_device(cr):
    """The checks run on the host first: the error is the geometry's, not
    CR_NO_DEVICE (point matrices: their points; their cells' geometry is
    checked on the device after the upload, tests/test_gpu_build.py)."""
    cells = np.stack([_lonlat_cell(cr, 0.0, 0.0, 4.0, 4.0)] * 2)
    cells[1, 2] = -cells[1, 1]
    with pytest.raises(cr.CRError) as e:
        cr.Grid.from_cells(cells, convex=True)
    assert e.value.status == cr._lib.CR_DEGENERATE_GEOMETRY
    pts = cr.lonlat_points(8, 4)
    pts[3, 2] *= 2.0
    with pytest.raises(cr.ArgumentError):
        cr.Grid.from_points(pts)


def test_configuration_grids_pass(cr):
    """No false alarms on the grids of the configurations and tests."""
    L = cr._lib

    def points_ok(P):
        c, nb, ni, nj = cr.grids._point_blocks(P)
        n = ctypes.c_int64(-1)
        st = L.lib.cr_validate_points(L.dptr(c), nb, ni, nj, ctypes.byref(n))
        assert st == 0, L.last_error()
        return n.value

    assert points_ok(cr.lonlat_points(90, 45)) == 0
    assert points_ok(cr.lonlat_points(72, 36)) == 0
    assert points_ok(cr.lonlat_points(1440, 720)) == 0
    assert points_ok(cr.cubed_sphere_points(48)) == 0
    from gpu_common import jittered_cubed_sphere
    assert points_ok(jittered_cubed_sphere(cr, 96, seed=4)) == 0
    assert points_ok(folded_points(cr, 40, 30, 62.0)) == 0
    verts, off, idx = geodesic_dual(3)
    cells, _ = cr.polygons_to_cells(verts, off, idx)
    st, n, msg = _validate(cr, cells, convex=True)
    assert st == 0 and n == 0, msg
    for step in (1, 4):
        st, n, msg = _validate(cr, cr.healpix_cells(16, "nested", step=step), convex=(step == 1))
        assert st == 0 and n == 0, msg


def test_manifold_resolution(cr):
    Planar, Spherical, Geodesic = cr.Planar, cr.Spherical, cr.Geodesic
    edges = (np.arange(0.0, 361.0, 45.0), np.arange(-90.0, 91.0, 45.0))
    pts = cr.lonlat_points(8, 4)
    assert cr.best_manifold(edges) == Planar()
    assert cr.best_manifold(pts) == Spherical()
    assert Spherical() != Spherical(radius=1.0)
    assert cr.resolve_manifold(pts, pts) == Spherical()
    # implicit promotion of the planar side (regridder.jl:110-115)
    assert cr.resolve_manifold(edges, pts) == Spherical()
    assert cr.resolve_manifold(pts, edges) == Spherical()
    assert cr.resolve_manifold(edges, edges) == Planar()

    class OnEllipsoid:
        manifold = Geodesic()
    with pytest.raises(cr.ErrorException, match="manifolds must be the same"):
        cr.resolve_manifold(OnEllipsoid(), pts)
    with pytest.raises(cr.ErrorException):
        cr.Regridder(pts, OnEllipsoid())
    # two planar grids build on the plane: the planar grid constructor is
    # reached (without a GPU it is the first call that needs the device)
    planar_pts = np.stack(np.meshgrid(edges[0], edges[1], indexing="ij"), axis=-1)
    assert cr.best_manifold(planar_pts) == Planar()
    assert cr.resolve_manifold(planar_pts, edges) == Planar()
    assert cr.resolve_manifold(planar_pts, pts) == Spherical()
    if cr.device_count() == 0:
        with pytest.raises(cr.CRError, match="from_planar_polygons"):
            cr.Regridder(edges, edges)


def test_promoted_planar_grid_is_the_lonlat_grid_of_its_edges(cr):
    from importlib import import_module
    M = import_module("crhip.manifolds")
    xs, ys = np.arange(0.0, 361.0, 4.0), np.arange(-90.0, 91.0, 4.0)
    P = M.on_sphere((xs, ys))
    np.testing.assert_array_equal(P, cr.lonlat_points(90, 45))
