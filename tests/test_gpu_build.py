"""Device Regridder build vs the CPU oracle (sparsity pattern, areas to 1e-12
of the cell area, cell areas to 1e-12), on the configurations' grids."""
import math

import numpy as np
import pytest

from gpu_common import jittered_cubed_sphere, pattern_and_values

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def test_c1_lonlat4_to_lonlat5_normalized(cr, orc):
    _print_case("c1", src="4deg", dst="5deg", normalize=True)
    sp_, dp_ = cr.lonlat_points(90, 45), cr.lonlat_points(72, 36)
    src, dst = cr.Grid.from_points(sp_), cr.Grid.from_points(dp_)
    R = cr.Regridder(dst, src, normalize=True, radius=1.0)
    ref = orc.build(cr.points_to_cells(dp_), cr.points_to_cells(sp_)).normalized()
    err, flips = pattern_and_values(R, ref)
    assert R.nnz == 12384 and flips == 0
    assert R.intersections.max() == 1.0
    st = R.stats()
    assert st["n_candidates"] >= R.nnz


def test_c2_lonlat1_to_c48(cr, orc):
    _print_case("c2", src="1deg", dst="C48")
    sp_, cp = cr.lonlat_points(360, 180), cr.cubed_sphere_points(48)
    R = cr.Regridder(cr.Grid.from_points(cp), cr.Grid.from_points(sp_), radius=1.0)
    ref = orc.build(cr.points_to_cells(cp), cr.points_to_cells(sp_))
    err, flips = pattern_and_values(R, ref)
    A = R.intersections
    np.testing.assert_allclose(np.asarray(A.sum(1)).ravel(), R.dst_areas, rtol=1e-12)
    np.testing.assert_allclose(np.asarray(A.sum(0)).ravel(), R.src_areas, rtol=1e-12)


def test_lonlat_to_healpix_16gons(cr, orc):
    _print_case("hpx32_step4", src="2deg", dst="healpix32 step4")
    sp_ = cr.lonlat_points(180, 90)
    dst = cr.Grid.healpix(32, "nested", step=4)
    assert not dst.convex
    R = cr.Regridder(dst, cr.Grid.from_points(sp_), radius=1.0)
    assert R.stats()["clip_with_dst"] == 0
    ref = orc.build(dst.cells(), cr.points_to_cells(sp_), clip_is_dst=False)
    pattern_and_values(R, ref)


def test_healpix_source_to_lonlat(cr, orc):
    _print_case("hpx_src", src="healpix16 ring step4", dst="3deg")
    dp_ = cr.lonlat_points(120, 60)
    src = cr.Grid.healpix(16, "ring", step=4)
    R = cr.Regridder(cr.Grid.from_points(dp_), src, radius=1.0)
    assert R.stats()["clip_with_dst"] == 1
    ref = orc.build(cr.points_to_cells(dp_), src.cells(), clip_is_dst=True)
    pattern_and_values(R, ref)


def test_c4_jittered_cubed_sphere_to_lonlat1(cr, orc):
    _print_case("c4", src="C96 jittered seed 4", dst="1deg")
    P = jittered_cubed_sphere(cr, 96, seed=4)
    dp_ = cr.lonlat_points(360, 180)
    R = cr.Regridder(cr.Grid.from_points(dp_), cr.Grid.from_points(P), radius=1.0)
    ref = orc.build(cr.points_to_cells(dp_), cr.points_to_cells(P))
    pattern_and_values(R, ref)
    assert R.stats()["flipped_cells"] == 0


def test_c3_full_size(cr, orc):
    _print_case("c3", src="0.25deg", dst="healpix256 nested step4")
    sp_ = cr.lonlat_points(1440, 720)
    dst = cr.Grid.healpix(256, "nested", step=4)
    R = cr.Regridder(dst, cr.Grid.from_points(sp_), radius=1.0)
    ref = orc.build(dst.cells(), cr.points_to_cells(sp_), clip_is_dst=False)
    err, flips = pattern_and_values(R, ref)
    print(f"c3 nnz={R.nnz} err={err:.2e} flips={flips} stats={R.stats()}", flush=True)


def test_device_healpix_matches_host_generator(cr):
    _print_case("hpx_gen")
    for order in ("nested", "ring"):
        g = cr.Grid.healpix(16, order, step=4)
        np.testing.assert_allclose(g.cells(), cr.healpix_cells(16, order, 4), atol=4e-16, rtol=0)


def test_clockwise_cells_are_reoriented(cr, orc):
    _print_case("cw")
    sp_, dp_ = cr.lonlat_points(36, 18), cr.lonlat_points(24, 12)
    cells = cr.points_to_cells(sp_)
    cw = cells[:, ::-1, :].copy()
    g = cr.Grid.from_cells(cw, convex=True)
    R1 = cr.Regridder(cr.Grid.from_points(dp_), g, radius=1.0)
    R2 = cr.Regridder(cr.Grid.from_points(dp_), cr.Grid.from_points(sp_), radius=1.0)
    assert R1.stats()["flipped_cells"] == cells.shape[0]
    A1, A2 = R1.intersections, R2.intersections
    assert (A1 != A2).nnz == 0 or np.max(np.abs((A1 - A2).data)) < 1e-15


def test_radius_scales_areas(cr):
    _print_case("radius")
    sp_, dp_ = cr.lonlat_points(36, 18), cr.lonlat_points(24, 12)
    R1 = cr.Regridder(cr.Grid.from_points(dp_), cr.Grid.from_points(sp_), radius=1.0)
    Re = cr.Regridder(cr.Grid.from_points(dp_), cr.Grid.from_points(sp_))
    r2 = cr.EARTH_RADIUS ** 2
    np.testing.assert_allclose(Re.dst_areas, R1.dst_areas * r2, rtol=1e-15)
    np.testing.assert_allclose(Re.intersections.data, R1.intersections.data * r2, rtol=1e-15)
    assert Re.src_areas.sum() == pytest.approx(4 * math.pi * r2, rel=1e-13)


def test_threshold_zero_keeps_touching_slivers(cr, orc):
    _print_case("thr0")
    sp_, dp_ = cr.lonlat_points(90, 45), cr.lonlat_points(72, 36)
    R0 = cr.Regridder(cr.Grid.from_points(dp_), cr.Grid.from_points(sp_), radius=1.0, threshold_rel=0.0)
    R = cr.Regridder(cr.Grid.from_points(dp_), cr.Grid.from_points(sp_), radius=1.0)
    assert R0.nnz >= R.nnz == 12384


def test_self_regrid_is_diagonal(cr):
    _print_case("self")
    p = cr.lonlat_points(72, 36)
    g = cr.Grid.from_points(p)
    R = cr.Regridder(g, g, radius=1.0)
    A = R.intersections
    assert R.nnz == g.ncells
    np.testing.assert_array_equal(A.indices, np.arange(g.ncells))
    np.testing.assert_allclose(A.diagonal(), R.dst_areas, rtol=1e-14)


def test_transposed_csr_is_consistent(cr):
    _print_case("transpose_csr")
    sp_, cp = cr.lonlat_points(90, 45), cr.cubed_sphere_points(12)
    R = cr.Regridder(cr.Grid.from_points(cp), cr.Grid.from_points(sp_), radius=1.0)
    A = R.intersections
    At = cr.transpose(R).intersections
    assert At.shape == (A.shape[1], A.shape[0])
    assert (At - A.T.tocsr()).nnz == 0 or np.max(np.abs((At - A.T).data)) == 0.0
    assert cr.transpose(R).src_areas is R.dst_areas


@pytest.mark.parametrize("coarse_dst", [True, False])
def test_coarse_cells_with_thousands_of_candidates(cr, orc, coarse_dst):
    """HEALPix nside 2 (48 cells of ~30 deg) against a 1 deg lon-lat grid: each
    coarse cell touches over a thousand fine cells, more than a search list
    holds, so the search's overflow routes (per-cluster walks, cells walked by
    a wave of their own) carry the candidates.  Sums of A are the cell areas."""
    _print_case("coarse_fine", coarse_dst=coarse_dst)
    ll = cr.lonlat_points(360, 180)
    hp = cr.Grid.healpix(2, "nested")
    if coarse_dst:
        R = cr.Regridder(hp, cr.Grid.from_points(ll), radius=1.0)
        ref = orc.build(hp.cells(), cr.points_to_cells(ll))
    else:
        R = cr.Regridder(cr.Grid.from_points(ll), hp, radius=1.0)
        ref = orc.build(cr.points_to_cells(ll), hp.cells())
    pattern_and_values(R, ref)
    A = R.intersections
    np.testing.assert_allclose(np.asarray(A.sum(1)).ravel(), R.dst_areas, rtol=1e-12)
    np.testing.assert_allclose(np.asarray(A.sum(0)).ravel(), R.src_areas, rtol=1e-12)
    if coarse_dst:
        assert np.diff(A.indptr).min() > 256 and np.diff(A.indptr).max() > 1024


def test_zero_area_cells_are_counted(cr, orc):
    """A destination cell collapsed to one point intersects nothing: its
    candidate pairs are dropped and counted, the rest of A equals the oracle's."""
    _print_case("zero_area_cell", src="3deg", dst="2deg + one collapsed cell")
    dcells = cr.points_to_cells(cr.lonlat_points(180, 90))
    ghost = dcells[5000, :1].repeat(4, axis=0)  # every vertex at one corner of cell 5000
    dcells = np.concatenate([dcells, ghost[None]])
    sp_ = cr.lonlat_points(120, 60)
    R = cr.Regridder(cr.Grid.from_cells(dcells, convex=True), cr.Grid.from_points(sp_), radius=1.0)
    st = R.stats()
    assert st["n_zero_area_cells"] == 1 and st["n_zero_cell_pairs"] >= 1
    assert st["n_nonfinite_pairs"] == 0 and st["n_nonfinite_cells"] == 0
    rp, _, _ = R.csr_arrays()
    assert rp[-1] == rp[-2]  # the collapsed cell's row is empty
    ref = orc.build(dcells, cr.points_to_cells(sp_))
    pattern_and_values(R, ref)


def test_point_grid_cells_are_checked_on_the_device(cr):
    """Grid.from_points checks the points on the host (a point that is not a
    finite unit vector fails before anything reaches the device) and the
    cells' geometry on the device: the same errors, naming the same lowest
    cell, as the host check cr_validate_points."""
    import ctypes
    import re
    _print_case("device_point_checks")
    L = cr._lib

    def host(P):
        c, nb, ni, nj = cr.grids._point_blocks(P)
        st = L.lib.cr_validate_points(L.dptr(c), nb, ni, nj, ctypes.byref(ctypes.c_int64()))
        return st, L.last_error()

    base = cr.lonlat_points(36, 18)
    anti = base.copy()
    anti[20, 9] = -anti[21, 9]  # an edge between antipodal points
    concave = base.copy()
    v = concave[7, 5] + 0.8 * (concave[8, 6] - concave[7, 5])  # a corner pulled across its cell
    concave[7, 5] = v / np.linalg.norm(v)
    cell = lambda m: re.search(r"cell (\d+)", m).group(1)  # noqa: E731
    for P in (anti, concave):
        st, msg = host(P)
        assert st == L.CR_DEGENERATE_GEOMETRY
        with pytest.raises(cr.CRError) as e:
            cr.Grid.from_points(P)
        assert e.value.status == L.CR_DEGENERATE_GEOMETRY
        assert cell(msg) == cell(str(e.value))
    for where, how in (((0, 0), "scaled"), ((20, 9), "nan"), ((36, 18), "inf"), ((5, 0), "scaled")):
        P = base.copy()
        if how == "scaled":
            P[where] *= 1.001
        else:
            P[where][1] = np.nan if how == "nan" else np.inf
        P[30, 10] = -P[31, 10]  # a later geometry error: the vertex error is named first
        st, msg = host(P)
        assert st == L.CR_INVALID_ARGUMENT
        with pytest.raises(cr.ArgumentError) as e:
            cr.Grid.from_points(P)
        assert "unit vector" in str(e.value) and cell(msg) == cell(str(e.value)), (where, how)
    assert host(base)[0] == 0
    cr.Grid.from_points(base)
