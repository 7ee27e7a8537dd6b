"""Whole-grid invariants of the CPU oracle build (what pins it where no
reference value exists: row/col sums vs geometric areas as in
lib/ConservativeRegriddingTestHelpers/src/ConservativeRegriddingTestHelpers.jl:70-73,
conservation as in test/usecases/*.jl, the candidate superset of
test/trees/quadtree_cursors.jl:55-81)."""
import math

import numpy as np
import pytest
import scipy.sparse as sp


def _csr(R):
    return sp.csr_matrix((R.val, R.col, R.rowptr), shape=(R.n_dst, R.n_src))


@pytest.fixture(scope="module")
def c1(cr, orc):
    """4x4 degree source -> 5x5 degree destination (poles included)."""
    src = cr.points_to_cells(cr.lonlat_points(90, 45))
    dst = cr.points_to_cells(cr.lonlat_points(72, 36))
    return src, dst, orc.build(dst, src, clip_is_dst=True)


def test_c1_exact_nnz(c1):
    """nnz(A) = 12384, counted by hand.

    Lon-lat rectangle overlaps: per 20 degrees, 4-degree and 5-degree intervals
    overlap in 8 pairs, so 144 longitude pairs x 72 latitude pairs = 10368.
    Great-circle slivers: at the 8 shared non-equatorial latitudes
    (+-20, 40, 60, 80) the edges of both grids bulge poleward by amounts that
    differ, so near every 5-degree vertex inside a 4-degree span (54 per
    circle) the source cell below overlaps the 2 destination cells above, and
    near every 4-degree vertex inside a 5-degree span (72 per circle) the 2
    source cells above overlap the destination cell below: 2*54 + 2*72 = 252
    pairs per latitude, 2016 in total.  The equator is itself a great circle
    and adds none.
    """
    _, _, R = c1
    assert R.nnz == 144 * 72 + 8 * (2 * 54 + 2 * 72) == 12384


def test_c1_row_col_sums_equal_cell_areas(c1):
    _, _, R = c1
    A = _csr(R)
    np.testing.assert_allclose(np.asarray(A.sum(1)).ravel(), R.a_dst, rtol=1e-14)
    np.testing.assert_allclose(np.asarray(A.sum(0)).ravel(), R.a_src, rtol=1e-14)
    assert A.sum() == pytest.approx(4 * math.pi, rel=1e-14)
    assert R.a_src.sum() == pytest.approx(4 * math.pi, rel=1e-14)


def test_c1_constant_field_and_conservation(c1, orc):
    _, _, R = c1
    ones = np.ones(R.n_src)
    np.testing.assert_allclose(R.apply(ones), 1.0, atol=1e-13)
    s = np.random.default_rng(0).random(R.n_src)
    d = R.apply(s)
    assert (d * R.a_dst).sum() == pytest.approx((s * R.a_src).sum(), rel=1e-13)
    T = R.transpose()
    back = T.apply(d)
    assert (back * R.a_src).sum() == pytest.approx((d * R.a_dst).sum(), rel=1e-13)


def test_swapping_grids_transposes_a(cr, orc, c1):
    src, dst, R = c1
    Rs = orc.build(src, dst, clip_is_dst=True)
    A, B = _csr(R), _csr(Rs).T.tocsr()
    B.sort_indices()
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    scale = np.minimum(R.a_dst[R.row_index()], R.a_src[R.col])
    assert np.max(np.abs(A.data - B.data) / scale) < 1e-14


def test_clip_side_does_not_change_a(c1, orc):
    src, dst, R = c1
    R2 = orc.build(dst, src, clip_is_dst=False)
    assert np.array_equal(R.rowptr, R2.rowptr) and np.array_equal(R.col, R2.col)
    scale = np.minimum(R.a_dst[R.row_index()], R.a_src[R.col])
    assert np.max(np.abs(R.val - R2.val) / scale) < 1e-14


def test_self_regrid_is_diagonal(cr, orc):
    cells = cr.points_to_cells(cr.lonlat_points(36, 18))
    R = orc.build(cells, cells)
    assert R.nnz == cells.shape[0]
    np.testing.assert_array_equal(R.col, np.arange(cells.shape[0]))
    np.testing.assert_allclose(R.val, R.a_dst, rtol=1e-14)


def test_threshold_zero_keeps_a_superset(c1, orc):
    src, dst, R = c1
    R0 = orc.build(dst, src, threshold_rel=0.0)
    assert R0.nnz >= R.nnz
    A0, A = _csr(R0), _csr(R)
    # every extra entry is below the drop threshold; kept ones agree exactly
    diff = (A0 - A).tocoo()
    amin = np.minimum(R.a_dst[diff.row], R.a_src[diff.col])
    assert np.all(np.abs(diff.data) <= 1e-14 * amin)


def test_row_range_sample_matches_full_build(c1, orc):
    src, dst, R = c1
    Rp = orc.build(dst, src, row_range=(100, 700))
    lo, hi = R.rowptr[100], R.rowptr[700]
    assert np.array_equal(Rp.col, R.col[lo:hi])
    np.testing.assert_array_equal(Rp.val, R.val[lo:hi])


def test_normalize_divides_by_max(c1):
    _, _, R = c1
    N = R.normalized()
    assert N.val.max() == 1.0
    s = np.random.default_rng(1).random(R.n_src)
    np.testing.assert_allclose(N.apply(s), R.apply(s), rtol=1e-14)


def test_candidates_are_a_superset_of_overlaps(c1):
    """Every kept pair was a candidate: nnz <= candidates (quadtree_cursors.jl:55-81)."""
    _, _, R = c1
    assert R.n_candidates >= R.nnz


def test_healpix_vs_lonlat_partitions(cr, orc):
    """HEALPix step-4 16-gons (non-convex) clipped by lon-lat cells."""
    dst = cr.healpix_cells(8, "nested", 4)
    src = cr.points_to_cells(cr.lonlat_points(72, 36))
    R = orc.build(dst, src, clip_is_dst=False)
    A = _csr(R)
    np.testing.assert_allclose(np.asarray(A.sum(1)).ravel(), R.a_dst, rtol=1e-13)
    np.testing.assert_allclose(np.asarray(A.sum(0)).ravel(), R.a_src, rtol=1e-13)
