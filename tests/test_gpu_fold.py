"""Tripolar grids with a RightCenterFolded top row on the device vs the oracle.

The reference's treeify keeps one polygon per physical fold-row cell and pads
the duplicates with zero-area ghosts (ext/ConservativeRegriddingOceananigansExt.jl:119-187);
finalize_regridding! mirrors each primary into its partner (:207-240).  The
reference tests (test/usecases/oceananigans.jl:60-80) regrid tripolar ->
lat-lon and check conservation to rtol 1e-7; here the synthetic folded grid
of ``polygrids.folded_points`` (Oceananigans is not in the image) is checked
in both directions against the oracle built from the same cells.
"""
import numpy as np
import pytest
import torch

from gpu_common import pattern_and_values
from polygrids import fold_primary, folded_cells, folded_points

pytestmark = pytest.mark.gpu
DEV = "cuda"
NX, NY = 72, 30


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _pairs():
    nh, nq = NX // 2, NX // 4
    prim = np.concatenate([np.arange(nq), np.arange(nh, nh + nq)])
    off = (NY - 1) * NX
    return off + prim, off + NX - 1 - prim


@pytest.fixture(scope="module")
def folded(cr):
    pts = folded_points(cr, NX, NY, 62.0)
    cells, ghost = folded_cells(cr, pts)
    return pts, cells, ghost


def test_tripolar_grid_cells(cr, folded):
    _case("tripolar_cells")
    pts, cells, ghost = folded
    g = cr.Grid.tripolar(pts)
    assert g.ncells == NX * NY and g.nfold == 2 * (NX // 4) and g.nverts == (NX + 1) * (NY + 1)
    np.testing.assert_array_equal(g.cells(), cells)
    a = g.unit_areas()
    assert np.all(a[ghost] == 0.0) and np.all(np.delete(a, ghost) > 0)
    with pytest.raises(cr.ArgumentError):
        cr.Grid.tripolar(pts[:-1])  # odd nx


def test_tripolar_source(cr, orc, folded):
    """Tripolar -> lat-lon (the reference's use case): parity, ghost columns
    empty, conservation of a random field."""
    _case("tripolar_source")
    pts, cells, ghost = folded
    src, dst = cr.Grid.tripolar(pts), cr.Grid.lonlat(90, 45)
    R = cr.Regridder(dst, src, radius=1.0)
    ref = orc.build(cr.points_to_cells(cr.lonlat_points(90, 45)), cells)
    pattern_and_values(R, ref)
    _, col, _ = R.csr_arrays()
    assert not np.isin(col, ghost).any()
    s = np.random.default_rng(3).random(NX * NY)
    d = cr.regrid(R, torch.tensor(s, device=DEV)).cpu().numpy()
    np.testing.assert_allclose(np.sum(d * R.dst_areas), np.sum(s * R.src_areas), rtol=1e-10)
    np.testing.assert_allclose(cr.regrid(R, torch.ones(NX * NY, dtype=torch.float64, device=DEV)).cpu().numpy(),
                               1.0, atol=1e-10)


def test_tripolar_destination_mirror(cr, orc, folded):
    """Lat-lon -> tripolar: ghosts receive their primary's value (FP64 and FP32
    batched, forward and through transpose(R) of the reverse regridder);
    without the mirror they hold the plain-vector result (NaN with normalize)."""
    _case("tripolar_destination")
    pts, cells, ghost = folded
    trip, ll = cr.Grid.tripolar(pts), cr.Grid.lonlat(90, 45)
    R = cr.Regridder(trip, ll, radius=1.0)
    ref = orc.build(cells, cr.points_to_cells(cr.lonlat_points(90, 45)))
    pattern_and_values(R, ref)
    prim, part = _pairs()
    s = np.random.default_rng(4).random(90 * 45)
    st = torch.tensor(s, device=DEV)
    d = cr.regrid(R, st).cpu().numpy()
    d_ref = ref.apply(s)
    real = np.setdiff1d(np.arange(NX * NY), part)
    np.testing.assert_allclose(d[real], d_ref[real], rtol=1e-12)
    np.testing.assert_array_equal(d[part], d[prim])
    np.testing.assert_allclose(np.sum(d * R.dst_areas), np.sum(s * R.src_areas), rtol=1e-10)
    plain = cr.regrid(R, st, mirror_folds=False).cpu().numpy()
    assert np.all(np.isnan(plain[part])) and np.array_equal(plain[real], d[real])
    # transpose(R') writes into the tripolar grid too (its source side)
    Rb = cr.Regridder(ll, trip, radius=1.0)
    dt = cr.regrid(cr.transpose(Rb), st).cpu().numpy()
    np.testing.assert_array_equal(dt[part], dt[prim])
    np.testing.assert_allclose(dt[real], d[real], rtol=1e-12)
    # FP32, 8 fields cells-fastest, one launch + one mirror launch
    S = torch.tensor(np.random.default_rng(5).random((8, 90 * 45)), device=DEV, dtype=torch.float32).T
    Y = torch.zeros((8, NX * NY), device=DEV, dtype=torch.float32).T
    cr.regrid_(Y, R, S)
    Y = Y.cpu().numpy()
    np.testing.assert_array_equal(Y[part], Y[prim])
    Y64 = np.stack([ref.apply(S[:, f].double().cpu().numpy()) for f in range(8)], axis=1)
    np.testing.assert_allclose(Y[real], Y64[real], rtol=1e-6)
    # a plan (graph-capturable launch) mirrors too
    Z = torch.zeros(NX * NY, device=DEV, dtype=torch.float64)
    cr.regrid_plan(Z, R, st)()
    np.testing.assert_array_equal(Z.cpu().numpy(), d)


def test_tripolar_destination_gradient(cr, folded):
    """Reverse mode through the mirror: a ghost's upstream gradient flows to
    its primary.  L = ||y - t||^2 is quadratic in s, so central differences
    through the device forward are exact up to rounding."""
    _case("tripolar_gradient")
    pts, cells, ghost = folded
    trip, ll = cr.Grid.tripolar(pts), cr.Grid.lonlat(36, 18)
    R = cr.Regridder(trip, ll, radius=1.0)
    rng = np.random.default_rng(6)
    s = torch.tensor(rng.random(36 * 18), device=DEV)
    t = torch.tensor(rng.random(NX * NY), device=DEV)
    y = cr.regrid(R, s)
    gy = 2.0 * (y - t)
    gx, g_dst, g_src = cr.regrid_backward(R, trip, ll, s, y, gy)
    assert torch.isfinite(g_dst).all() and torch.isfinite(g_src).all()

    def loss(sv):
        return float(((cr.regrid(R, sv) - t) ** 2).sum())

    h = 1e-3
    for j in rng.choice(36 * 18, 16, replace=False):
        e = torch.zeros_like(s)
        e[j] = h
        fd = (loss(s + e) - loss(s - e)) / (2 * h)
        assert abs(fd - float(gx[j])) <= 1e-8 * max(1.0, abs(fd)), (j, fd, float(gx[j]))
    # without the mirror's pullback the ghosts' gradient would be lost
    gx0, _, _ = cr.regrid_backward(R, trip, ll, s, y, gy, mirror_folds=False, vertex_grad=False)
    assert not torch.allclose(gx0, gx)
