"""Generic grid adapter on the device vs the CPU oracle.

Ragged hexagon/pentagon grids (the dual of a subdivided icosahedron, the shape
of ISEA-type grids) as source and as destination, RingGrids full grids
(test/usecases/fullclenshaw.jl), a user tree given only by ncells/getcell
with clockwise closed lon/lat rings (the faces of test/usecases/speedyweather.jl),
and vertex gradients on shared vertices of a ragged grid.
"""
import numpy as np
import pytest
import torch

from gpu_common import pattern_and_values
from polygrids import geodesic_dual, rotation

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _hexes(cr, level, seed):
    v, off, idx = geodesic_dual(level, rotate=rotation(seed))
    return (v, off, idx), cr.polygons_to_cells(v, off, idx)[0]


@pytest.mark.parametrize("hex_is_dst", [True, False])
def test_ragged_grid_parity(cr, orc, hex_is_dst):
    _case("ragged_parity", hex_is_dst=hex_is_dst)
    poly, hcells = _hexes(cr, 4, seed=2)
    P = cr.lonlat_points(72, 36)
    hg, lg = cr.Grid.from_polygons(*poly), cr.Grid.from_points(P)
    assert hg.nv == 6 and hg.ncells == 2562 and hg.nverts == poly[0].shape[0]
    lcells = cr.points_to_cells(P)
    if hex_is_dst:
        R = cr.Regridder(hg, lg, radius=1.0)
        ref = orc.build(hcells, lcells, clip_is_dst=True)
    else:
        R = cr.Regridder(lg, hg, radius=1.0)
        ref = orc.build(lcells, hcells, clip_is_dst=True)
    err, flips = pattern_and_values(R, ref)
    print(f"  nnz {ref.nnz} err {err:.2e} flips {flips}", flush=True)
    # full sphere to full sphere: sums of A are the cell areas; constants stay
    rp, col, val = R.csr_arrays()
    np.testing.assert_allclose(np.add.reduceat(val, rp[:-1]), R.dst_areas, rtol=1.5e-8)
    one = torch.ones(R.shape[1], dtype=torch.float64, device=DEV)
    np.testing.assert_allclose(cr.regrid(R, one).cpu().numpy(), 1.0, atol=1e-10)
    # transpose apply against the oracle's transpose
    x = np.random.default_rng(1).random(R.shape[0])
    y = cr.regrid(cr.transpose(R), torch.tensor(x, device=DEV)).cpu().numpy()
    T = ref.transpose()
    y_ref = orc.spmv(T.rowptr, T.col, T.val, x, area=ref.a_src)
    np.testing.assert_allclose(y, y_ref, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("src_kind,src_half,dst_kind,dst_half", [
    ("clenshaw", 12, "clenshaw", 18), ("gaussian", 12, "clenshaw", 16), ("gaussian", 16, "clenshaw", 24)])
def test_ring_full_grids(cr, orc, src_kind, src_half, dst_kind, dst_half):
    """test/usecases/fullclenshaw.jl: constants preserved (atol 1e-10), the
    area-weighted mean of 2 + 0.5 sin(2 lon) cos(3 lat) conserved (rtol 1e-10)."""
    _case("ring_full", src=(src_kind, src_half), dst=(dst_kind, dst_half))
    s_lat, s_lon = cr.ring_latitudes(src_kind, src_half), cr.ring_longitudes(src_half)
    d_lat, d_lon = cr.ring_latitudes(dst_kind, dst_half), cr.ring_longitudes(dst_half)
    src, dst = cr.Grid.ring_full(s_lat, s_lon), cr.Grid.ring_full(d_lat, d_lon)
    R = cr.Regridder(dst, src, radius=1.0)
    assert R.shape == (d_lat.size * d_lon.size, s_lat.size * s_lon.size)
    ref = orc.build(cr.polygons_to_cells(*cr.ring_full_polygons(d_lat, d_lon))[0],
                    cr.polygons_to_cells(*cr.ring_full_polygons(s_lat, s_lon))[0])
    pattern_and_values(R, ref)
    one = torch.ones(src.ncells, dtype=torch.float64, device=DEV)
    np.testing.assert_allclose(cr.regrid(R, one).cpu().numpy(), 1.0, atol=1e-10)
    lam, phi = np.meshgrid(np.radians(s_lon), np.radians(s_lat))  # (ring, lon): ring-major
    f = (2.0 + 0.5 * np.sin(2 * lam) * np.cos(3 * phi)).reshape(-1)
    d = cr.regrid(R, torch.tensor(f, device=DEV)).cpu().numpy()
    np.testing.assert_allclose(np.sum(d * R.dst_areas), np.sum(f * R.src_areas), rtol=1e-10)


def test_tree_adapter_lonlat_faces(cr, orc):
    """A grid known only through ncells/getcell, with clockwise closed lon/lat
    rings (E, S, W, N, E), builds the same A as the lon-lat point grid."""
    _case("tree_adapter")
    ni, nj = 45, 30
    dl, dp = 360.0 / ni, 180.0 / nj

    class Faces:
        def ncells(self):
            return ni * nj

        def getcell(self, c):
            i, j = c % ni, c // ni
            w, e, s, n = i * dl, (i + 1) * dl, -90.0 + j * dp, -90.0 + (j + 1) * dp
            return [(e, s), (w, s), (w, n), (e, n), (e, s)]  # clockwise seen from outside

    tg = cr.Grid.from_tree(Faces())
    assert tg.ncells == ni * nj and tg.nv == 4
    fg = cr.Grid.from_lonlat_cells(np.array([Faces().getcell(c) for c in range(ni * nj)]))
    dst = cr.Grid.healpix(8)
    ref = orc.build(cr.healpix_cells(8), cr.points_to_cells(cr.lonlat_points(ni, nj)))
    for g in (tg, fg):
        R = cr.Regridder(dst, g, radius=1.0)
        pattern_and_values(R, ref)


def test_ragged_vertex_gradient(cr, orc):
    """dL/dV on the shared vertices of a ragged source grid against central
    differences of the oracle-built loss (every dual vertex sits in 3 cells)."""
    _case("ragged_vertex_grad")
    (v, off, idx), hcells = _hexes(cr, 3, seed=6)
    D = cr.lonlat_points(48, 24, lon0=0.5)
    dcells = cr.points_to_cells(D)
    rng = np.random.default_rng(9)
    s, t = rng.random(off.size - 1), rng.random(48 * 24)
    src, dst = cr.Grid.from_polygons(v, off, idx), cr.Grid.from_points(D)
    R = cr.Regridder(dst, src, radius=1.0)
    st = torch.tensor(s, device=DEV)
    d = cr.regrid(R, st)
    gy = 2.0 * (d - torch.tensor(t, device=DEV))
    _, _, g_src = cr.regrid_backward(R, dst, src, st, d, gy, field_grad=False)
    g_src = g_src.cpu().numpy()
    assert g_src.shape == v.shape

    def loss(vv):
        ref = orc.build(dcells, cr.polygons_to_cells(vv, off, idx)[0])
        return float(np.sum((ref.apply(s) - t) ** 2))

    h, worst = 1e-6, 0.0
    for p in rng.choice(v.shape[0], 12, replace=False):
        for k in range(3):
            vp, vm = v.copy(), v.copy()
            vp[p, k] += h
            vm[p, k] -= h
            vp[p] /= np.linalg.norm(vp[p])
            vm[p] /= np.linalg.norm(vm[p])
            fd = (loss(vp) - loss(vm)) / (2 * h)
            scale = max(abs(fd), np.abs(g_src[p]).max(), 1e-3)
            worst = max(worst, abs(g_src[p, k] - fd) / scale)
    assert worst < 1e-5, worst


def test_hexagon_clip_side_gradient(cr, orc):
    """Vertex gradients with hexagons and pentagons (6 slots, pentagons padded)
    as the clip side: the intersection areas do not depend on which side
    clips, so dL/dV of both grids equals the quad-clipped result, and central
    differences of the oracle's loss built with the hexagons clipping."""
    _case("hexagon_clip_grad")
    (v, off, idx), hcells = _hexes(cr, 3, seed=7)
    D = cr.lonlat_points(48, 24, lon0=0.5)
    dcells = cr.points_to_cells(D)
    rng = np.random.default_rng(11)
    s, t = rng.random(off.size - 1), rng.random(48 * 24)
    src, dst = cr.Grid.from_polygons(v, off, idx), cr.Grid.from_points(D)
    st = torch.tensor(s, device=DEV)
    out = {}
    for side in ("dst", "src"):
        R = cr.Regridder(dst, src, radius=1.0, clip_side=side)
        assert R.stats()["clip_with_dst"] == (side == "dst")
        d = cr.regrid(R, st)
        gy = 2.0 * (d - torch.tensor(t, device=DEV))
        _, g_dst, g_src = cr.regrid_backward(R, dst, src, st, d, gy, field_grad=False)
        out[side] = (g_dst.cpu().numpy(), g_src.cpu().numpy())
    for a, b in zip(out["dst"], out["src"]):
        np.testing.assert_allclose(b, a, rtol=0, atol=1e-9 * np.abs(a).max())
    g_src = out["src"][1]

    def loss(vv):
        ref = orc.build(dcells, cr.polygons_to_cells(vv, off, idx)[0], clip_is_dst=False)
        return float(np.sum((ref.apply(s) - t) ** 2))

    h, worst = 1e-6, 0.0
    for p in rng.choice(v.shape[0], 8, replace=False):
        for k in range(3):
            vp, vm = v.copy(), v.copy()
            vp[p, k] += h
            vm[p, k] -= h
            vp[p] /= np.linalg.norm(vp[p])
            vm[p] /= np.linalg.norm(vm[p])
            fd = (loss(vp) - loss(vm)) / (2 * h)
            scale = max(abs(fd), np.abs(g_src[p]).max(), 1e-3)
            worst = max(worst, abs(g_src[p, k] - fd) / scale)
    assert worst < 1e-5, worst
