"""The Planar manifold on the device against the planar oracle and the
reference's planar known answers.

* test/usecases/simple.jl: Regridder(Planar(), squares, diamond + triangles;
  normalize=false) has exactly the hand-worked matrix, row sums == the squares'
  areas, column sums == the sources' areas, conservation == in both directions
  (regrid! forward and through transpose(R));
* test/regridding.jl:299-376: 16 x 16 over 8 x 8 unit grids, size (256, 64),
  every entry 1/256 and sum(A) == 1;
* two tilings of a rectangle (jittered quads, triangles), both directions and
  both clip sides: the device CSR equals the oracle's (pattern exact, areas to
  1e-12 of the smaller cell), areas to 1e-12, FP64 regrid vs the oracle SpMV;
* the Python entry points: edge tuples and planar point matrices build on the
  plane, a planar grid next to a spherical one is promoted (lon/lat degrees),
  spherical-only operations refuse planar grids.
"""
import numpy as np
import pytest
import torch

from gpu_common import pattern_and_values
from planar_cases import SIMPLE_A, jittered_points, jittered_quads, simple_polys, table, triangles, unit_square_grid

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _dense(R):
    rp, col, val = R.csr_arrays()
    A = np.zeros(R.shape)
    for r in range(R.shape[0]):
        A[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    return A


def test_simple_known_answers(cr):
    _case("planar_simple")
    squares, sources = simple_polys()
    dst = cr.Grid.from_planar_rings(squares)
    src = cr.Grid.from_planar_rings(sources)
    assert dst.manifold == cr.Planar() and src.planar
    R = cr.Regridder(dst, src, normalize=False, threshold_rel=0.0)
    A = _dense(R)
    assert np.array_equal(A, SIMPLE_A)
    assert np.array_equal(A.sum(axis=1), R.dst_areas) and np.array_equal(R.dst_areas, np.ones(4))
    assert np.array_equal(A.sum(axis=0), R.src_areas) and np.array_equal(R.src_areas, [2.0, 0.5, 0.5, 0.5, 0.5])
    v2 = torch.tensor([0.0, 0.0, 5.0, 0.0, 0.0], dtype=torch.float64, device=DEV)
    v1 = cr.regrid(R, v2)
    a1 = torch.tensor(R.dst_areas, device=DEV)
    a2 = torch.tensor(R.src_areas, device=DEV)
    assert float((v1 * a1).sum()) == float((v2 * a2).sum())
    back = cr.regrid(cr.transpose(R), v1)
    assert float((back * a2).sum()) == float((v2 * a2).sum())
    s1 = torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=torch.float64, device=DEV)
    d2 = cr.regrid(cr.transpose(R), s1)
    assert d2.cpu().numpy().tolist() == [2.5, 1.0, 3.0, 4.0, 2.0]
    assert float((d2 * a2).sum()) == float((s1 * a1).sum())
    # the explicit manifold and the default threshold give the same matrix here
    R2 = cr.Regridder(dst, src, manifold=cr.Planar())
    assert np.array_equal(_dense(R2), SIMPLE_A)


def test_nested_unit_grids(cr):
    _case("planar_unit_grids", dst="16x16", src="8x8")
    dst = cr.Grid.from_planar_rings(unit_square_grid(16, 16))
    src = cr.Grid.from_planar_rings(unit_square_grid(8, 8))
    R = cr.Regridder(dst, src)
    assert R.shape == (256, 64) and R.nnz == 256
    rp, col, val = R.csr_arrays()
    assert np.all(val == 1.0 / 256) and val.sum() == 1.0
    np.testing.assert_array_equal(R.dst_areas, np.full(256, 1.0 / 256))


@pytest.mark.parametrize("flip", [False, True])
@pytest.mark.parametrize("clip_side", ["auto", "dst", "src"])
def test_tilings_match_oracle(cr, orc, flip, clip_side):
    _case("planar_tilings", flip=flip, clip_side=clip_side)
    q_rings = jittered_quads(23, 17, 3.0, 2.0, seed=5)
    t_rings = triangles(11, 9, 3.0, 2.0)
    dst_r, src_r = (t_rings, q_rings) if flip else (q_rings, t_rings)
    dst, src = cr.Grid.from_planar_rings(dst_r), cr.Grid.from_planar_rings(src_r)
    R = cr.Regridder(dst, src, clip_side=clip_side)
    clip_is_dst = {"auto": True, "dst": True, "src": False}[clip_side]
    ref = orc.build_planar(table(dst_r, nv=4), table(src_r, nv=4), clip_is_dst=clip_is_dst)
    err, flips = pattern_and_values(R, ref, tol=1e-12)
    assert flips == 0
    np.testing.assert_allclose(dst.unit_areas(), ref.a_dst, rtol=1e-12)
    x = np.random.default_rng(1).random(R.shape[1])
    y = cr.regrid(R, torch.tensor(x, device=DEV)).cpu().numpy()
    np.testing.assert_allclose(y, ref.apply(x), rtol=1e-12, atol=1e-14)
    xt = np.random.default_rng(2).random(R.shape[0])
    yt = cr.regrid(cr.transpose(R), torch.tensor(xt, device=DEV)).cpu().numpy()
    np.testing.assert_allclose(yt, ref.transpose().apply(xt), rtol=1e-12, atol=1e-14)
    # conservation: the tilings cover the same rectangle
    assert float(R.dst_areas.sum()) == pytest.approx(6.0, rel=1e-13)


def test_entry_points(cr, orc):
    _case("planar_entry_points")
    xs, ys = np.linspace(0.0, 3.0, 13), np.linspace(0.0, 2.0, 9)
    P = jittered_points(7, 5, 3.0, 2.0, seed=9)
    R = cr.Regridder((xs, ys), P)  # an edge tuple and a planar point matrix: both Planar
    assert R.shape == (12 * 8, 7 * 5)
    ref = orc.build_planar(table([np.array([(xs[i], ys[j]), (xs[i + 1], ys[j]), (xs[i + 1], ys[j + 1]),
                                            (xs[i], ys[j + 1])]) for j in range(8) for i in range(12)]),
                           table([np.array([P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]])
                                  for j in range(5) for i in range(7)]))
    pattern_and_values(R, ref, tol=1e-12)
    # a planar grid next to a spherical one is promoted (its points read as lon/lat degrees)
    lon = np.linspace(0.0, 30.0, 7)
    lat = np.linspace(-10.0, 10.0, 5)
    pg = cr.Grid.from_planar_points(np.stack(np.meshgrid(lon, lat, indexing="ij"), axis=-1))
    sg = cr.Grid.from_points(cr.lonlat_points(6, 4, lon0=0.0, lat0=-10.0, dlon=5.0, dlat=5.0))
    Rs = cr.Regridder(pg, sg, radius=1.0)
    rp, col, val = Rs.csr_arrays()
    assert Rs.nnz == 24 and np.array_equal(col, np.arange(24))  # the same lon-lat cells: a diagonal
    np.testing.assert_allclose(val, Rs.src_areas[col], rtol=1e-12)
    # spherical-only operations refuse planar grids
    dst = cr.Grid.from_planar_rings(unit_square_grid(4, 4))
    src = cr.Grid.from_planar_rings(unit_square_grid(3, 3))
    Rp = cr.Regridder(dst, src)
    x = torch.rand(9, dtype=torch.float64, device=DEV)
    y = cr.regrid(Rp, x)
    # the field adjoint needs no vertex geometry: dL/dx = A^T (gy ./ a_dst)
    gx, gd, gs = cr.regrid_backward(Rp, dst, src, x, y, torch.ones_like(y), vertex_grad=False)
    rp, col, val = Rp.csr_arrays()
    A = np.zeros(Rp.shape)
    for r in range(Rp.shape[0]):
        A[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    np.testing.assert_allclose(gx.cpu().numpy(), A.T @ (1.0 / Rp.dst_areas), rtol=1e-12)
    with pytest.raises(cr.MethodError):
        cr.regrid_backward(Rp, dst, src, x, y, torch.ones_like(y))
    with pytest.raises(cr.MethodError):
        cr.cell_averages(dst, cr.SinusoidField())
    with pytest.raises(cr.ErrorException):
        cr.Regridder(dst, sg, manifold=cr.Planar())
