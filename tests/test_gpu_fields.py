"""Device cell averages of the example fields (cr_cell_averages) against the
oracle's restatement of the test helpers' integrator, and the reference's
sweat-test integral check run on them (test/sweat.jl:93-182, field path of
test/sweat_field.jl):

  * device averages equal the oracle's to 1e-12 of the field's scale, every
    field, on a 4 degree lon-lat grid, HEALPix nside 8 (ring order, 8-gons)
    and a cubed sphere C6 (the longitude field only where no cell straddles
    the 180 degree meridian, across which the field jumps by 360 and where a
    rule point on the meridian may land on either side);
  * for every ordered pair of four grids (2 degree lon-lat, HEALPix 16 nested,
    cubed sphere C16, a RingGrids full Gaussian grid), Regridder(dst, src;
    normalize=false), and every field: sum |regridded| a_dst and sum
    |analytical| a_dst over covered destination cells agree to rtol 1e-2
    (5e-2 for the longitude field when a grid has cells across the 180 degree
    meridian), with the intersection row and column sums equal to the cell
    areas to sqrt(eps).
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


FIELDS = [("longitude", lambda cr: cr.LongitudeField(), {}),
          ("sinusoid", lambda cr: cr.SinusoidField(), {}),
          ("harmonic", lambda cr: cr.HarmonicField(), {}),
          ("gulf_stream", lambda cr: cr.GulfStreamField(), {}),
          ("vortex", lambda cr: cr.VortexField(lat0_rad=math.radians(80)), dict(lat0_rad=math.radians(80)))]


def _grids(cr):
    return {
        "lonlat4": (cr.Grid.from_points(cr.lonlat_points(90, 45)), False),
        "healpix8_ring_8gon": (cr.Grid.healpix(8, "ring", step=2), True),
        "cubed6": (cr.Grid.from_points(cr.cubed_sphere_points(6)), True),
    }


@pytest.mark.parametrize("gname", ["lonlat4", "healpix8_ring_8gon", "cubed6"])
def test_cell_averages_match_oracle(cr, orc, gname):
    grid, crosses = _grids(cr)[gname]
    cells = grid.cells()
    for name, make, kw in FIELDS:
        if name == "longitude" and crosses:
            continue
        _case("cell_averages", grid=gname, field=name)
        got = cr.cell_averages(grid, make(cr)).cpu().numpy()
        f = lambda lon, lat: orc.oracle_field(name, lon, lat, **kw)  # noqa: E731
        want = np.array([orc.cell_average(c, f) for c in cells])
        scale = np.max(np.abs(want))
        assert np.max(np.abs(got - want)) <= 1e-12 * scale, (gname, name)


def test_cell_averages_arguments(cr):
    grid = cr.Grid.from_points(cr.lonlat_points(36, 18))
    a = cr.cell_averages(grid, cr.SinusoidField(coef=3.0))
    b = cr.cell_averages(grid, cr.SinusoidField())
    torch.testing.assert_close(a, b + 1.0, rtol=0, atol=1e-13)  # coef shifts the field
    d4 = cr.cell_averages(grid, cr.HarmonicField(), degree=4)
    d8 = cr.cell_averages(grid, cr.HarmonicField())
    assert float((d4 - d8).abs().max()) < 1e-3  # lower degree, close values
    with pytest.raises(TypeError):
        cr.cell_averages(grid, lambda lon, lat: lon)
    with pytest.raises(cr.MethodError):  # no rule past degree 8 (reference_rule's table)
        cr.cell_averages(grid, cr.HarmonicField(), degree=9)


def _sweat_grids(cr):
    lat = cr.ring_latitudes("gaussian", 24)
    return [("lonlat2", cr.Grid.from_points(cr.lonlat_points(180, 90)), False),
            ("healpix16", cr.Grid.healpix(16, "nested", step=1), True),
            ("cubed16", cr.Grid.from_points(cr.cubed_sphere_points(16)), True),
            ("gaussian24", cr.Grid.ring_full(lat, cr.ring_longitudes(24)), True)]


def test_sweat_integral_conservation(cr):
    grids = _sweat_grids(cr)
    for n1, g1, c1 in grids:
        for n2, g2, c2 in grids:
            if n1 == n2:
                continue
            R = cr.Regridder(g2, g1, radius=1.0, normalize=False)
            A = R.intersections
            np.testing.assert_allclose(np.asarray(A.sum(axis=1)).ravel(), R.dst_areas, rtol=math.sqrt(2.2e-16))
            np.testing.assert_allclose(np.asarray(A.sum(axis=0)).ravel(), R.src_areas, rtol=math.sqrt(2.2e-16))
            covered = torch.tensor(np.asarray(A.sum(axis=1)).ravel() > 0, device=DEV)
            w2 = torch.tensor(R.dst_areas, device=DEV)
            for name, make, _ in FIELDS:
                _case("sweat", src=n1, dst=n2, field=name)
                f = make(cr)
                v1 = cr.cell_averages(g1, f)
                v2 = cr.regrid(R, v1)
                ana = cr.cell_averages(g2, f)
                lhs = float((v2.abs() * w2)[covered].sum())
                rhs = float((ana.abs() * w2)[covered].sum())
                tol = 5e-2 if (name == "longitude" and (c1 or c2)) else 1e-2
                assert lhs == pytest.approx(rhs, rel=tol), (n1, n2, name, lhs, rhs)
