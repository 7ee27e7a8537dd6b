"""The oracle's example fields (src/utils/example_data.jl) and its restatement
of the test helpers' cell-average integrator, pinned by closed forms; and the
package's host-side field callables against the oracle.

Closed forms: the harmonic field at 45 N (3 and 1 at cos 16 lon = +-1) and on
the equator (2); the sinusoid at (0, 0) (coef - 1) and at (180, 0)
(coef - cos(pi / 1.2)); the vortex at its rotated pole (2); the Gulf Stream
field far from the stream equals the sinusoid's fcos term and never falls
below it.  Integrator: int dOmega of small cells = their solid angle (Van
Oosterom-Strackee, orc_polygon_area) to 1e-12; a constant averages to itself;
the harmonic field's integral over the sphere, summed over a 2 degree grid's
cells, is 8 pi (its cos 16 lon part integrates to 0); zero-area cells give 0."""
import math

import numpy as np
import pytest

from oracle import oracle as o


def _ring(lon0, lon1, lat0, lat1):
    pts = [(lon0, lat0), (lon1, lat0), (lon1, lat1), (lon0, lat1)]
    return np.array([[math.cos(math.radians(b)) * math.cos(math.radians(a)),
                      math.cos(math.radians(b)) * math.sin(math.radians(a)), math.sin(math.radians(b))]
                     for a, b in pts])


def test_field_closed_forms():
    assert o.oracle_field("longitude", 123.5, -7.0) == 123.5
    assert o.oracle_field("harmonic", 0.0, 45.0) == pytest.approx(3.0, abs=1e-14)
    assert o.oracle_field("harmonic", 11.25, 45.0) == pytest.approx(1.0, abs=1e-14)
    assert o.oracle_field("harmonic", 77.0, 0.0) == pytest.approx(2.0, abs=1e-14)
    assert o.oracle_field("sinusoid", 0.0, 0.0) == pytest.approx(1.0, abs=1e-14)
    assert o.oracle_field("sinusoid", 180.0, 0.0) == pytest.approx(2.0 - math.cos(math.pi / 1.2), abs=1e-13)
    assert o.oracle_field("sinusoid", 0.0, 0.0, coef=3.0, coefmult=2.0) == pytest.approx(4.0, abs=1e-14)
    pole = (math.degrees(5.5), math.degrees(0.2))
    assert o.oracle_field("vortex", *pole) == pytest.approx(2.0, abs=1e-7)
    lat0 = math.radians(80.0)
    assert o.oracle_field("vortex", math.degrees(5.5), 80.0, lat0_rad=lat0) == pytest.approx(2.0, abs=1e-7)
    rng = np.random.default_rng(0)
    lon, lat = rng.uniform(-180, 180, 2000), rng.uniform(-90, 90, 2000)
    gs, fc = o.oracle_field("gulf_stream", lon, lat), o.oracle_field("sinusoid", lon, lat)
    assert np.all(gs >= fc - 1e-12)
    far = (lon > 60) | (lat < -30)
    np.testing.assert_allclose(gs[far], fc[far], rtol=0, atol=1e-13)
    assert np.any(gs > fc + 0.1)  # the current is there


def test_package_fields_match_oracle(cr):
    rng = np.random.default_rng(1)
    lon, lat = rng.uniform(-180, 180, 5000), rng.uniform(-90, 90, 5000)
    pairs = [(cr.LongitudeField(), "longitude", {}), (cr.SinusoidField(), "sinusoid", {}),
             (cr.SinusoidField(dp_length=3.0, coef=1.5, coefmult=0.5), "sinusoid",
              dict(dp_length=3.0, coef=1.5, coefmult=0.5)),
             (cr.HarmonicField(), "harmonic", {}), (cr.GulfStreamField(), "gulf_stream", {}),
             (cr.VortexField(lat0_rad=math.radians(80)), "vortex", dict(lat0_rad=math.radians(80))),
             (cr.VortexField(), "vortex", {})]
    for f, name, kw in pairs:
        np.testing.assert_allclose(f(lon, lat), o.oracle_field(name, lon, lat, **kw), rtol=1e-13, atol=1e-13,
                                   err_msg=name)


def test_integrator_solid_angle_and_constants():
    for cell in (_ring(0, 2, 0, 2), _ring(10, 10.5, 60, 61), _ring(-3, 4, -80, -78)):
        num, den = o.cell_average(cell, lambda lon, lat: np.ones_like(lon), parts=True)
        assert den == pytest.approx(o.polygon_area(cell), rel=1e-12)
        assert num == pytest.approx(den, rel=1e-15)
        assert o.cell_average(cell, lambda lon, lat: np.full_like(lon, 3.25)) == pytest.approx(3.25, rel=1e-14)
    # a zero-area cell (all four corners at one point)
    p = _ring(5, 5, 5, 5)
    assert o.cell_average(p, lambda lon, lat: lon + 1.0) == 0.0


def test_integrator_global_integral_of_harmonic_field():
    total = 0.0
    f = lambda lon, lat: o.oracle_field("harmonic", lon, lat)  # noqa: E731
    for j in range(90):
        for i in range(180):
            cell = _ring(-180 + 2 * i, -180 + 2 * i + 2, -90 + 2 * j, -90 + 2 * j + 2)
            num, _ = o.cell_average(cell, f, parts=True)
            total += num
    assert total == pytest.approx(8.0 * math.pi, rel=1e-9)
