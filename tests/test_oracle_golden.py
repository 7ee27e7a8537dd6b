"""The CPU oracle against golden vectors and closed forms.

* tests/golden/clip_areas.json -- 50-digit brute-force clip areas (vertex
  enumeration + Girard's theorem), see tests/golden/gen_clip_fixtures.py
* tests/golden/readme_example.json -- the reference README's regridder and its
  printed transpose-regrid result (README.md:70-127)
* closed forms: octant, polar triangles, spherical-triangle excess in mpmath
"""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest
import scipy.sparse as sp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _pts(hexes):
    return np.array([[float.fromhex(x) for x in p] for p in hexes])


CLIP = _load("clip_areas.json")["cases"]


@pytest.mark.parametrize("case", CLIP, ids=[c["name"] for c in CLIP])
def test_clip_area_matches_brute_force(orc, case):
    S, C = _pts(case["subject"]), _pts(case["clip"])
    ref = float(case["area"])
    scale = min(float(case["subject_area"]), float(case["clip_area"]))
    got = orc.clip_area(S, C)
    # tolerance relative to the smaller cell, the scale of the 1e-14 drop threshold
    assert abs(got - ref) <= 2e-14 * scale, (got, ref)


@pytest.mark.parametrize("case", CLIP[:12], ids=[c["name"] for c in CLIP[:12]])
def test_cell_areas_match_girard(orc, case):
    for key, arr in (("subject_area", case["subject"]), ("clip_area", case["clip"])):
        assert abs(orc.polygon_area(_pts(arr)) - float(case[key])) <= 1e-15 * max(1.0, float(case[key]))


def test_octant_is_half_pi(orc):
    assert orc.polygon_area(np.eye(3)) == pytest.approx(math.pi / 2, rel=1e-15)
    # clockwise ring gives the negative area
    assert orc.polygon_area(np.eye(3)[::-1]) == pytest.approx(-math.pi / 2, rel=1e-15)


@pytest.mark.parametrize("dlon", [1.0, 7.5, 45.0, 90.0])
def test_equator_to_pole_triangle_is_dlon(orc, dlon):
    # cell [0, dlon] x [0, 90]: two right angles on the equator, apex angle dlon
    lam = math.radians(dlon)
    ring = np.array([[1, 0, 0], [math.cos(lam), math.sin(lam), 0], [0, 0, 1], [0, 0, 1]])
    assert orc.polygon_area(ring) == pytest.approx(lam, rel=1e-14)


def test_random_triangles_match_mpmath_excess(orc):
    rng = np.random.default_rng(7)
    mp.mp.dps = 40
    for _ in range(50):
        c = rng.normal(size=3)
        c /= np.linalg.norm(c)
        size = 10 ** rng.uniform(-4, -0.5)
        v = c + size * rng.normal(size=(3, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        a, b, d = ([mp.mpf(float(x)) for x in p] for p in v)
        na, nb, nd = (mp.sqrt(sum(x * x for x in p)) for p in (a, b, d))
        a, b, d = ([x / n for x in p] for p, n in ((a, na), (b, nb), (d, nd)))
        trip = (a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0])
                + a[2] * (b[0] * d[1] - b[1] * d[0]))
        den = 1 + sum(x * y for x, y in zip(a, b)) + sum(x * y for x, y in zip(b, d)) + sum(
            x * y for x, y in zip(d, a))
        ref = float(2 * mp.atan2(trip, den))
        assert orc.polygon_area(v) == pytest.approx(ref, rel=1e-12, abs=1e-300)


def test_clip_is_symmetric_for_convex_pairs(orc):
    for case in CLIP:
        S, C = _pts(case["subject"]), _pts(case["clip"])
        if len(C) != 4:
            continue
        scale = min(float(case["subject_area"]), float(case["clip_area"]))
        assert abs(orc.clip_area(S, C) - orc.clip_area(C, S)) <= 2e-14 * scale


def test_readme_example_apply_and_transpose_are_exact(orc):
    d = _load("readme_example.json")
    A = sp.csr_matrix(np.array(d["A_dense"]))
    a_dst, a_src = np.array(d["dst_areas"]), np.array(d["src_areas"])
    orange = np.array(d["orange"])
    blue = orc.spmv(A.indptr, A.indices, A.data, orange, a_dst)
    # mean conservation holds exactly for these dyadic weights (README.md:96-105)
    assert (orange * a_src).sum() / a_src.sum() == (blue * a_dst).sum() / a_dst.sum()
    trp, tcol, tval = orc.transpose(A.indptr, A.indices, A.data, 4, 5)
    tilde = orc.spmv(trp, tcol, tval, blue, a_src)
    np.testing.assert_array_equal(tilde, np.array(d["orange_tilde"]))


def test_spmv_without_normalize_returns_raw_product(orc):
    d = _load("readme_example.json")
    A = sp.csr_matrix(np.array(d["A_dense"]))
    x = np.arange(1.0, 6.0)
    np.testing.assert_array_equal(orc.spmv(A.indptr, A.indices, A.data, x), A @ x)
