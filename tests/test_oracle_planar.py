"""The planar oracle (oracle/cr_oracle_planar.c) against what fixes it:

* the known answers of test/usecases/simple.jl (row sums == the squares'
  areas, column sums == the sources' areas, conservation ==, both directions),
  with the matrix worked out by hand (each square meets half the diamond and
  one corner triangle);
* test/regridding.jl:299-376's nested 16 x 16 / 8 x 8 unit grids (every entry
  1/256, sum(A) == 1);
* axis-aligned rectangles: overlap = product of the interval overlaps;
* random convex polygons with integer vertices against an exact rational
  intersection made another way (vertices inside the other polygon plus
  edge-edge crossings, their convex hull, then the shoelace formula);
* a non-convex (L-shaped) subject against the sum of its rectangles' overlaps;
* shoelace areas of triangles and regular polygons (closed forms);
* two tilings of one rectangle: row and column sums equal the cell areas.
"""
from fractions import Fraction

import math

import numpy as np
import pytest

from planar_cases import SIMPLE_A, ccw, jittered_quads, simple_polys, table, triangles, unit_square_grid


def test_simple_known_answers(orc):
    squares, sources = simple_polys()
    R = orc.build_planar(table(squares), table(sources, nv=4), clip_is_dst=True, threshold_rel=0.0)
    A = np.zeros((4, 5))
    for r in range(4):
        A[r, R.col[R.rowptr[r]:R.rowptr[r + 1]]] = R.val[R.rowptr[r]:R.rowptr[r + 1]]
    assert np.array_equal(A, SIMPLE_A)
    assert np.array_equal(A.sum(axis=1), R.a_dst) and np.array_equal(R.a_dst, np.ones(4))
    assert np.array_equal(A.sum(axis=0), R.a_src) and np.array_equal(R.a_src, [2.0, 0.5, 0.5, 0.5, 0.5])
    v2 = np.array([0.0, 0.0, 5.0, 0.0, 0.0])
    v1 = A @ v2 / R.a_dst
    assert np.sum(v1 * R.a_dst) == np.sum(v2 * R.a_src)
    back = A.T @ v1 / R.a_src
    assert np.sum(back * R.a_src) == np.sum(v2 * R.a_src)
    # the transposed regrid of simple.jl:66-74
    s1 = np.array([1.0, 2.0, 3.0, 4.0])
    d2 = A.T @ s1 / R.a_src
    assert np.array_equal(d2, [2.5, 1.0, 3.0, 4.0, 2.0])
    assert np.sum(d2 * R.a_src) == np.sum(s1 * R.a_dst)
    # the other side clipping gives the same matrix
    R2 = orc.build_planar(table(squares), table(sources, nv=4), clip_is_dst=False, threshold_rel=0.0)
    assert np.array_equal(R2.val, R.val) and np.array_equal(R2.col, R.col)


def test_nested_unit_grids(orc):
    R = orc.build_planar(table(unit_square_grid(16, 16)), table(unit_square_grid(8, 8)), threshold_rel=0.0)
    assert R.nnz == 256
    assert np.all(R.val == 1.0 / 256) and R.val.sum() == 1.0
    assert np.all(np.diff(R.rowptr) == 1)


def test_rectangle_overlaps(orc):
    rng = np.random.default_rng(3)
    for _ in range(300):
        a = np.sort(rng.uniform(-5, 5, 2)), np.sort(rng.uniform(-5, 5, 2))
        b = np.sort(rng.uniform(-5, 5, 2)), np.sort(rng.uniform(-5, 5, 2))
        ra = np.array([(a[0][0], a[1][0]), (a[0][1], a[1][0]), (a[0][1], a[1][1]), (a[0][0], a[1][1])])
        rb = np.array([(b[0][0], b[1][0]), (b[0][1], b[1][0]), (b[0][1], b[1][1]), (b[0][0], b[1][1])])
        ox = max(0.0, min(a[0][1], b[0][1]) - max(a[0][0], b[0][0]))
        oy = max(0.0, min(a[1][1], b[1][1]) - max(a[1][0], b[1][0]))
        got = orc.planar_clip_area(ra, rb)
        assert got == pytest.approx(ox * oy, rel=1e-13, abs=1e-13)


def _hull(points):
    """Convex hull (Andrew's monotone chain), counter-clockwise, exact."""
    pts = sorted(set(points))
    if len(pts) < 3:
        return pts

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])
    lower, upper = [], []
    for p in pts:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(pts):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    return lower[:-1] + upper[:-1]


def _inside(p, poly):
    n = len(poly)
    for k in range(n):
        a, b = poly[k], poly[(k + 1) % n]
        if (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0]) < 0:
            return False
    return True


def _exact_area(P, Q):
    """Exact area of two convex counter-clockwise polygons' intersection."""
    pts = [p for p in P if _inside(p, Q)] + [q for q in Q if _inside(q, P)]
    for i in range(len(P)):
        p1, p2 = P[i], P[(i + 1) % len(P)]
        for j in range(len(Q)):
            q1, q2 = Q[j], Q[(j + 1) % len(Q)]
            r = (p2[0] - p1[0], p2[1] - p1[1])
            s = (q2[0] - q1[0], q2[1] - q1[1])
            den = r[0] * s[1] - r[1] * s[0]
            if den == 0:
                continue
            qp = (q1[0] - p1[0], q1[1] - p1[1])
            t = Fraction(qp[0] * s[1] - qp[1] * s[0], den)
            u = Fraction(qp[0] * r[1] - qp[1] * r[0], den)
            if 0 <= t <= 1 and 0 <= u <= 1:
                pts.append((p1[0] + t * r[0], p1[1] + t * r[1]))
    h = _hull([(Fraction(x), Fraction(y)) for x, y in pts])
    if len(h) < 3:
        return Fraction(0)
    return sum(h[i][0] * h[(i + 1) % len(h)][1] - h[(i + 1) % len(h)][0] * h[i][1] for i in range(len(h))) / 2


def test_convex_pairs_exact(orc):
    rng = np.random.default_rng(7)
    done = 0
    while done < 150:
        P = _hull([tuple(int(v) for v in rng.integers(0, 21, 2)) for _ in range(8)])
        Q = _hull([tuple(int(v) for v in rng.integers(0, 21, 2)) for _ in range(8)])
        if len(P) < 3 or len(Q) < 3:
            continue
        want = float(_exact_area(P, Q))
        got = orc.planar_clip_area(np.array(P, float), np.array(Q, float))
        assert got == pytest.approx(want, rel=1e-12, abs=1e-12), (P, Q)
        done += 1


def test_nonconvex_subject(orc):
    # L = [0,4]x[0,1] + [0,1]x[1,4], counter-clockwise
    L = np.array([(0, 0), (4, 0), (4, 1), (1, 1), (1, 4), (0, 4)], float)
    rng = np.random.default_rng(11)
    for _ in range(200):
        x0, y0 = rng.uniform(-1, 4, 2)
        w, h = rng.uniform(0.2, 3, 2)
        sq = np.array([(x0, y0), (x0 + w, y0), (x0 + w, y0 + h), (x0, y0 + h)])

        def ov(a0, a1, b0, b1):
            return max(0.0, min(a1, b1) - max(a0, b0))
        want = ov(0, 4, x0, x0 + w) * ov(0, 1, y0, y0 + h) + ov(0, 1, x0, x0 + w) * ov(1, 4, y0, y0 + h)
        assert orc.planar_clip_area(L, sq) == pytest.approx(want, rel=1e-12, abs=1e-12)


def test_shoelace_closed_forms(orc):
    rng = np.random.default_rng(2)
    for _ in range(50):
        t = rng.uniform(-3, 3, (3, 2))
        want = 0.5 * abs((t[1, 0] - t[0, 0]) * (t[2, 1] - t[0, 1]) - (t[1, 1] - t[0, 1]) * (t[2, 0] - t[0, 0]))
        assert orc.planar_polygon_area(ccw(t)) == pytest.approx(want, rel=1e-13)
    for n in (3, 5, 8, 17):
        r = 2.5
        ring = np.array([(r * math.cos(2 * math.pi * k / n), r * math.sin(2 * math.pi * k / n)) for k in range(n)])
        assert orc.planar_polygon_area(ring) == pytest.approx(0.5 * n * r * r * math.sin(2 * math.pi / n), rel=1e-14)
        assert orc.planar_polygon_area(ring[::-1]) == pytest.approx(-0.5 * n * r * r * math.sin(2 * math.pi / n),
                                                                    rel=1e-14)


def test_two_tilings_partition(orc):
    q = table(jittered_quads(23, 17, 3.0, 2.0, seed=5))
    t = table(triangles(11, 9, 3.0, 2.0), nv=4)
    for dst, src, clip_is_dst in ((q, t, True), (t, q, False), (q, t, False)):
        R = orc.build_planar(dst, src, clip_is_dst=clip_is_dst, threshold_rel=0.0)
        rows = np.add.reduceat(R.val, R.rowptr[:-1]) if R.nnz else np.zeros(R.n_dst)
        # the shoelace of coordinates ~3 rounds at ~1e-15 absolute, ~1e-13 of these cells
        np.testing.assert_allclose(rows, R.a_dst, rtol=1e-12)
        cols = np.bincount(R.col, weights=R.val, minlength=R.n_src)
        np.testing.assert_allclose(cols, R.a_src, rtol=1e-12)
        assert R.a_dst.sum() == pytest.approx(6.0, rel=1e-14)
