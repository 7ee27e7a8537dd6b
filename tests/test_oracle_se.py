"""Pins of the oracle's spectral-element weights (oracle/cr_oracle_se.c).

ClimaCore is not in the image, so its maps and quadratures are pinned by
their definitions and by the reference's own tests of them:
* Dunavant rules integrate x^a y^b exactly on the reference triangle for
  a + b <= degree, weights sum to 1/2 (test/triangle_quadrature.jl:5-18);
* GLL nodes/weights: closed form for 4 points, exactness to degree 2n-3;
* Lagrange basis: Kronecker delta at nodes, partition of unity off-node
  (test/usecases/climacore.jl:15-47), reproduction of x^k, k < n;
* inverse element map: round trip at GLL nodes and random points
  (test/extensions/climacore.jl:57-82, atol 1e-10; here 1e-13);
* moments: sum over (i, j) = intersection area (test/extensions/climacore.jl:108-131,
  rtol 1e-10; here 1e-13), and, for a cell that is the image of a rectangle
  [xa, xb] x [ya, yb] of the reference square (lines of constant xi or eta map
  to great circles under the normalised bilinear map, so the cell has exact
  great-circle edges), B[i, j] ~ int int phi_i phi_j J dxi deta, computed
  independently by tensor Gauss-Legendre with the analytic Jacobian, to the
  O(h^2) accuracy of the reference's fan rule (checked to converge).
"""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("degree", range(1, 9))
def test_dunavant_exact_monomials(orc, degree):
    bary, w = orc.dunavant(degree)
    assert abs(w.sum() - 0.5) < 1e-12
    np.testing.assert_allclose(bary.sum(axis=1), 1.0, atol=1e-14)
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            got = np.sum(w * bary[:, 1] ** a * bary[:, 2] ** b)
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(got - exact) < 1e-12, (degree, a, b)
    # one degree more is not integrated exactly (the rule is the degree asked)
    worst = max(abs(np.sum(w * bary[:, 1] ** a * bary[:, 2] ** (degree + 1 - a))
                    - math.factorial(a) * math.factorial(degree + 1 - a) / math.factorial(degree + 3))
                for a in range(degree + 2))
    assert worst > 1e-9


def test_gll(orc):
    x, w = orc.gll(4)
    np.testing.assert_allclose(x, [-1, -1 / math.sqrt(5), 1 / math.sqrt(5), 1], atol=1e-15)
    np.testing.assert_allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6], atol=1e-15)
    for n in range(2, 10):
        x, w = orc.gll(n)
        assert x[0] == -1.0 and x[-1] == 1.0 and np.all(np.diff(x) > 0)
        for k in range(2 * n - 2):
            exact = 0.0 if k % 2 else 2.0 / (k + 1)
            assert abs(np.sum(w * x ** k) - exact) < 1e-13, (n, k)


def test_lagrange(orc):
    for n in (3, 4, 6):
        x, _ = orc.gll(n)
        for p in range(n):
            np.testing.assert_allclose(orc.lagrange(x, x[p]), np.eye(n)[p], atol=1e-14)
        for t in (-0.7, -0.3, 0.0, 0.42, 0.91):
            phi = orc.lagrange(x, t)
            assert abs(phi.sum() - 1.0) < 1e-13
            for k in range(n):
                assert abs(np.sum(phi * x ** k) - t ** k) < 1e-13


def _element(cr, ne=4, face=0, i=1, j=2):
    P = cr.cubed_sphere_points(ne)
    return np.array([P[face, i, j], P[face, i + 1, j], P[face, i + 1, j + 1], P[face, i, j + 1]])


def test_inverse_element_map_round_trip(cr, orc):
    rng = np.random.default_rng(1)
    nodes, _ = orc.gll(4)
    for face in range(6):
        E = _element(cr, 4, face, 1, 2)
        pts = [(a, b) for a in nodes for b in nodes] + [tuple(rng.uniform(-1, 1, 2)) for _ in range(20)]
        for xi, eta in pts:
            x = orc.element_map(E, xi, eta)
            assert abs(np.linalg.norm(x) - 1.0) < 1e-15
            gx, ge = orc.inverse_element_map(E, x)
            assert abs(gx - xi) < 1e-13 and abs(ge - eta) < 1e-13, (face, xi, eta, gx, ge)


def test_jacobian_integrates_to_area(orc, cr):
    """Tensor Gauss-Legendre of J over the reference square = element area."""
    E = _element(cr, 4, 2, 0, 3)
    g, gw = np.polynomial.legendre.leggauss(40)
    tot = sum(gw[a] * gw[b] * orc.element_jacobian(E, g[a], g[b]) for a in range(40) for b in range(40))
    assert abs(tot - orc.polygon_area(E)) < 1e-13 * tot


def _rect_cell(orc, E, xa, xb, ya, yb):
    return np.array([orc.element_map(E, xa, ya), orc.element_map(E, xb, ya), orc.element_map(E, xb, yb),
                     orc.element_map(E, xa, yb)])


def _tensor_moments(orc, E, nq, xa, xb, ya, yb, ng=30):
    nodes, _ = orc.gll(nq)
    g, gw = np.polynomial.legendre.leggauss(ng)
    xs, ys = 0.5 * (xb - xa) * g + 0.5 * (xb + xa), 0.5 * (yb - ya) * g + 0.5 * (yb + ya)
    B = np.zeros((nq, nq))
    for a in range(ng):
        px = orc.lagrange(nodes, xs[a])
        for b in range(ng):
            py = orc.lagrange(nodes, ys[b])
            wj = gw[a] * gw[b] * 0.25 * (xb - xa) * (yb - ya) * orc.element_jacobian(E, xs[a], ys[b])
            B += wj * np.outer(px, py)
    return B


@pytest.mark.parametrize("rect,tol", [((-1, 1, -1, 1), 3e-4), ((-0.6, 0.35, -0.9, -0.2), 5e-6),
                                      ((0.1, 0.9, -0.3, 0.95), 1e-5)])
def test_moments_of_mapped_rectangles(cr, orc, rect, tol):
    """B of a cell that is the image of a reference-square rectangle against
    the tensor-product integral.  The fan rule weights each triangle's points
    by 2 A_t w without the radial projection's Jacobian, so B is exact in its
    sum but carries an O(h^2) error per entry (h = triangle size); the
    tolerances are that error at 2.8 degree elements, far below what a
    transposed (i, j), a swapped (xi, eta) or a missing factor would give."""
    E = _element(cr, 32, 4, 3, 5)
    xa, xb, ya, yb = rect
    cell = _rect_cell(orc, E, xa, xb, ya, yb)
    for nq in (3, 4):
        B = orc.se_moments(E, cell, nq)
        ref = _tensor_moments(orc, E, nq, xa, xb, ya, yb)
        m = np.max(np.abs(ref))
        assert abs(B.sum() - orc.clip_area(E, cell)) < 1e-13 * abs(B.sum())
        assert np.max(np.abs(B - ref)) < tol * m
        if rect != (-1, 1, -1, 1):  # a transposed B would fail
            assert np.max(np.abs(B - B.T)) > 10 * tol * m
    # the per-entry error falls with the element size (about as h^2; the whole
    # element, whose fan has only 4 triangles, at least as h): 4x larger
    # elements give 5-20x the error, where a wrong factor would not converge
    E8 = _element(cr, 8, 4, 3, 5)
    cell8 = _rect_cell(orc, E8, xa, xb, ya, yb)
    ref8 = _tensor_moments(orc, E8, 4, xa, xb, ya, yb)
    e8 = np.max(np.abs(orc.se_moments(E8, cell8, 4) - ref8)) / np.max(np.abs(ref8))
    e32 = np.max(np.abs(orc.se_moments(E, cell, 4) - _tensor_moments(orc, E, 4, xa, xb, ya, yb))) / m
    assert 4 < e8 / e32 < 32, (e8, e32)


def test_moments_sum_to_intersection_area(cr, orc):
    """Partition of unity: sum B = area(element ∩ cell), for cells that cut the
    element anywhere (lon-lat cells over a cubed-sphere element)."""
    E = _element(cr, 6, 1, 2, 2)
    ll = cr.points_to_cells(cr.lonlat_points(72, 36))
    hits = 0
    for cell in ll:
        a = orc.clip_area(E, cell)
        if a <= 0:
            continue
        B = orc.se_moments(E, cell, 4)
        assert abs(B.sum() - a) < 1e-13 * orc.polygon_area(E)
        hits += 1
    assert hits >= 9


def _cs_cells(cr, ne):
    return cr.points_to_cells(cr.cubed_sphere_points(ne))


def test_oracle_se_to_fv_invariants(cr, orc):
    """test/usecases/climacore.jl:93-111,160-186: a constant SE field regrids to
    1 in every covered cell (1e-10), the source integral sum_k colsum(A) is the
    sphere's area and the destination integral equals it (rtol 1e-12);
    test/extensions/climacore.jl:31-40: sum W = 4 pi (rtol 1e-5)."""
    E, C = _cs_cells(cr, 4), cr.points_to_cells(cr.lonlat_points(36, 18))
    S = orc.se_to_fv(C, E, 4)
    d = S.apply(np.ones(S.n_src))
    assert np.max(np.abs(d - 1.0)) < 1e-10
    src_int = np.bincount(S.col, weights=S.val, minlength=S.n_src).sum()
    assert abs(src_int - 4 * math.pi) < 1e-12 * 4 * math.pi
    assert abs(np.sum(d * S.a_dst) - src_int) < 1e-12 * src_int
    assert abs(S.a_src.sum() - 4 * math.pi) < 1e-5 * 4 * math.pi


def test_oracle_fv_to_se_invariants(cr, orc):
    """test/usecases/climacore.jl:127-140: a constant FV field projects to 1
    at every node (1e-10); the round trip SE -> FV -> SE keeps the integral
    of the latitude field (rtol 1e-2, :188-205)."""
    E, C = _cs_cells(cr, 4), cr.points_to_cells(cr.lonlat_points(36, 18))
    F = orc.fv_to_se(E, C, 4)
    assert np.max(np.abs(F.apply(np.ones(F.n_src)) - 1.0)) < 1e-10
    S = orc.se_to_fv(C, E, 4)
    # latitude field at the nodes
    x, _ = orc.gll(4)
    lat = np.array([[[math.asin(orc.element_map(e, x[i], x[j])[2]) for i in range(4)] for j in range(4)]
                    for e in E]).reshape(-1)
    fv = S.apply(lat)
    assert abs(np.sum(fv * S.a_dst) - np.sum(lat * S.a_src)) < 1e-2 * np.sum(np.abs(lat) * S.a_src)
    back = F.apply(fv)
    assert abs(np.sum(back * S.a_src) - np.sum(lat * S.a_src)) < 1e-2 * np.sum(np.abs(lat) * S.a_src)


def test_product_rules_match_definitions(cr, orc):
    """The library's own Dunavant table and GLL nodes (host code, no GPU)."""
    for degree in range(1, 9):
        bary, w = cr.se.tri_rule(degree)
        assert abs(w.sum() - 0.5) < 1e-12
        for a in range(degree + 1):
            for b in range(degree + 1 - a):
                exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
                assert abs(np.sum(w * bary[:, 1] ** a * bary[:, 2] ** b) - exact) < 1e-12
    with pytest.raises(cr.MethodError):
        cr.se.tri_rule(9)
    for n in range(2, 13):
        x, w = cr.se.gll(n)
        ox, ow = orc.gll(n)
        np.testing.assert_allclose(x, ox, atol=1e-15)
        np.testing.assert_allclose(w, ow, rtol=1e-14)
