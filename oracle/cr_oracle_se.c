/*
 * cr_oracle_se.c -- CPU FP64 restatement of the spectral-element weights of
 * ConservativeRegridding.jl's ClimaCore extension.
 *
 * TEST INFRASTRUCTURE ONLY (see cr_oracle.c): the checker for libcrhip.so's
 * spectral-element kernels; the product never links or calls it.
 *
 * What it restates (reference files relative to the reference repo root):
 *   reference_rule(degree)            src/regridder/triangle_quadrature.jl:10-156
 *                                     (Dunavant 1985 symmetric rules, barycentric
 *                                      points, weights summing to 1/2; a missing
 *                                      degree takes the next one up)
 *   accumulate_principled_b           ext/ConservativeRegriddingClimaCoreExt.jl:248-298
 *                                     (clip element by cell, fan-triangulate the
 *                                      intersection from its normalised vertex mean,
 *                                      after merging repeated vertices: a clip repeats
 *                                      a vertex where a cell corner lies on a clip
 *                                      edge, as many times as rounding decides,
 *                                      per triangle: area A_t (skip if <= 0), points
 *                                      normalize(l1 c + l2 v1 + l3 v2), weight 2 A_t w,
 *                                      B[i,j] += w phi_i(xi) phi_j(eta))
 *   inverse_element_map               ext/ConservativeRegriddingClimaCoreExt.jl:188-202
 *   element_jacobian_at               ext/ConservativeRegriddingClimaCoreExt.jl:218-233
 *
 * Libraries the reference calls that are not vendored, restated from their
 * published definitions (parity to those packages unpinned):
 *   - ClimaCore's NormalizedBilinearMap: x(xi, eta) = p / |p| with p the bilinear
 *     interpolation of the element's four corners c1 (-1,-1), c2 (1,-1),
 *     c3 (1,1), c4 (-1,1); its inverse is solved here in closed form (the two
 *     tangent-plane components of p x x = 0 are bilinear in (xi, eta); eliminating
 *     eta leaves a quadratic in xi);
 *   - Gauss-Lobatto-Legendre nodes (-1, the roots of P'_{n-1}, 1) and weights
 *     2 / (n (n-1) P_{n-1}(x)^2);
 *   - Lagrange basis on those nodes (product formula).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { double x, y, z; } sv3;

static sv3 s_add(sv3 a, sv3 b) { sv3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static sv3 s_sub(sv3 a, sv3 b) { sv3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static sv3 s_scale(sv3 a, double s) { sv3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static double s_dot(sv3 a, sv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static sv3 s_cross(sv3 a, sv3 b) {
    sv3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static sv3 s_unit(sv3 a) { return s_scale(a, 1.0 / sqrt(s_dot(a, a))); }
static sv3 s_load(const double *p) { sv3 r = {p[0], p[1], p[2]}; return r; }

/* ------------------------------------------------------ Dunavant rules -- */
/* Orbit generators of Dunavant's symmetric rules (published table, Dunavant
 * 1985): an orbit is the centroid (1 point), (a, b, b) with its 3 rotations,
 * or (a, b, c) with its 6 permutations, in the point order of the reference's
 * table.  Weights are for a triangle of area 1/2. */
static int emit3(double a, double b, double w, double *bary, double *wt, int k) {
    const double p[3][3] = {{a, b, b}, {b, a, b}, {b, b, a}};
    for (int r = 0; r < 3; ++r) {
        for (int d = 0; d < 3; ++d) bary[3 * (k + r) + d] = p[r][d];
        wt[k + r] = w;
    }
    return k + 3;
}

/* an (a, b, c) orbit: its six permutations in the table's order */
static int emit6(const double p[6][3], double w, double *bary, double *wt, int k) {
    for (int r = 0; r < 6; ++r) {
        for (int d = 0; d < 3; ++d) bary[3 * (k + r) + d] = p[r][d];
        wt[k + r] = w;
    }
    return k + 6;
}

static int rule_points(int degree, double *bary, double *wt) {
    int k = 0;
    const double third = 1.0 / 3.0;
    switch (degree) {
    case 1:
        bary[0] = bary[1] = bary[2] = third; wt[0] = 0.5; return 1;
    case 2:
        return emit3(2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, bary, wt, 0);
    case 3:
        bary[0] = bary[1] = bary[2] = third; wt[0] = -27.0 / 96.0;
        return emit3(3.0 / 5.0, 1.0 / 5.0, 25.0 / 96.0, bary, wt, 1);
    case 4: {
        const double a = 0.445948490915965, b = 0.091576213509771;
        k = emit3(1 - 2 * a, a, 0.111690794839005, bary, wt, 0);
        return emit3(1 - 2 * b, b, 0.054975871827661, bary, wt, k);
    }
    case 5: {
        const double s15 = sqrt(15.0);
        const double a = (6 + s15) / 21, b = (6 - s15) / 21;
        bary[0] = bary[1] = bary[2] = third; wt[0] = 9.0 / 80.0;
        k = emit3(1 - 2 * a, a, (155 + s15) / 2400, bary, wt, 1);
        return emit3(1 - 2 * b, b, (155 - s15) / 2400, bary, wt, k);
    }
    case 6: {
        k = emit3(0.873821971016996, 0.063089014491502, 0.025422453185103, bary, wt, 0);
        k = emit3(0.501426509658179, 0.249286745170910, 0.058393137863189, bary, wt, k);
        const double a = 0.636502499121399, b = 0.310352451033785, c = 0.053145049844816;
        const double p[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
        return emit6(p, 0.041425537809187, bary, wt, k);
    }
    case 7: {
        bary[0] = bary[1] = bary[2] = third; wt[0] = -0.074785022233835;
        k = emit3(0.479308067841920, 0.260345966079040, 0.087807628716602, bary, wt, 1);
        k = emit3(0.869739794195568, 0.065130102902216, 0.026673617804419, bary, wt, k);
        const double a = 0.048690315425316, b = 0.312865496004875, c = 0.638444188569809;
        const double p[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
        return emit6(p, 0.038556880445129, bary, wt, k);
    }
    case 8: {
        bary[0] = bary[1] = bary[2] = third; wt[0] = 0.072157803838894;
        k = emit3(0.081414823414554, 0.459292588292723, 0.047545817133643, bary, wt, 1);
        k = emit3(0.658861384496480, 0.170569307751760, 0.051608685267359, bary, wt, k);
        k = emit3(0.898905543365938, 0.050547228317031, 0.016229248811599, bary, wt, k);
        const double a = 0.008394777409958, b = 0.263112829634638, c = 0.728492392955404;
        const double p[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
        return emit6(p, 0.013615157087218, bary, wt, k);
    }
    default:
        return -1;
    }
}

/* reference_rule(degree): the rule of that degree or the next one up; -1 if
 * none (max 8).  bary holds 3*npts values, w npts (at most 16 points). */
int32_t orc_dunavant(int32_t degree, double *bary, double *w) {
    if (degree < 1) degree = 1;
    for (int d = degree; d <= 8; ++d) {
        int n = rule_points(d, bary, w);
        if (n > 0) return n;
    }
    return -1;
}

/* ----------------------------------------------------- GLL and Lagrange -- */
/* Legendre P_{n}(x) and P_{n-1}(x) by the three-term recurrence. */
static void legendre(int n, double x, double *pn, double *pn1) {
    double p0 = 1.0, p1 = x;
    if (n == 0) { *pn = 1.0; *pn1 = 0.0; return; }
    for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    *pn = p1;
    *pn1 = p0;
}

/* n >= 2 Gauss-Lobatto-Legendre nodes ascending on [-1, 1] and weights. */
void orc_gll(int32_t n, double *x, double *w) {
    const int N = n - 1;
    for (int i = 0; i < n; ++i) {
        /* Chebyshev-Gauss-Lobatto start, Newton on (1 - x^2) P'_N(x) */
        double t = -cos(M_PI * i / N);
        if (i > 0 && i < N) {
            for (int it = 0; it < 100; ++it) {
                double pn, pn1;
                legendre(N, t, &pn, &pn1);
                /* P'_N = N (P_{N-1} - x P_N) / (1 - x^2); f = P'_N,
                 * f' = (2 x P'_N - N (N+1) P_N) / (1 - x^2) */
                double dp = N * (pn1 - t * pn) / (1.0 - t * t);
                double d2 = (2.0 * t * dp - N * (N + 1.0) * pn) / (1.0 - t * t);
                double step = dp / d2;
                t -= step;
                if (fabs(step) < 1e-16) break;
            }
        }
        x[i] = t;
        double pn, pn1;
        legendre(N, t, &pn, &pn1);
        w[i] = 2.0 / (N * (N + 1.0) * pn * pn);
    }
    /* exact symmetry */
    for (int i = 0; i < n / 2; ++i) {
        double a = 0.5 * (x[n - 1 - i] - x[i]);
        x[i] = -a;
        x[n - 1 - i] = a;
        double b = 0.5 * (w[i] + w[n - 1 - i]);
        w[i] = w[n - 1 - i] = b;
    }
    if (n % 2) x[n / 2] = 0.0;
}

/* Lagrange basis phi_i(t) on nodes x[0..n-1] (product formula). */
void orc_lagrange(int32_t n, const double *x, double t, double *phi) {
    for (int i = 0; i < n; ++i) {
        double v = 1.0;
        for (int k = 0; k < n; ++k)
            if (k != i) v *= (t - x[k]) / (x[i] - x[k]);
        phi[i] = v;
    }
}

/* ------------------------------------------------ element maps -- */
/* p(xi, eta) = a + b xi + c eta + d xi eta for corners c1..c4. */
static void bilinear_coeffs(const double *cn, sv3 *a, sv3 *b, sv3 *c, sv3 *d) {
    sv3 c1 = s_load(cn), c2 = s_load(cn + 3), c3 = s_load(cn + 6), c4 = s_load(cn + 9);
    *a = s_scale(s_add(s_add(c1, c2), s_add(c3, c4)), 0.25);
    *b = s_scale(s_sub(s_add(c2, c3), s_add(c1, c4)), 0.25);
    *c = s_scale(s_sub(s_add(c3, c4), s_add(c1, c2)), 0.25);
    *d = s_scale(s_sub(s_add(c1, c3), s_add(c2, c4)), 0.25);
}

void orc_element_map(const double *corners, double xi, double eta, double *xyz) {
    sv3 a, b, c, d;
    bilinear_coeffs(corners, &a, &b, &c, &d);
    sv3 p = s_add(s_add(a, s_scale(b, xi)), s_add(s_scale(c, eta), s_scale(d, xi * eta)));
    p = s_unit(p);
    xyz[0] = p.x; xyz[1] = p.y; xyz[2] = p.z;
}

/* |dx/dxi x dx/deta| of x = p / |p| (unit sphere). */
double orc_element_jacobian(const double *corners, double xi, double eta) {
    sv3 a, b, c, d;
    bilinear_coeffs(corners, &a, &b, &c, &d);
    sv3 p = s_add(s_add(a, s_scale(b, xi)), s_add(s_scale(c, eta), s_scale(d, xi * eta)));
    double np = sqrt(s_dot(p, p));
    sv3 x = s_scale(p, 1.0 / np);
    sv3 pxi = s_add(b, s_scale(d, eta)), peta = s_add(c, s_scale(d, xi));
    sv3 txi = s_scale(s_sub(pxi, s_scale(x, s_dot(x, pxi))), 1.0 / np);
    sv3 teta = s_scale(s_sub(peta, s_scale(x, s_dot(x, peta))), 1.0 / np);
    sv3 n = s_cross(txi, teta);
    return sqrt(s_dot(n, n));
}

/* (xi, eta) with p(xi, eta) parallel to x, p . x > 0.  Closed form: with t1, t2
 * spanning the plane orthogonal to x, t_k . p = 0 gives f_k + g_k xi + h_k eta
 * + q_k xi eta = 0; eliminating eta leaves A xi^2 + B xi + C = 0.  Returns 0 on
 * success, -1 if no root lies in the (slightly widened) reference square. */
int32_t orc_inverse_element_map(const double *corners, const double *xyz, double *xi_out, double *eta_out) {
    sv3 a, b, c, d;
    bilinear_coeffs(corners, &a, &b, &c, &d);
    sv3 x = s_unit(s_load(xyz));
    sv3 e = fabs(x.x) < 0.6 ? (sv3){1, 0, 0} : (fabs(x.y) < 0.6 ? (sv3){0, 1, 0} : (sv3){0, 0, 1});
    sv3 t1 = s_unit(s_cross(x, e)), t2 = s_cross(x, t1);
    double f1 = s_dot(t1, a), g1 = s_dot(t1, b), h1 = s_dot(t1, c), q1 = s_dot(t1, d);
    double f2 = s_dot(t2, a), g2 = s_dot(t2, b), h2 = s_dot(t2, c), q2 = s_dot(t2, d);
    double A = g2 * q1 - q2 * g1;
    double B = f2 * q1 + g2 * h1 - h2 * g1 - q2 * f1;
    double C = f2 * h1 - h2 * f1;
    double roots[2];
    int nr = 0;
    double scale = fabs(A) + fabs(B) + fabs(C);
    if (fabs(A) <= 1e-14 * scale) {
        roots[nr++] = -C / B;
    } else {
        double disc = B * B - 4 * A * C;
        if (disc < 0) disc = 0;
        double sq = sqrt(disc);
        double qq = -0.5 * (B + (B >= 0 ? sq : -sq));
        roots[nr++] = qq / A;
        if (qq != 0.0) roots[nr++] = C / qq;
    }
    const double tol = 1e-6;
    for (int r = 0; r < nr; ++r) {
        double xi = roots[r];
        if (!(xi >= -1 - tol && xi <= 1 + tol)) continue;
        double den1 = h1 + q1 * xi, den2 = h2 + q2 * xi;
        double eta = fabs(den1) >= fabs(den2) ? -(f1 + g1 * xi) / den1 : -(f2 + g2 * xi) / den2;
        if (!(eta >= -1 - tol && eta <= 1 + tol)) continue;
        sv3 p = s_add(s_add(a, s_scale(b, xi)), s_add(s_scale(c, eta), s_scale(d, xi * eta)));
        if (s_dot(p, x) <= 0) continue;
        *xi_out = xi;
        *eta_out = eta;
        return 0;
    }
    return -1;
}

/* ------------------------------------------------------------ moments -- */
/* defined in cr_oracle.c */
int32_t orc_clip_polygon(const double *subj, int32_t ns, const double *clip, int32_t nc, double *out,
                         int32_t maxout);

static double tri_area(sv3 a, sv3 b, sv3 c) {
    double num = s_dot(a, s_cross(s_sub(b, a), s_sub(c, a)));
    double den = 1.0 + s_dot(a, b) + s_dot(b, c) + s_dot(c, a);
    return 2.0 * atan2(num, den);
}

/* B[i + j*nq] = integral over (element ∩ cell) of phi_i(xi) phi_j(eta) on the
 * unit sphere.  The element (4 corners, map order) is the subject clipped by the
 * convex cell (nc vertices).  Returns the intersection's vertex count, or -1. */
int32_t orc_se_moments(const double *elem, const double *cell, int32_t nc, int32_t nq, int32_t degree,
                       const double *nodes, double *B) {
    double poly[3 * 64];
    double bary[48], w[16], phx[16], phy[16];
    memset(B, 0, sizeof(double) * nq * nq);
    int m = orc_clip_polygon(elem, 4, cell, nc, poly, 64);
    if (m > 1) { /* merge cyclically consecutive vertices closer than 1e-13 (chord) */
        int wv = 1;
        for (int v = 1; v < m; ++v) {
            sv3 d = s_sub(s_load(poly + 3 * v), s_load(poly + 3 * (wv - 1)));
            if (sqrt(s_dot(d, d)) >= 1e-13) { memmove(poly + 3 * wv, poly + 3 * v, 3 * sizeof(double)); ++wv; }
        }
        while (wv > 1) {
            sv3 d = s_sub(s_load(poly + 3 * (wv - 1)), s_load(poly));
            if (sqrt(s_dot(d, d)) >= 1e-13) break;
            --wv;
        }
        m = wv;
    }
    if (m < 3) return m;
    int npts = orc_dunavant(degree, bary, w);
    if (npts < 0) return -1;
    sv3 cen = {0, 0, 0};
    for (int k = 0; k < m; ++k) cen = s_add(cen, s_load(poly + 3 * k));
    cen = s_unit(s_scale(cen, 1.0 / m));
    for (int k = 0; k < m; ++k) {
        sv3 v1 = s_load(poly + 3 * k), v2 = s_load(poly + 3 * ((k + 1) % m));
        double At = tri_area(cen, v1, v2);
        if (At <= 0) continue;
        for (int q = 0; q < npts; ++q) {
            sv3 x = s_unit(s_add(s_add(s_scale(cen, bary[3 * q]), s_scale(v1, bary[3 * q + 1])),
                                 s_scale(v2, bary[3 * q + 2])));
            double xs[3] = {x.x, x.y, x.z}, xi, eta;
            if (orc_inverse_element_map(elem, xs, &xi, &eta) != 0) return -1;
            orc_lagrange(nq, nodes, xi, phx);
            orc_lagrange(nq, nodes, eta, phy);
            double wa = w[q] * 2.0 * At;
            for (int j = 0; j < nq; ++j)
                for (int i = 0; i < nq; ++i) B[i + j * nq] += wa * phx[i] * phy[j];
        }
    }
    return m;
}
