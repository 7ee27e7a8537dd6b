// cr_oracle_ad.cpp -- forward-mode dual-number instantiation of the oracle's
// clip and area (cr_oracle.c), the CPU stand-in for differentiating the
// reference with ForwardDiff.
//
// TEST INFRASTRUCTURE ONLY (see cr_oracle.c): only tests/ load it.
//
// The differentiated function is area(normalize(v_1), ..., normalize(v_n)):
// inputs are projected onto the unit sphere first, so derivatives are the
// tangential (on-sphere) ones and radial perturbations have no effect.  The
// algorithm is the oracle's buffered Sutherland-Hodgman with the same sign
// tests (decided on the real parts) and the same published formulas.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

struct Dual {
    double v, d;
};
inline Dual mk(double v, double d = 0.0) { return Dual{v, d}; }
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual dsqrt(Dual a) {
    double s = std::sqrt(a.v);
    return {s, a.d / (2.0 * s)};
}
inline Dual datan2(Dual y, Dual x) {
    double r2 = x.v * x.v + y.v * y.v;
    return {std::atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / r2};
}

struct V {
    Dual x, y, z;
};
inline V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V add(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V scl(V a, Dual s) { return {a.x * s, a.y * s, a.z * s}; }
inline Dual dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V unit(V a) { return scl(a, mk(1.0) / dsqrt(dot(a, a))); }

constexpr int MAXV = 256;
constexpr double EDGE_EPS = 1e-13;

Dual tri_excess(V a, V b, V c) {
    Dual num = dot(a, cross(sub(b, a), sub(c, a)));
    Dual den = mk(1.0) + dot(a, b) + dot(b, c) + dot(c, a);
    return mk(2.0) * datan2(num, den);
}

Dual ring_area(const V *p, int n) {
    Dual s = mk(0.0);
    for (int i = 1; i + 1 < n; ++i) s = s + tri_excess(p[0], p[i], p[i + 1]);
    return s;
}

int sh_clip(const V *S, int ns, const V *C, int nc, V *out, V *tmp) {
    V *cur = tmp, *nxt = out;
    int m = ns;
    for (int i = 0; i < ns; ++i) cur[i] = S[i];
    for (int k = 0; k < nc && m > 0; ++k) {
        V a = C[k], b = C[(k + 1) % nc];
        V n = cross(a, sub(b, a));
        if (std::sqrt(dot(n, n).v) < EDGE_EPS) continue;
        int w = 0;
        V p = cur[m - 1];
        Dual sp = dot(n, sub(p, a));
        for (int i = 0; i < m; ++i) {
            V q = cur[i];
            Dual sq = dot(n, sub(q, a));
            if (sq.v >= 0.0) {
                if (sp.v < 0.0) nxt[w++] = unit(add(p, scl(sub(q, p), sp / (sp - sq))));
                nxt[w++] = q;
            } else if (sp.v >= 0.0) {
                nxt[w++] = unit(add(p, scl(sub(q, p), sp / (sp - sq))));
            }
            if (w >= MAXV - 2) return -1;
            p = q;
            sp = sq;
        }
        m = w;
        V *t = cur;
        cur = nxt;
        nxt = t;
    }
    if (cur != out)
        for (int i = 0; i < m; ++i) out[i] = cur[i];
    return m;
}

void load(const double *xyz, int n, V *out, int seed) {
    for (int i = 0; i < n; ++i) {
        V v = {mk(xyz[3 * i], seed == 3 * i), mk(xyz[3 * i + 1], seed == 3 * i + 1),
               mk(xyz[3 * i + 2], seed == 3 * i + 2)};
        out[i] = unit(v);
    }
}

}  // namespace

extern "C" {

// d(area)/d(vertex coordinates) of an open ring; returns the area.
double orc_polygon_area_grad(const double *xyz, int32_t n, double *grad) {
    V p[MAXV];
    double area = 0.0;
    if (n > MAXV) return NAN;
    for (int seed = 0; seed < 3 * n; ++seed) {
        load(xyz, n, p, seed);
        Dual a = ring_area(p, n);
        grad[seed] = a.d;
        area = a.v;
    }
    return area;
}

// Area of subject ∩ clip and its derivatives w.r.t. every coordinate of both
// polygons (one forward-mode pass per coordinate).
double orc_clip_area_grad(const double *subj, int32_t ns, const double *clip, int32_t nc,
                          double *grad_subj, double *grad_clip) {
    V S[MAXV], C[MAXV], out[MAXV], tmp[MAXV];
    if (ns > MAXV / 2 || nc > MAXV / 2) return NAN;
    double area = 0.0;
    for (int seed = 0; seed < 3 * (ns + nc); ++seed) {
        load(subj, ns, S, seed);
        load(clip, nc, C, seed - 3 * ns);
        int m = sh_clip(S, ns, C, nc, out, tmp);
        if (m < 0) return NAN;
        Dual a = m >= 3 ? ring_area(out, m) : mk(0.0);
        if (seed < 3 * ns) grad_subj[seed] = a.d;
        else grad_clip[seed - 3 * ns] = a.d;
        area = a.v;
    }
    return area;
}

}  // extern "C"
