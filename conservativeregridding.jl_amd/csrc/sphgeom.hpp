// sphgeom.hpp -- FP64 spherical polygon geometry for the gfx950 kernels.
//
// Replaces the per-pair operator of the reference,
//   DefaultIntersectionOperator{Spherical}(p1, p2) =
//       GO.area(m, GO.intersection(ConvexConvexSutherlandHodgman(m), p1, p2))
//   (src/regridder/regridder.jl:96-103),
// and the cell areas of areas() (regridder.jl:165-178).
//
// Design for registers, not buffers: Sutherland-Hodgman is run as the
// re-entrant pipeline of its original paper -- one stage per clip edge, each
// stage holding only (first, previous) vertex and side value.  Subject vertices
// are streamed from memory one at a time and every vertex a stage emits is
// pushed straight into the next stage; the last stage feeds the area
// accumulator.  State is O(clip edges) doubles no matter how many vertices the
// subject polygon has (16 for a HEALPix step-4 cell), so the kernel needs no
// per-thread scratch arrays.
//
// Numerics (shared with the CPU oracle as published formulas, not as code):
//   edge a->b: great circle with normal n = a x (b - a); side(v) = n . (v - a),
//   inside iff side >= 0.  Crossing of edge p->q: the positive combination
//   |s_q| p + |s_p| q, renormalised (it lies on the plane for opposite signs).
//   Area: triangle solid angle tan(E/2) = a.((b-a)x(c-a)) / (1 + a.b + b.c + c.a)
//   (Van Oosterom & Strackee 1983), fan from the first vertex.  The fan's angles
//   are summed as the argument of a product of complex numbers, so each polygon
//   needs one atan2 instead of one per triangle.
#pragma once

#include "common.hpp"

namespace cr {

struct d3 {
    double x, y, z;
};

CR_HD d3 mk3(double x, double y, double z) { return d3{x, y, z}; }
CR_HD d3 operator-(const d3 &a, const d3 &b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CR_HD d3 operator+(const d3 &a, const d3 &b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CR_HD d3 operator*(double s, const d3 &a) { return d3{s * a.x, s * a.y, s * a.z}; }
CR_HD double dot(const d3 &a, const d3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CR_HD d3 cross(const d3 &a, const d3 &b) {
    return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
CR_HD d3 load3(const double *p) { return d3{p[0], p[1], p[2]}; }
CR_HD d3 normalized(const d3 &a) { return (1.0 / sqrt(dot(a, a))) * a; }

// Squared length below which a clip edge is treated as degenerate (coincident
// vertices, e.g. the two pole corners of a lon-lat polar cell).
constexpr double kDegenerateEdge2 = 1e-26;

// Fan-area accumulator: angle sum as arg of a running complex product.
struct AreaAcc {
    d3 w0, prev;
    double re, im;
    int n;
    CR_HD void init() { re = 1.0; im = 0.0; n = 0; }
    CR_HD void push(const d3 &w) {
        if (n >= 2) {
            d3 e1 = prev - w0, e2 = w - w0;
            double num = dot(w0, cross(e1, e2));
            double den = 1.0 + dot(w0, prev) + dot(prev, w) + dot(w, w0);
            // scale by 1/4 (exact) to keep the product near unit magnitude
            num *= 0.25;
            den *= 0.25;
            double r = re * den - im * num;
            double i = re * num + im * den;
            re = r;
            im = i;
        } else if (n == 0) {
            w0 = w;
        }
        prev = w;
        ++n;
    }
    CR_HD double area() const { return n >= 3 ? 2.0 * atan2(im, re) : 0.0; }
};

// One clip stage per clip-polygon edge, unrolled at compile time (NC stages).
template <int NC>
struct ClipPipe {
    d3 anchor[NC];
    d3 nrm[NC];
    d3 first[NC];
    d3 prev[NC];
    double sfirst[NC];
    double sprev[NC];
    bool started[NC];
    AreaAcc acc;

    // Stage k clips against the great circle of edge (a -> b).  A degenerate
    // edge gets a zero normal: every vertex is then "inside" (side 0) and the
    // stage passes the polygon through unchanged.
    CR_HD void set_edge(int k, const d3 &a, const d3 &b) {
        d3 n = cross(a, b - a);
        if (dot(n, n) < kDegenerateEdge2) n = mk3(0.0, 0.0, 0.0);
        anchor[k] = a;
        nrm[k] = n;
        started[k] = false;
    }
    CR_HD void init() { acc.init(); }
};

// 1/sqrt(v): the device's rsqrt (one instruction sequence, no IEEE division);
// the host build divides.  Either is within an ulp or two, which only moves a
// renormalised point along its ray.
CR_HD double inv_norm(double v) {
#ifdef __HIP_DEVICE_COMPILE__
    return rsqrt(v);
#else
    return 1.0 / sqrt(v);
#endif
}

CR_HD d3 crossing(const d3 &p, double sp, const d3 &q, double sq) {
    d3 x = fabs(sq) * p + fabs(sp) * q;
    return inv_norm(dot(x, x)) * x;
}

template <int K, int NC>
CR_HD void pipe_push(ClipPipe<NC> &P, const d3 &v);

// Sutherland-Hodgman rule for edge p->q at stage K: a crossing when the side
// changes, then q when it is inside.  The (at most two) outputs are forwarded
// by a loop that is not unrolled, so each stage's downstream code exists once
// and the kernel grows linearly, not exponentially, with the number of stages.
template <int K, int NC>
CR_HD void pipe_edge(ClipPipe<NC> &P, const d3 &p, double sp, const d3 &q, double sq) {
    const bool qin = sq >= 0.0;
    const bool crosses = (sp >= 0.0) != qin;
    d3 x = q;
    if (crosses) x = crossing(p, sp, q, sq);
    const int m = (int)crosses + (int)qin;
#pragma unroll 1
    for (int e = 0; e < m; ++e) {
        d3 u = (e == 0) ? x : q;
        pipe_push<K + 1, NC>(P, u);
    }
}

template <int K, int NC>
CR_HD void pipe_push(ClipPipe<NC> &P, const d3 &v) {
    if constexpr (K == NC) {
        P.acc.push(v);
    } else {
        double s = dot(P.nrm[K], v - P.anchor[K]);
        if (!P.started[K]) {
            P.started[K] = true;
            P.first[K] = v;
            P.sfirst[K] = s;
        } else {
            pipe_edge<K, NC>(P, P.prev[K], P.sprev[K], v, s);
        }
        P.prev[K] = v;
        P.sprev[K] = s;
    }
}

// Close every stage in order: the closing edge (last -> first) of stage K may
// emit vertices into stage K+1 before that stage is closed.
template <int K, int NC>
CR_HD void pipe_close(ClipPipe<NC> &P) {
    if constexpr (K < NC) {
        if (P.started[K]) pipe_edge<K, NC>(P, P.prev[K], P.sprev[K], P.first[K], P.sfirst[K]);
        pipe_close<K + 1, NC>(P);
    }
}

// Area of (subject ∩ clip) for a clip polygon of at most NC vertices (padded by
// repeating its last vertex) and a subject of ns vertices read from memory.
template <int NC>
CR_HD double clip_area_stream(const double *clip, int nc, const double *subj, int ns) {
    ClipPipe<NC> P;
    P.init();
    d3 c[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) c[k] = load3(clip + 3 * (k < nc ? k : nc - 1));
#pragma unroll
    for (int k = 0; k < NC; ++k) P.set_edge(k, c[k], c[(k + 1) % NC]);
    // subject vertices are loaded two ahead of the one being clipped, so their
    // latency overlaps the pipeline's work
    d3 v0 = load3(subj), v1 = load3(subj + 3 * (ns > 1 ? 1 : 0));
    for (int i = 0; i < ns; ++i) {
        const d3 v2 = load3(subj + 3 * (i + 2 < ns ? i + 2 : ns - 1));
        pipe_push<0, NC>(P, v0);
        v0 = v1;
        v1 = v2;
    }
    pipe_close<0, NC>(P);
    return P.acc.area();
}

// The same with a pipeline of NS stages, one per clip edge whose bit is set in
// `mask` (exactly NS bits, in edge order).  An edge that has every subject
// vertex on its inside passes the clipped polygon through unchanged (the
// polygon stays inside the subject, which is inside that edge's hemisphere),
// so dropping its stage changes nothing but rounding.
template <int K, int NS, int NC>
CR_HD void pipe_set_masked(ClipPipe<NS> &P, const double *clip, int nc, unsigned mask) {
    if constexpr (K < NS) {
        const int e = __builtin_ctz(mask), e1 = (e + 1) % NC;
        P.set_edge(K, load3(clip + 3 * (e < nc ? e : nc - 1)), load3(clip + 3 * (e1 < nc ? e1 : nc - 1)));
        pipe_set_masked<K + 1, NS, NC>(P, clip, nc, mask & (mask - 1));
    }
}

template <int NS, int NC>
CR_HD double clip_area_stream_masked(const double *clip, int nc, unsigned mask, const double *subj, int ns) {
    ClipPipe<NS> P;
    P.init();
    pipe_set_masked<0, NS, NC>(P, clip, nc, mask);
    d3 v0 = load3(subj), v1 = load3(subj + 3 * (ns > 1 ? 1 : 0));
    for (int i = 0; i < ns; ++i) {
        const d3 v2 = load3(subj + 3 * (i + 2 < ns ? i + 2 : ns - 1));
        pipe_push<0, NS>(P, v0);
        v0 = v1;
        v1 = v2;
    }
    pipe_close<0, NS>(P);
    return P.acc.area();
}

// `mask` with its lowest clear bits set until it has n bits (n <= NC).
CR_HD unsigned pad_mask(unsigned mask, int n) {
    while (__builtin_popcount(mask) < n) mask |= ~mask & (mask + 1u);
    return mask;
}

// Where the subject lies against the clip polygon, by the pipeline's own side
// test n.(v - a) (a degenerate edge has n = 0: everything is on its inside):
// kSubjectOutside when every subject vertex is strictly outside one clip edge
// (the intersection is empty), kSubjectInside when every vertex is inside
// every edge (the pipeline would pass the subject through unchanged: the
// intersection is the subject), kSubjectCrosses otherwise.
enum { kSubjectOutside = 0, kSubjectInside = 1, kSubjectCrosses = 2 };

template <int NC>
CR_HD int classify_subject(const double *clip, int nc, const double *subj, int ns, unsigned *crossed = nullptr) {
    d3 a[NC], n[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) a[k] = load3(clip + 3 * (k < nc ? k : nc - 1));
    bool out[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        d3 nk = cross(a[k], a[(k + 1) % NC] - a[k]);
        if (dot(nk, nk) < kDegenerateEdge2) nk = mk3(0.0, 0.0, 0.0);
        out[k] = true;
        n[k] = nk;
    }
    bool in = true;
    unsigned cut = 0;  // edges with a subject vertex outside
    for (int i = 0; i < ns; ++i) {
        const d3 v = load3(subj + 3 * i);
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const double s = dot(n[k], v - a[k]);
            out[k] = out[k] && s < 0.0;
            cut |= (s < 0.0 ? 1u : 0u) << k;
        }
    }
    in = cut == 0;
    if (crossed) *crossed = cut;
    bool any = false;
#pragma unroll
    for (int k = 0; k < NC; ++k) any = any || out[k];
    return any ? kSubjectOutside : (in ? kSubjectInside : kSubjectCrosses);
}

// Signed area of an open ring of n vertices.
CR_HD double ring_area(const double *xyz, int n) {
    AreaAcc a;
    a.init();
    for (int i = 0; i < n; ++i) a.push(load3(xyz + 3 * i));
    return a.area();
}


// ============================================================================
// Backward: d(area of subject ∩ clip) / d(vertices), by provenance.
//
// Every vertex of the clipped polygon is, in exact arithmetic, a subject
// vertex, a clip vertex, or the crossing of two great circles (a subject edge
// with a clip edge, or two clip edges).  The pipeline below is the forward one
// with a label carried next to each vertex; a crossing's label comes from the
// line the clipped segment lies on (the line its two endpoints share).  The
// area stage differentiates each fan triangle,
//   E = 2 atan2(T, D),  T = a.(b x c),  D = 1 + a.b + b.c + c.a,
//   dE/da = 2 (D (b x c) - T (b + c)) / (T^2 + D^2)   (and cyclic),
// and hands (label, position, gradient) to a sink, which chains through the
// crossing x = y/|y|, y = +-(n1 x n2), n = p x q, to the input vertices.
// ============================================================================

// label: bits 28-29 type, bits 8-27 and 0-7 indices (a subject vertex uses
// bits 0-27); -1 (sign bit) marks a vertex without provenance
constexpr int32_t kLabS = 0, kLabC = 1, kLabXF = 2, kLabXE = 3;
constexpr int32_t kLabIndexMask = 0x0fffffff;
CR_HD int32_t lab_make(int32_t type, int32_t a, int32_t b) { return (type << 28) | (a << 8) | b; }
CR_HD int32_t lab_type(int32_t l) { return (l >> 28) & 3; }
CR_HD int32_t lab_hi(int32_t l) { return (l >> 8) & 0xfffff; }
CR_HD int32_t lab_lo(int32_t l) { return l & 0xff; }

// line ids: subject edge e -> e, clip edge k -> kClipLine | k
constexpr int32_t kClipLine = 1 << 27;

template <int NC>
CR_HD void lab_lines(int32_t l, int ns, int32_t &l1, int32_t &l2) {
    int32_t t = lab_type(l), a = lab_hi(l), b = lab_lo(l);
    if (t == kLabS) {
        int32_t e = l & kLabIndexMask;
        l1 = e == 0 ? ns - 1 : e - 1;
        l2 = e;
    }
    else if (t == kLabC) { l1 = kClipLine | ((b + NC - 1) % NC); l2 = kClipLine | b; }
    else if (t == kLabXF) { l1 = a; l2 = kClipLine | b; }
    else { l1 = kClipLine | a; l2 = kClipLine | b; }
}

// Label of the crossing of segment p->q with clip edge K.
template <int NC>
CR_HD int32_t lab_crossing(int32_t lp, int32_t lq, int K, int ns) {
    int32_t p1, p2, q1, q2;
    lab_lines<NC>(lp, ns, p1, p2);
    lab_lines<NC>(lq, ns, q1, q2);
    int32_t line = (p1 == q1 || p1 == q2) ? p1 : ((p2 == q1 || p2 == q2) ? p2 : -1);
    if (line < 0) return -1;  // no shared line (degenerate input): no gradient
    if (!(line & kClipLine)) return lab_make(kLabXF, line, K);
    int a = line & 0xff;
    if (a == (K + NC - 1) % NC) return lab_make(kLabC, 0, K);
    if (a == (K + 1) % NC) return lab_make(kLabC, 0, (K + 1) % NC);
    return lab_make(kLabXE, a, K);
}

CR_HD d3 tangential(const d3 &g, const d3 &x) { return g - dot(g, x) * x; }

template <int NC, class Sink>
struct GradPipe {
    d3 anchor[NC];
    d3 *nrm;  // the NC edge normals, in caller memory (LDS) to spare registers
    d3 first[NC], prev[NC];
    double sfirst[NC], sprev[NC];
    int32_t lfirst[NC], lprev[NC];
    bool started[NC];
    int ns;
    double coef;
    Sink *sink;
    // area stage: fan triangles (w0, wp, w); the previous vertex's part from
    // the last triangle waits in pend until the next triangle completes it,
    // so each vertex reaches the sink once (the sink is linear in g)
    d3 w0, wp, g0, pend;
    int32_t l0, lp;
    int n;

    CR_HD void area_push(const d3 &w, int32_t l) {
        if (n == 0) {
            w0 = w;
            l0 = l;
            g0 = mk3(0, 0, 0);
            pend = mk3(0, 0, 0);
        } else if (n >= 2) {
            d3 a = w0, b = wp, c = w;
            double T = dot(a, cross(b, c));
            double D = 1.0 + dot(a, b) + dot(b, c) + dot(c, a);
            double f = 2.0 / (T * T + D * D);
            g0 = g0 + f * (D * cross(b, c) - T * (b + c));
            (*sink)(lp, b, pend + (coef * f) * (D * cross(c, a) - T * (a + c)));
            pend = (coef * f) * (D * cross(a, b) - T * (a + b));
        }
        wp = w;
        lp = l;
        ++n;
    }
    CR_HD void area_close() {
        if (n < 3) return;
        (*sink)(lp, wp, pend);
        (*sink)(l0, w0, coef * g0);
    }
};

template <int K, int NC, class Sink>
CR_HD void gpipe_push(GradPipe<NC, Sink> &P, const d3 &v, int32_t l);

template <int K, int NC, class Sink>
CR_HD void gpipe_edge(GradPipe<NC, Sink> &P, const d3 &p, double sp, int32_t lp, const d3 &q,
                      double sq, int32_t lq) {
    const bool qin = sq >= 0.0;
    const bool crosses = (sp >= 0.0) != qin;
    d3 x = q;
    int32_t lx = lq;
    if (crosses) {
        x = crossing(p, sp, q, sq);
        lx = lab_crossing<NC>(lp, lq, K, P.ns);
    }
    const int m = (int)crosses + (int)qin;
#pragma unroll 1
    for (int e = 0; e < m; ++e) {
        d3 u = (e == 0) ? x : q;
        int32_t lu = (e == 0) ? lx : lq;
        gpipe_push<K + 1, NC, Sink>(P, u, lu);
    }
}

template <int K, int NC, class Sink>
CR_HD void gpipe_push(GradPipe<NC, Sink> &P, const d3 &v, int32_t l) {
    if constexpr (K == NC) {
        P.area_push(v, l);
    } else {
        double s = dot(P.nrm[K], v - P.anchor[K]);
        if (!P.started[K]) {
            P.started[K] = true;
            P.first[K] = v;
            P.sfirst[K] = s;
            P.lfirst[K] = l;
        } else {
            gpipe_edge<K, NC, Sink>(P, P.prev[K], P.sprev[K], P.lprev[K], v, s, l);
        }
        P.prev[K] = v;
        P.sprev[K] = s;
        P.lprev[K] = l;
    }
}

template <int K, int NC, class Sink>
CR_HD void gpipe_close(GradPipe<NC, Sink> &P) {
    if constexpr (K < NC) {
        if (P.started[K])
            gpipe_edge<K, NC, Sink>(P, P.prev[K], P.sprev[K], P.lprev[K], P.first[K], P.sfirst[K],
                                    P.lfirst[K]);
        gpipe_close<K + 1, NC, Sink>(P);
    }
}

// Scatter coef * d(area)/d(vertex) of (subject ∩ clip) through `sink`;
// nrm_store holds NC normals (the caller's, e.g. a per-thread slice of LDS).
template <int NC, class Sink>
CR_HD void clip_area_backward(const double *clip, int nc, const double *subj, int ns, double coef,
                              Sink &sink, d3 *nrm_store) {
    GradPipe<NC, Sink> P;
    P.nrm = nrm_store;
    P.ns = ns;
    P.coef = coef;
    P.sink = &sink;
    P.n = 0;
    d3 c[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) c[k] = load3(clip + 3 * (k < nc ? k : nc - 1));
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        d3 n = cross(c[k], c[(k + 1) % NC] - c[k]);
        if (dot(n, n) < kDegenerateEdge2) n = mk3(0.0, 0.0, 0.0);
        P.anchor[k] = c[k];
        P.nrm[k] = n;
        P.started[k] = false;
    }
    for (int i = 0; i < ns; ++i) gpipe_push<0, NC, Sink>(P, load3(subj + 3 * i), lab_make(kLabS, 0, 0) | i);
    gpipe_close<0, NC, Sink>(P);
    P.area_close();
}

// crossing_backward for a crossing that lies on the arc p1 -> q1, with the
// gradient already made tangential at the crossing: the sign of y is that of
// y.(p1 + q1).
CR_HD bool crossing_backward_arc(const d3 &gt, const d3 &p1, const d3 &q1, const d3 &p2, const d3 &q2, d3 &gp1,
                                 d3 &gq1, d3 &gp2, d3 &gq2) {
    d3 n1 = cross(p1, q1), n2 = cross(p2, q2);
    d3 y = cross(n1, n2);
    double L = sqrt(dot(y, y));
    if (!(L > 1e-300)) return false;
    double sigma = dot(y, p1 + q1) >= 0.0 ? 1.0 : -1.0;
    d3 yb = (sigma / L) * gt;
    d3 n1b = cross(n2, yb), n2b = cross(yb, n1);
    gp1 = cross(q1, n1b);
    gq1 = cross(n1b, p1);
    gp2 = cross(q2, n2b);
    gq2 = cross(n2b, p2);
    return true;
}

// Gradient w.r.t. the great-circle crossing x = y/|y| (y = +-(n1 x n2),
// n1 = p1 x q1, n2 = p2 x q2) pushed back to p1, q1, p2, q2.
CR_HD bool crossing_backward(const d3 &x, const d3 &g, const d3 &p1, const d3 &q1, const d3 &p2,
                             const d3 &q2, d3 &gp1, d3 &gq1, d3 &gp2, d3 &gq2) {
    d3 n1 = cross(p1, q1), n2 = cross(p2, q2);
    d3 y = cross(n1, n2);
    double L = sqrt(dot(y, y));
    if (!(L > 1e-300)) return false;
    double sigma = dot(y, x) >= 0.0 ? 1.0 : -1.0;
    d3 yb = (sigma / L) * tangential(g, x);
    d3 n1b = cross(n2, yb), n2b = cross(yb, n1);
    gp1 = cross(q1, n1b);
    gq1 = cross(n1b, p1);
    gp2 = cross(q2, n2b);
    gq2 = cross(n2b, p2);
    return true;
}

// ============================================================================
// Backward by boundary decomposition (no clip pipeline): edges_plan, then
// edges_emit.
//
// area(P), P = subject ∩ clip, is the sum over P's boundary arcs a -> b of the
// fan triangle E(O, a, b) for any fixed point O (the area does not depend on
// it; O is taken on P's boundary so that every triangle is as small as P).
// The arcs are (i) the part of each subject edge inside the convex clip
// polygon -- one interval, from the edge's side values against every clip
// line: the latest entry and the earliest exit (Liang-Barsky on the chord
// parameter; the crossings are the forward's crossing() of the edge's own end
// points) -- and (ii) the parts of the clip edges inside the subject, which
// begin where a subject edge leaves the clip polygon through that edge's line
// or at a clip vertex inside the subject, and end where a subject edge enters
// or at the next clip vertex.  A clip edge carries at most two such points
// here.  Each arc's two end points take their derivatives: a subject vertex
// or clip vertex directly, a crossing through crossing_backward (labels as in
// clip_area_backward, g not yet tangential, coef applied).  Every loop runs
// over the ring sizes: no clip stages, no recursion.
//
// edges_plan returns false (and edges_emit then emits nothing for the pair)
// when the pair is near a configuration where the area is not differentiable
// or where the decisions above could differ from the forward clip's by
// rounding -- a subject vertex within ~1e-10 of a clip edge's length from
// that edge (on its line between the lines through its corners), a crossing
// within that distance of a clip corner or of another crossing on its clip
// edge, an edge entering and leaving through neighbouring lines at one point
// -- when a clip edge meets the subject's boundary more than twice, or when a
// clip edge is degenerate.  The caller then uses clip_area_backward.
// ============================================================================
constexpr double kEdgeTol = 1e-10;

// d E(O, a, b) / da and / db for the fan triangle O -> a -> b (O fixed).
CR_HD void fan_grad(const d3 &O, const d3 &a, const d3 &b, double coef, d3 &ga, d3 &gb) {
    const double T = dot(O, cross(a, b));
    const double D = 1.0 + dot(O, a) + dot(a, b) + dot(b, O);
    const double f = coef * 2.0 / (T * T + D * D);
    ga = f * (D * cross(b, O) - T * (O + b));
    gb = f * (D * cross(O, a) - T * (O + a));
}

// The crossings of one subject edge with the clip lines: the latest entry
// (k_in, t_in) and the earliest exit (k_out, t_out); empty when an end point
// pair lies outside one line.
// The winners' side values (and, with corners, their clip edge's end points)
// are carried out of the unrolled loop, so no array is indexed by a run-time
// line number (which would put it in scratch memory).
template <int NC>
struct EdgeCut {
    double t_in, t_out, ap_in, aq_in, ap_out, aq_out;
    // POS: the end points' side values against the lines before and after
    // the winning line (through its clip edge's first and last corner), and
    // their tolerances: at the crossing they place it along the clip edge
    double bp_in, bq_in, fp_in, fq_in, btol_in, ftol_in;
    double bp_out, bq_out, fp_out, fq_out, btol_out, ftol_out;
    int k_in, k_out;
    bool empty;
    template <bool POS>
    CR_HD void make(const double (&ap)[NC], const double (&aq)[NC], const double (&tol)[NC]) {
        t_in = 0.0;
        t_out = 1.0;
        ap_in = aq_in = ap_out = aq_out = 0.0;
        k_in = -1;
        k_out = -1;
        empty = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const bool pin = ap[k] >= 0.0, qin = aq[k] >= 0.0;
            if (pin && qin) continue;
            if (!pin && !qin) {
                empty = true;
                continue;
            }
            const double t = fabs(ap[k]) / (fabs(ap[k]) + fabs(aq[k]));
            const int b = (k + NC - 1) % NC, f = (k + 1) % NC;
            if (!pin && t > t_in) {
                t_in = t;
                k_in = k;
                ap_in = ap[k];
                aq_in = aq[k];
                if (POS) {
                    bp_in = ap[b];
                    bq_in = aq[b];
                    fp_in = ap[f];
                    fq_in = aq[f];
                    btol_in = tol[b];
                    ftol_in = tol[f];
                }
            }
            if (pin && t < t_out) {
                t_out = t;
                k_out = k;
                ap_out = ap[k];
                aq_out = aq[k];
                if (POS) {
                    bp_out = ap[b];
                    bq_out = aq[b];
                    fp_out = ap[f];
                    fq_out = aq[f];
                    btol_out = tol[b];
                    ftol_out = tol[f];
                }
            }
        }
    }
    CR_HD bool inside() const { return !empty && t_in < t_out; }
};

// The crossing of subject edge j with the clip line through ck with normal
// nk: the forward's crossing() of the edge's own end points.
CR_HD d3 edge_cut_point(const double *subj, int ns, int j, const d3 &ck, const d3 &nk) {
    const d3 p = load3(subj + 3 * j), q = load3(subj + 3 * (j + 1 == ns ? 0 : j + 1));
    return crossing(p, dot(nk, p - ck), q, dot(nk, q - ck));
}

// What pass 1 leaves for pass 2: the clip ring and its lines, the crossings
// on every clip edge (at most two, in order along the edge; kind 1 = a
// clip-edge arc begins there), and O.
template <int NC>
struct EdgesPlan {
    d3 c[NC], n[NC], O;
    int32_t lab_lo[NC], lab_hi[NC];
    int kind_lo[NC], kind_hi[NC], cnt[NC];
    bool any_out;
};

// Pass 1: the checks (false: use clip_area_backward) and the plan.
template <int NC>
CR_HD bool edges_plan_(const double *clip, int nc, const double *subj, int ns, EdgesPlan<NC> &P) {
    double tol[NC];
    // a declined plan still drives edges_emit in step with the other lanes
    // (emitting nothing): it reads no crossing labels (no counts) and finite lines
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        P.c[k] = load3(clip + 3 * (k < nc ? k : nc - 1));
        P.n[k] = mk3(0.0, 0.0, 0.0);
        P.lab_lo[k] = P.lab_hi[k] = -1;
        P.kind_lo[k] = P.kind_hi[k] = 0;
        P.cnt[k] = 0;
    }
    P.any_out = false;
    P.O = load3(subj);
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const d3 e = P.c[(k + 1) % NC] - P.c[k];
        const d3 nk = cross(P.c[k], e);
        const double nn = dot(nk, nk);
        // a degenerate clip edge (a polar cell's coincident corners, a padded
        // triangle) is no constraint in the forward clip, whose corner there is
        // the crossing of the two neighbouring lines: clip_area_backward
        // differentiates that corner; this decomposition would not
        if (nn < kDegenerateEdge2) return false;
        P.n[k] = nk;
        tol[k] = kEdgeTol * nn;  // |side| below this: within ~1e-10 edge lengths of the line
    }
    double u_lo[NC], u_hi[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) u_lo[k] = u_hi[k] = -1.0;
    bool have_O = false, overflow = false;
    // a subject vertex on a clip line is a degenerate point only on the
    // clip edge's segment (inside the lines through its two corners); off it,
    // both clips take the line's side test the same way
    auto on_edge = [&](const double (&a)[NC]) {
        bool hit = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const int b = (k + NC - 1) % NC, f = (k + 1) % NC;
            hit = hit || (fabs(a[k]) <= tol[k] && a[b] >= -tol[b] && a[f] >= -tol[f]);
        }
        return hit;
    };
    d3 p = load3(subj);
    double ap[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) ap[k] = dot(P.n[k], p - P.c[k]);
    if (on_edge(ap)) return false;
    for (int j = 0; j < ns; ++j) {
        const d3 q = load3(subj + 3 * (j + 1 == ns ? 0 : j + 1));
        double aq[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            aq[k] = dot(P.n[k], q - P.c[k]);
            P.any_out = P.any_out || aq[k] < 0.0;
        }
        if (on_edge(aq)) return false;
        EdgeCut<NC> cut;
        cut.template make<true>(ap, aq, tol);
        if (cut.k_in >= 0 && cut.k_out >= 0 && fabs(cut.t_in - cut.t_out) <= kEdgeTol) return false;
        if (cut.inside()) {
            if (!have_O) {
                P.O = cut.k_in >= 0 ? crossing(p, cut.ap_in, q, cut.aq_in) : p;
                have_O = true;
            }
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int k = side == 0 ? cut.k_in : cut.k_out;
                if (k < 0) continue;
                // the crossing's place along clip edge k: its side values (linear
                // along the chord) against the line through the edge's first
                // corner (0 there, growing along the edge) and through its last
                const double t = side == 0 ? cut.t_in : cut.t_out;
                const double u = side == 0 ? (1.0 - t) * cut.bp_in + t * cut.bq_in
                                           : (1.0 - t) * cut.bp_out + t * cut.bq_out;
                const double v = side == 0 ? (1.0 - t) * cut.fp_in + t * cut.fq_in
                                           : (1.0 - t) * cut.fp_out + t * cut.fq_out;
                if (u <= (side == 0 ? cut.btol_in : cut.btol_out) || v <= (side == 0 ? cut.ftol_in : cut.ftol_out))
                    return false;  // within ~1e-10 of a clip corner
                // entering the clip polygon through line k ends a clip-edge arc there (kind 0),
                // leaving through it begins one (kind 1)
                const int32_t lab = lab_make(kLabXF, j, k);
#pragma unroll
                for (int m = 0; m < NC; ++m) {
                    if (m != k) continue;
                    if (P.cnt[m] == 0) {
                        u_lo[m] = u;
                        P.lab_lo[m] = lab;
                        P.kind_lo[m] = side;
                    } else if (P.cnt[m] == 1) {
                        if (fabs(u - u_lo[m]) <= tol[(m + NC - 1) % NC]) overflow = true;
                        if (u > u_lo[m]) {
                            u_hi[m] = u;
                            P.lab_hi[m] = lab;
                            P.kind_hi[m] = side;
                        } else {
                            u_hi[m] = u_lo[m];
                            P.lab_hi[m] = P.lab_lo[m];
                            P.kind_hi[m] = P.kind_lo[m];
                            u_lo[m] = u;
                            P.lab_lo[m] = lab;
                            P.kind_lo[m] = side;
                        }
                    } else {
                        overflow = true;
                    }
                    P.cnt[m] += 1;
                }
            }
        }
        p = q;
#pragma unroll
        for (int k = 0; k < NC; ++k) ap[k] = aq[k];
    }
    if (overflow) return false;
#pragma unroll
    for (int k = 0; k < NC; ++k)
        if (P.cnt[k] == 2 && P.kind_lo[k] == P.kind_hi[k]) return false;  // not an arc's two ends
    if (!have_O) P.O = P.any_out ? P.c[0] : load3(subj);  // P is the clip polygon, or the subject
    return true;
}

template <int NC>
CR_HD bool edges_plan(const double *clip, int nc, const double *subj, int ns, EdgesPlan<NC> &P) {
    if (edges_plan_<NC>(clip, nc, subj, ns, P)) return true;
#pragma unroll
    for (int k = 0; k < NC; ++k) P.cnt[k] = 0;
    return false;
}

// The arcs of clip edge K: the state at c_K (inside the subject or not) is
// carried from the last crossing before it along the clip boundary; with no
// crossing anywhere P is the clip polygon (some subject vertex outside it) or
// the subject.  Clip-vertex parts go to gc, crossings to emit.
template <int K, int NC, class Emit>
CR_HD void clip_arcs(const EdgesPlan<NC> &P, const double *subj, int ns, double coef, bool any_ev,
                     d3 (&gc)[NC], Emit &emit) {
    constexpr int K1 = (K + 1) % NC;
    const d3 &O = P.O;
    int in_start = P.any_out ? 1 : 0;
    if (any_ev) {
        bool found = false;
#pragma unroll
        for (int s = 1; s <= NC; ++s) {
            const int m = (K - s + NC) % NC;
            if (!found && P.cnt[m] > 0) {
                in_start = P.cnt[m] == 2 ? P.kind_hi[m] : P.kind_lo[m];
                found = true;
            }
        }
    }
    if (P.cnt[K] == 0) {
        if (in_start) {
            d3 ga, gb;
            fan_grad(O, P.c[K], P.c[K1], coef, ga, gb);
            gc[K] = gc[K] + ga;
            gc[K1] = gc[K1] + gb;
        }
        return;
    }
    const int32_t l1 = P.lab_lo[K];
    const d3 x1 = edge_cut_point(subj, ns, lab_hi(l1), P.c[K], P.n[K]);
    if (P.kind_lo[K] == 0) {  // an arc from c_K ends at the first crossing
        d3 ga, gb;
        fan_grad(O, P.c[K], x1, coef, ga, gb);
        gc[K] = gc[K] + ga;
        emit(l1, x1, gb);
    }
    if (P.cnt[K] == 1) {
        if (P.kind_lo[K] == 1) {  // an arc begins there and runs to c_K+1
            d3 ga, gb;
            fan_grad(O, x1, P.c[K1], coef, ga, gb);
            emit(l1, x1, ga);
            gc[K1] = gc[K1] + gb;
        }
        return;
    }
    const int32_t l2 = P.lab_hi[K];
    const d3 x2 = edge_cut_point(subj, ns, lab_hi(l2), P.c[K], P.n[K]);
    d3 ga, gb;
    if (P.kind_lo[K] == 1) {  // one arc between the two crossings
        fan_grad(O, x1, x2, coef, ga, gb);
        emit(l1, x1, ga);
        emit(l2, x2, gb);
    } else {  // the second crossing begins an arc to c_K+1
        fan_grad(O, x2, P.c[K1], coef, ga, gb);
        emit(l2, x2, ga);
        gc[K1] = gc[K1] + gb;
    }
}

template <int K, int NC, class Emit, class Sync>
CR_HD void clip_arcs_from(const EdgesPlan<NC> &P, const double *subj, int ns, double coef, bool any_ev,
                          d3 (&gc)[NC], Emit &emit, Sync &sync) {
    clip_arcs<K, NC>(P, subj, ns, coef, any_ev, gc, emit);
    sync();
    if constexpr (K + 1 < NC) clip_arcs_from<K + 1, NC>(P, subj, ns, coef, any_ev, gc, emit, sync);
}

// Pass 2: every arc's end-point derivatives through emit.  sync() is called
// after every subject edge and every clip edge, the same number of times
// whatever the pair (a caller running lanes in step may act there).
template <int NC, class Emit, class Sync>
CR_HD void edges_emit(const EdgesPlan<NC> &P, const double *subj, int ns, double coef, Emit &emit, Sync &&sync) {
    const d3 O = P.O;
    // the subject arcs
    const int32_t labS = lab_make(kLabS, 0, 0);
    d3 p = load3(subj), g_first = mk3(0.0, 0.0, 0.0), g_pend = mk3(0.0, 0.0, 0.0);
    double ap[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) ap[k] = dot(P.n[k], p - P.c[k]);
    for (int j = 0; j < ns; ++j) {
        const d3 q = load3(subj + 3 * (j + 1 == ns ? 0 : j + 1));
        double aq[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) aq[k] = dot(P.n[k], q - P.c[k]);
        EdgeCut<NC> cut;
        cut.template make<false>(ap, aq, ap);  // no positions: the third argument is not read
        d3 g_p = mk3(0.0, 0.0, 0.0), g_q = mk3(0.0, 0.0, 0.0);  // this edge's parts of s_j and s_j+1
        if (cut.inside()) {
            const d3 xa = cut.k_in >= 0 ? crossing(p, cut.ap_in, q, cut.aq_in) : p;
            const d3 xb = cut.k_out >= 0 ? crossing(p, cut.ap_out, q, cut.aq_out) : q;
            d3 ga, gb;
            fan_grad(O, xa, xb, coef, ga, gb);
            if (cut.k_in >= 0) emit(lab_make(kLabXF, j, cut.k_in), xa, ga);
            else g_p = ga;
            if (cut.k_out >= 0) emit(lab_make(kLabXF, j, cut.k_out), xb, gb);
            else g_q = gb;
        }
        if (j == 0) {
            g_first = g_p;
        } else {
            const d3 g = g_pend + g_p;  // subject vertex j: from edges j - 1 and j
            if (g.x != 0.0 || g.y != 0.0 || g.z != 0.0) emit(labS | j, p, g);
        }
        g_pend = g_q;
        p = q;
#pragma unroll
        for (int k = 0; k < NC; ++k) ap[k] = aq[k];
        sync();
    }
    {
        const d3 g = g_first + g_pend;  // vertex 0: from the first and the closing edge
        if (g.x != 0.0 || g.y != 0.0 || g.z != 0.0) emit(labS, p, g);
    }
    // the clip arcs, one clip edge at a time (K a template argument: every
    // index stays compile-time whatever the compiler unrolls)
    bool any_ev = false;
#pragma unroll
    for (int k = 0; k < NC; ++k) any_ev = any_ev || P.cnt[k] > 0;
    d3 gc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) gc[k] = mk3(0.0, 0.0, 0.0);
    clip_arcs_from<0, NC>(P, subj, ns, coef, any_ev, gc, emit, sync);
#pragma unroll
    for (int k = 0; k < NC; ++k)
        if (gc[k].x != 0.0 || gc[k].y != 0.0 || gc[k].z != 0.0) emit(lab_make(kLabC, 0, k), P.c[k], gc[k]);
}

// coef * d(signed ring area)/d(vertices) through emit(slot, grad): one call per
// vertex (a vertex's parts from the two fan triangles that share it are summed
// before the call).
template <class Emit>
CR_HD void ring_area_backward(const double *xyz, int n, double coef, Emit &&emit) {
    if (n < 3) return;
    d3 a = load3(xyz);
    d3 ga = mk3(0, 0, 0), pend = mk3(0, 0, 0);
    d3 b = load3(xyz + 3);
    for (int i = 1; i + 1 < n; ++i) {
        const d3 c = load3(xyz + 3 * (i + 1));
        double T = dot(a, cross(b, c));
        double D = 1.0 + dot(a, b) + dot(b, c) + dot(c, a);
        double f = coef * 2.0 / (T * T + D * D);
        ga = ga + f * (D * cross(b, c) - T * (b + c));
        emit(i, tangential(pend + f * (D * cross(c, a) - T * (a + c)), b));
        pend = f * (D * cross(a, b) - T * (a + b));
        b = c;
    }
    emit(n - 1, tangential(pend, b));
    emit(0, tangential(ga, a));
}

// Part of ring_area_backward's fan, unrolled (the slots are compile-time
// indices): the triangles (v0, v_i, v_{i+1}) for I0 <= i < I1.  ga (not made
// tangential) gets their derivative w.r.t. v0, g[j] (made
// tangential) w.r.t. v_{I0+j}, 0 <= j <= I1 - I0.  Two parts that meet at a
// vertex each hold a share of its derivative; the projection is linear, so
// the shares add.
template <int I0, int I1>
CR_HD void ring_area_backward_part(const double *xyz, double coef, d3 &ga, d3 (&g)[I1 - I0 + 1]) {
    const d3 a = load3(xyz);
    d3 pend = mk3(0, 0, 0);
    d3 b = load3(xyz + 3 * I0);
#pragma unroll
    for (int i = I0; i < I1; ++i) {
        const d3 c = load3(xyz + 3 * (i + 1));
        const double T = dot(a, cross(b, c));
        const double D = 1.0 + dot(a, b) + dot(b, c) + dot(c, a);
        const double f = coef * 2.0 / (T * T + D * D);
        ga = ga + f * (D * cross(b, c) - T * (b + c));
        g[i - I0] = g[i - I0] + tangential(pend + f * (D * cross(c, a) - T * (a + c)), b);
        pend = f * (D * cross(a, b) - T * (a + b));
        b = c;
    }
    g[I1 - I0] = g[I1 - I0] + tangential(pend, b);
}


}  // namespace cr
