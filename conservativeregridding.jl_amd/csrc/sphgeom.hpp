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

CR_HD d3 crossing(const d3 &p, double sp, const d3 &q, double sq) {
    d3 x = fabs(sq) * p + fabs(sp) * q;
    return (1.0 / sqrt(dot(x, x))) * x;
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
    for (int i = 0; i < ns; ++i) pipe_push<0, NC>(P, load3(subj + 3 * i));
    pipe_close<0, NC>(P);
    return P.acc.area();
}

// Signed area of an open ring of n vertices.
CR_HD double ring_area(const double *xyz, int n) {
    AreaAcc a;
    a.init();
    for (int i = 0; i < n; ++i) a.push(load3(xyz + 3 * i));
    return a.area();
}

}  // namespace cr
