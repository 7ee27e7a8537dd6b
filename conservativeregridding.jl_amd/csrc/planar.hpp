// planar.hpp -- geometry of the Planar manifold (Regridder(GO.Planar(), dst, src)).
//
// The reference's planar intersection operator clips with GeometryOps'
// FosterHormannClipping and measures with GO.area(Planar())
// (src/regridder/regridder.jl:87-94).  With one side's cells convex (the
// condition the device build already puts on the spherical clip), the
// intersection polygon is the Sutherland-Hodgman one, so this file clips the
// subject ring by the convex ring's half-planes and takes the shoelace area.
//
// Planar grids keep their caller vertices (x, y) beside the unit-vector images
// the candidate search runs on (cr_grid_from_planar_polygons): a cell's ring is
// gathered through the grid's slot -> vertex table, so the slots follow the
// orientation the build gave the cell (counter-clockwise).
#pragma once

#include "common.hpp"

namespace cr {

struct p2 {
    double x, y;
};

CR_HD p2 planar_vertex(const double *xy, const int32_t *vidx, int64_t slot) {
    const int64_t v = vidx[slot];
    return p2{xy[2 * v], xy[2 * v + 1]};
}

// Signed area of an open ring (counter-clockwise positive): the fan from the
// first vertex, products of coordinates relative to it.
CR_HD double planar_ring_area(const p2 *r, int n) {
    if (n < 3) return 0.0;
    double s = 0.0;
    const double x0 = r[0].x, y0 = r[0].y;
    double ax = r[1].x - x0, ay = r[1].y - y0;
    for (int i = 2; i < n; ++i) {
        const double bx = r[i].x - x0, by = r[i].y - y0;
        s += ax * by - ay * bx;
        ax = bx;
        ay = by;
    }
    return 0.5 * s;
}

// Side of p against the directed line a -> a + e: > 0 left (inside of a
// counter-clockwise ring's edge), 0 on the line.
CR_HD double planar_side(const p2 &a, double ex, double ey, const p2 &p) {
    return ex * (p.y - a.y) - ey * (p.x - a.x);
}

// Area of (subject ∩ clip), clip convex and counter-clockwise, by
// Sutherland-Hodgman over the clip's edges (a point with side >= 0 is inside;
// an edge of zero length, the padding of a narrower cell, clips nothing).  A
// crossing is p + t (q - p), t = s_p / (s_p - s_q).  NaN when a ring would
// pass MAXV vertices (the build then reports the pair).
template <int MAXV>
CR_HD double planar_clip_area(const p2 *clip, int nc, const p2 *subj, int ns) {
    p2 bufA[MAXV], bufB[MAXV];
    if (ns > MAXV) return __builtin_nan("");
    for (int i = 0; i < ns; ++i) bufA[i] = subj[i];
    p2 *cur = bufA, *nxt = bufB;
    int m = ns;
    for (int k = 0; k < nc && m > 0; ++k) {
        const p2 a = clip[k], b = clip[k + 1 == nc ? 0 : k + 1];
        const double ex = b.x - a.x, ey = b.y - a.y;
        if (ex == 0.0 && ey == 0.0) continue;
        int w = 0;
        p2 p = cur[m - 1];
        double sp = planar_side(a, ex, ey, p);
        for (int i = 0; i < m; ++i) {
            const p2 q = cur[i];
            const double sq = planar_side(a, ex, ey, q);
            const bool qin = sq >= 0.0;
            if (qin != (sp >= 0.0)) {
                if (w >= MAXV) return __builtin_nan("");
                const double t = sp / (sp - sq);
                nxt[w++] = p2{p.x + t * (q.x - p.x), p.y + t * (q.y - p.y)};
            }
            if (qin) {
                if (w >= MAXV) return __builtin_nan("");
                nxt[w++] = q;
            }
            p = q;
            sp = sq;
        }
        m = w;
        p2 *t = cur;
        cur = nxt;
        nxt = t;
    }
    return m >= 3 ? planar_ring_area(cur, m) : 0.0;
}

}  // namespace cr
