/*
 * cr_oracle_planar.c -- CPU FP64 restatement of the build on the Planar manifold.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the device's planar path; the
 * product never links or calls it.
 *
 * What it restates (reference files relative to the reference repo root):
 *   Regridder(GO.Planar(), dst, src)   src/regridder/regridder.jl:125-163
 *   per-pair operator                  src/regridder/regridder.jl:87-94:
 *       GO.intersection(FosterHormannClipping(Planar()), p1, p2) then
 *       GO.area(Planar(), polygons)
 *   keep iff area > 0, sparse(dst, src) src/regridder/intersection_areas.jl:4-32, 115-121
 *   areas(Planar(), tree)              src/regridder/regridder.jl:165-178
 *
 * FosterHormannClipping lives in GeometryOps (compat 0.1.33 in the reference's
 * Project.toml; not vendored).  Its published method is the Greiner-Hormann
 * boundary walk with Foster, Hormann & Popa's degenerate-case rules ("Clipping
 * simple polygons with degenerate intersections", Computers & Graphics X 2,
 * 2019).  Where one polygon is convex, the intersection it returns is the
 * intersection of the other polygon with the convex one's half-planes, which
 * this file computes with Sutherland & Hodgman's re-entrant clip (CACM 17(1),
 * 1974):
 *   - rings are open (no repeated closing vertex) and counter-clockwise; the
 *     inside of edge a -> b holds the points with (b - a) x (p - a) >= 0;
 *   - a crossing of p -> q is p + t (q - p) with t = s_p / (s_p - s_q);
 *   - an edge of zero length clips nothing;
 *   - area: the shoelace formula, sum of (x_i y_{i+1} - x_{i+1} y_i) / 2.
 * Candidates are the pairs whose bounding boxes meet (closed boxes), a
 * superset of the intersecting pairs, so A does not depend on the search.
 *
 * Build: make -C oracle (gcc, -O2, no FMA contraction).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PMAXV 256

typedef struct { double x, y; } pt;

/* Shoelace area of an open ring of n points (signed, counter-clockwise > 0). */
double orc_planar_polygon_area(const double *xy, int32_t n) {
    if (n < 3) return 0.0;
    double s = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t j = i + 1 == n ? 0 : i + 1;
        s += xy[2 * i] * xy[2 * j + 1] - xy[2 * j] * xy[2 * i + 1];
    }
    return 0.5 * s;
}

static int planar_clip(const pt *S, int ns, const pt *C, int nc, pt *out, pt *tmp) {
    pt *cur = tmp, *nxt = out;
    if (ns > PMAXV) return -1;
    memcpy(cur, S, sizeof(pt) * ns);
    int m = ns;
    for (int k = 0; k < nc && m > 0; ++k) {
        const pt a = C[k], b = C[(k + 1) % nc];
        const double ex = b.x - a.x, ey = b.y - a.y;
        if (ex == 0.0 && ey == 0.0) continue;
        int w = 0;
        pt p = cur[m - 1];
        double sp = ex * (p.y - a.y) - ey * (p.x - a.x);
        for (int i = 0; i < m; ++i) {
            const pt q = cur[i];
            const double sq = ex * (q.y - a.y) - ey * (q.x - a.x);
            if ((sq >= 0.0) != (sp >= 0.0)) {
                if (w >= PMAXV) return -1;
                const double t = sp / (sp - sq);
                nxt[w].x = p.x + t * (q.x - p.x);
                nxt[w].y = p.y + t * (q.y - p.y);
                ++w;
            }
            if (sq >= 0.0) {
                if (w >= PMAXV) return -1;
                nxt[w++] = q;
            }
            p = q;
            sp = sq;
        }
        m = w;
        pt *t = cur; cur = nxt; nxt = t;
    }
    if (cur != out) memcpy(out, cur, sizeof(pt) * (m > 0 ? m : 0));
    return m;
}

/* Area of (subject ∩ clip), both open rings of (x, y) points, clip convex and
 * counter-clockwise.  NaN when a ring would pass PMAXV points. */
double orc_planar_clip_area(const double *subj, int32_t ns, const double *clip, int32_t nc) {
    pt S[PMAXV], C[PMAXV], out[PMAXV], tmp[PMAXV];
    if (ns > PMAXV || nc > PMAXV) return NAN;
    memcpy(S, subj, sizeof(pt) * ns);
    memcpy(C, clip, sizeof(pt) * nc);
    const int m = planar_clip(S, ns, C, nc, out, tmp);
    if (m < 0) return NAN;
    return m >= 3 ? orc_planar_polygon_area((const double *)out, m) : 0.0;
}

typedef struct {
    int64_t n_dst, n_src, nnz, n_candidates;
    int64_t *rowptr;
    int32_t *col;
    double *val, *a_dst, *a_src;
    int32_t status;
} orc_planar_result;

typedef struct { double x0, x1, y0, y1; } box_t;

static box_t ring_box(const double *xy, int nv) {
    box_t b = {INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int k = 0; k < nv; ++k) {
        const double x = xy[2 * k], y = xy[2 * k + 1];
        b.x0 = x < b.x0 ? x : b.x0;
        b.x1 = x > b.x1 ? x : b.x1;
        b.y0 = y < b.y0 ? y : b.y0;
        b.y1 = y > b.y1 ? y : b.y1;
    }
    return b;
}

/*
 * Build A (n_dst x n_src, CSR, columns ascending per row) from planar polygon
 * tables (ncells x nv x 2, open counter-clockwise rings; narrower cells repeat
 * their last point).  clip_is_dst picks the convex side.  A pair is kept iff
 * its area > threshold_rel * min(a_src, a_dst) and both cells have positive
 * area (threshold_rel = 0: every area > 0).
 */
orc_planar_result *orc_planar_build(const double *dst_xy, int64_t n_dst, int32_t dst_nv, const double *src_xy,
                                    int64_t n_src, int32_t src_nv, int32_t clip_is_dst, double threshold_rel) {
    orc_planar_result *R = calloc(1, sizeof(orc_planar_result));
    R->n_dst = n_dst;
    R->n_src = n_src;
    if (dst_nv > PMAXV / 2 || src_nv > PMAXV / 2) { R->status = 1; return R; }
    R->a_dst = malloc(sizeof(double) * (n_dst ? n_dst : 1));
    R->a_src = malloc(sizeof(double) * (n_src ? n_src : 1));
    box_t *bd = malloc(sizeof(box_t) * (n_dst ? n_dst : 1)), *bs = malloc(sizeof(box_t) * (n_src ? n_src : 1));
    for (int64_t d = 0; d < n_dst; ++d) {
        R->a_dst[d] = orc_planar_polygon_area(dst_xy + 2 * dst_nv * d, dst_nv);
        bd[d] = ring_box(dst_xy + 2 * dst_nv * d, dst_nv);
    }
    for (int64_t s = 0; s < n_src; ++s) {
        R->a_src[s] = orc_planar_polygon_area(src_xy + 2 * src_nv * s, src_nv);
        bs[s] = ring_box(src_xy + 2 * src_nv * s, src_nv);
    }
    int64_t cap = 1024, nnz = 0;
    R->rowptr = malloc(sizeof(int64_t) * (n_dst + 1));
    R->col = malloc(sizeof(int32_t) * cap);
    R->val = malloc(sizeof(double) * cap);
    R->rowptr[0] = 0;
    for (int64_t d = 0; d < n_dst; ++d) {
        for (int64_t s = 0; s < n_src; ++s) {  /* columns ascend with s */
            if (bs[s].x0 > bd[d].x1 || bs[s].x1 < bd[d].x0 || bs[s].y0 > bd[d].y1 || bs[s].y1 < bd[d].y0) continue;
            ++R->n_candidates;
            const double amin = R->a_src[s] < R->a_dst[d] ? R->a_src[s] : R->a_dst[d];
            if (!(amin > 0.0)) continue;
            const double *dp = dst_xy + 2 * dst_nv * d, *sp = src_xy + 2 * src_nv * s;
            const double a = clip_is_dst ? orc_planar_clip_area(sp, src_nv, dp, dst_nv)
                                         : orc_planar_clip_area(dp, dst_nv, sp, src_nv);
            if (isnan(a)) R->status = 2;
            if (!(a > threshold_rel * amin)) continue;
            if (nnz == cap) {
                cap *= 2;
                R->col = realloc(R->col, sizeof(int32_t) * cap);
                R->val = realloc(R->val, sizeof(double) * cap);
            }
            R->col[nnz] = (int32_t)s;
            R->val[nnz] = a;
            ++nnz;
        }
        R->rowptr[d + 1] = nnz;
    }
    R->nnz = nnz;
    free(bd);
    free(bs);
    return R;
}

int64_t orc_planar_nnz(const orc_planar_result *R) { return R->nnz; }
int64_t orc_planar_candidates(const orc_planar_result *R) { return R->n_candidates; }
int32_t orc_planar_status(const orc_planar_result *R) { return R->status; }
void orc_planar_copy(const orc_planar_result *R, int64_t *rowptr, int32_t *col, double *val, double *a_dst,
                     double *a_src) {
    memcpy(rowptr, R->rowptr, sizeof(int64_t) * (R->n_dst + 1));
    memcpy(col, R->col, sizeof(int32_t) * R->nnz);
    memcpy(val, R->val, sizeof(double) * R->nnz);
    memcpy(a_dst, R->a_dst, sizeof(double) * R->n_dst);
    memcpy(a_src, R->a_src, sizeof(double) * R->n_src);
}
void orc_planar_free(orc_planar_result *R) {
    if (!R) return;
    free(R->rowptr); free(R->col); free(R->val); free(R->a_dst); free(R->a_src); free(R);
}
