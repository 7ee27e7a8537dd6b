/*
 * cr_oracle.c -- CPU FP64 restatement of ConservativeRegridding.jl's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * (libcrhip.so) and the timed CPU baseline.  Only tests/, the smoke() entry and
 * bench.py's cpu_baseline leg load it; the product never links or calls it.
 *
 * What it restates (reference files relative to the reference repo root):
 *   Regridder(dst, src)          src/regridder/regridder.jl:125-163
 *     candidate search           src/regridder/intersection_areas.jl:48-65
 *                                (dual-tree DFS over bounding caps; any superset
 *                                 of the intersecting pairs gives the same A, so
 *                                 this file uses a sorted 3-D bucket grid of caps)
 *     per-pair clip + area       src/regridder/intersection_areas.jl:4-32,
 *                                regridder.jl:96-103 (ConvexConvexSutherlandHodgman
 *                                + spherical area; keep iff area > 0, here
 *                                area > threshold_rel * min(a_src, a_dst))
 *     sparse(dst, src, area)     src/regridder/intersection_areas.jl:115-121
 *                                (rows = dst, cols = src, columns sorted per row)
 *     areas(manifold, tree)      src/regridder/regridder.jl:165-178 (geometric)
 *     normalize!                 src/regridder/regridder.jl:54-62
 *   regrid! / transpose          src/regridder/regrid.jl:63-118,
 *                                regridder.jl:49-50  (d = A s ./ a_d, s~ = A' d ./ a_s)
 *
 * Geometry conventions (GeometryOps is not vendored with the reference, so its
 * exact formulas cannot be restated; these are the published ones):
 *   - points are unit 3-vectors; polygons are open rings (no repeated closing
 *     vertex), counter-clockwise seen from outside the sphere;
 *   - edge a->b lies on the great circle with normal n = a x (b - a); a point v is
 *     inside the edge's half-space iff n . (v - a) >= 0;
 *   - the crossing of edge p->q with that great circle is the chord point
 *     p + t (q - p), t = s_p / (s_p - s_q), normalised back to the sphere;
 *   - triangle solid angle (Van Oosterom & Strackee 1983):
 *       E = 2 atan2(a . ((b - a) x (c - a)), 1 + a.b + b.c + c.a);
 *     polygon area = sum of E over the fan from vertex 0 (signed).
 *   - an edge with |a x (b - a)| < 1e-13 (coincident pole vertices) is skipped.
 *
 * Build: make -C oracle   (gcc, -O2, pthreads).  No FMA contraction, so results
 * are reproducible across hosts.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct { double x, y, z; } v3;

/* FP64 operation counting for the roofline of the device clip: when g_fl is
 * set (orc_set_flop_counting), the clip and area below add the add, sub, mul,
 * div and sqrt operations they perform to *g_fl (atan2 counts as one). */
static _Thread_local int64_t *g_fl = NULL;
static int g_count_flops = 0;
#define FL(n) do { if (g_fl) *g_fl += (n); } while (0)

static inline v3 v3_sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline v3 v3_add(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline v3 v3_scale(v3 a, double s) { v3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static inline double v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 v3_cross(v3 a, v3 b) {
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static inline v3 v3_normalize(v3 a) { return v3_scale(a, 1.0 / sqrt(v3_dot(a, a))); }
static inline v3 v3_load(const double *p) { v3 r = {p[0], p[1], p[2]}; return r; }

#define EDGE_EPS 1e-13
#define MAXV 256

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ---------------------------------------------------------------- areas -- */

/* Solid angle of the spherical triangle (a, b, c), signed (positive if CCW). */
static double tri_excess(v3 a, v3 b, v3 c) {
    FL(40); /* 2 subs, cross, dot; 3 dots + 3 adds; atan2, scale */
    double num = v3_dot(a, v3_cross(v3_sub(b, a), v3_sub(c, a)));
    double den = 1.0 + v3_dot(a, b) + v3_dot(b, c) + v3_dot(c, a);
    return 2.0 * atan2(num, den);
}

static double ring_area(const v3 *p, int n) {
    double s = 0.0;
    for (int i = 1; i + 1 < n; ++i) s += tri_excess(p[0], p[i], p[i + 1]);
    FL(n > 2 ? n - 2 : 0);
    return s;
}

/* Signed area (steradians) of the open ring xyz[0..n-1]. */
double orc_polygon_area(const double *xyz, int32_t n) {
    v3 p[MAXV];
    if (n > MAXV) return NAN;
    for (int i = 0; i < n; ++i) p[i] = v3_load(xyz + 3 * i);
    return ring_area(p, n);
}

/* ----------------------------------------------------------------- clip -- */

/* Sutherland-Hodgman: clip subject polygon S (any simple CCW ring) by the
 * half-spaces of the convex CCW ring C.  Returns the number of output vertices
 * (written to out); buffers must hold MAXV vertices. */
static int sh_clip(const v3 *S, int ns, const v3 *C, int nc, v3 *out, v3 *tmp) {
    v3 *cur = tmp, *nxt = out;
    int m = ns;
    if (ns > MAXV / 2) return -1;
    memcpy(cur, S, sizeof(v3) * ns);
    for (int k = 0; k < nc && m > 0; ++k) {
        v3 a = C[k], b = C[(k + 1) % nc];
        v3 n = v3_cross(a, v3_sub(b, a));
        FL(18); /* normal (sub, cross), its norm (dot, sqrt) */
        if (sqrt(v3_dot(n, n)) < EDGE_EPS) continue; /* degenerate edge */
        int w = 0;
        v3 p = cur[m - 1];
        double sp = v3_dot(n, v3_sub(p, a));
        FL(8 + 8 * m); /* one side test (sub, dot) per vertex, plus the first */
        for (int i = 0; i < m; ++i) {
            v3 q = cur[i];
            double sq = v3_dot(n, v3_sub(q, a));
            if (sq >= 0.0) {
                if (sp < 0.0) {
                    double t = sp / (sp - sq);
                    nxt[w++] = v3_normalize(v3_add(p, v3_scale(v3_sub(q, p), t)));
                    FL(21); /* t (sub, div), lerp (sub, scale, add), normalise (dot, sqrt, div, scale) */
                }
                nxt[w++] = q;
            } else if (sp >= 0.0) {
                double t = sp / (sp - sq);
                nxt[w++] = v3_normalize(v3_add(p, v3_scale(v3_sub(q, p), t)));
                FL(21);
            }
            if (w >= MAXV - 2) return -1;
            p = q;
            sp = sq;
        }
        m = w;
        v3 *t = cur; cur = nxt; nxt = t;
    }
    if (cur != out) memcpy(out, cur, sizeof(v3) * (m > 0 ? m : 0));
    return m;
}

/* Area of (subject ∩ clip); clip must be convex.  NaN on buffer overflow. */
double orc_clip_area(const double *subj, int32_t ns, const double *clip, int32_t nc) {
    v3 S[MAXV], C[MAXV], out[MAXV], tmp[MAXV];
    if (ns > MAXV / 2 || nc > MAXV) return NAN;
    for (int i = 0; i < ns; ++i) S[i] = v3_load(subj + 3 * i);
    for (int i = 0; i < nc; ++i) C[i] = v3_load(clip + 3 * i);
    int m = sh_clip(S, ns, C, nc, out, tmp);
    if (m < 0) return NAN;
    return m >= 3 ? ring_area(out, m) : 0.0;
}

/* Writes the clipped ring (up to cap vertices) and returns its vertex count. */
int32_t orc_clip_polygon(const double *subj, int32_t ns, const double *clip, int32_t nc,
                         double *out_xyz, int32_t cap) {
    v3 S[MAXV], C[MAXV], out[MAXV], tmp[MAXV];
    if (ns > MAXV / 2 || nc > MAXV) return -1;
    for (int i = 0; i < ns; ++i) S[i] = v3_load(subj + 3 * i);
    for (int i = 0; i < nc; ++i) C[i] = v3_load(clip + 3 * i);
    int m = sh_clip(S, ns, C, nc, out, tmp);
    for (int i = 0; i < m && i < cap; ++i) {
        out_xyz[3 * i] = out[i].x; out_xyz[3 * i + 1] = out[i].y; out_xyz[3 * i + 2] = out[i].z;
    }
    return m;
}

/* ----------------------------------------------------------- candidates -- */

/* Bounding cap of a cell as a ball in R^3: centre = normalised vertex mean,
 * radius = max chord to a vertex.  For cells smaller than a hemisphere the cap
 * is spherically convex, so it holds every great-circle edge and the interior. */
typedef struct { double cx, cy, cz, r; } cap_t;

static cap_t cell_cap(const double *xyz, int nv) {
    v3 s = {0, 0, 0};
    for (int k = 0; k < nv; ++k) s = v3_add(s, v3_load(xyz + 3 * k));
    double L = sqrt(v3_dot(s, s));
    cap_t c;
    if (L < 1e-12) { c.cx = 0; c.cy = 0; c.cz = 1; c.r = 2.0; return c; }
    v3 m = v3_scale(s, 1.0 / L);
    double r = 0;
    for (int k = 0; k < nv; ++k) {
        v3 d = v3_sub(v3_load(xyz + 3 * k), m);
        double dd = sqrt(v3_dot(d, d));
        if (dd > r) r = dd;
    }
    /* chord radius > sqrt(2) means the cap exceeds a hemisphere: no pruning */
    if (r > 1.4) r = 2.0;
    c.cx = m.x; c.cy = m.y; c.cz = m.z; c.r = r * (1.0 + 1e-9) + 1e-12;
    return c;
}

typedef struct { int64_t key; int64_t idx; } keyidx_t;

static int cmp_keyidx(const void *a, const void *b) {
    const keyidx_t *x = a, *y = b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

typedef struct {
    int64_t nb;      /* buckets per axis */
    double h;        /* bucket width */
    double rmax;     /* largest src cap radius */
    keyidx_t *sorted;
    int64_t n;
} bucket_grid_t;

static inline int64_t bucket_coord(double v, double h, int64_t nb) {
    int64_t i = (int64_t)floor((v + 1.0) / h);
    if (i < 0) i = 0;
    if (i >= nb) i = nb - 1;
    return i;
}

static void bucket_grid_build(bucket_grid_t *g, const cap_t *caps, int64_t n) {
    double rmax = 0, rsum = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (caps[i].r > rmax) rmax = caps[i].r;
        rsum += caps[i].r;
    }
    double ravg = n ? rsum / n : 1.0;
    double h = 2.0 * ravg;
    if (h < 1e-6) h = 1e-6;
    int64_t nb = (int64_t)ceil(2.0 / h) + 1;
    if (nb > 100000) nb = 100000;
    h = 2.0 / (double)(nb - 1);
    g->nb = nb; g->h = h; g->rmax = rmax; g->n = n;
    g->sorted = malloc(sizeof(keyidx_t) * (n ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        int64_t ix = bucket_coord(caps[i].cx, h, nb), iy = bucket_coord(caps[i].cy, h, nb),
                iz = bucket_coord(caps[i].cz, h, nb);
        g->sorted[i].key = (ix * nb + iy) * nb + iz;
        g->sorted[i].idx = i;
    }
    qsort(g->sorted, n, sizeof(keyidx_t), cmp_keyidx);
}

static int64_t lower_bound_key(const keyidx_t *a, int64_t n, int64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (a[mid].key < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* ---------------------------------------------------------------- build -- */

typedef struct {
    int64_t n_dst, n_src, nnz, n_candidates;
    int64_t *rowptr;
    int32_t *col;
    double *val;
    double *a_dst, *a_src;
    double t_total, t_areas, t_pairs;
    int64_t clip_flops;      /* FP64 operations of every candidate's clip and area (when counted) */
    int64_t clip_flops_kept; /* the same over the candidates kept as non-zeros */
    int32_t status;
} orc_result;

typedef struct { int32_t col; double val; } entry_t;

typedef struct {
    /* inputs */
    const double *dst_xyz, *src_xyz;
    int32_t dst_nv, src_nv;
    int32_t clip_is_dst;
    double threshold_rel;
    const cap_t *dst_caps, *src_caps;
    const bucket_grid_t *grid;
    const double *a_dst, *a_src;
    int64_t row_begin, row_end;
    /* outputs: per-row entry lists */
    entry_t **row_entries;
    int32_t *row_count;
    _Atomic int64_t next_row;
    _Atomic int64_t n_candidates;
    _Atomic int64_t clip_flops, clip_flops_kept;
    _Atomic int32_t failed;
} build_ctx_t;

static int cmp_entry(const void *a, const void *b) {
    const entry_t *x = a, *y = b;
    return (x->col > y->col) - (x->col < y->col);
}

#define ROW_CHUNK 64

static void *build_worker(void *arg) {
    build_ctx_t *c = arg;
    const bucket_grid_t *g = c->grid;
    int64_t *cand = malloc(sizeof(int64_t) * 1024);
    int64_t cand_cap = 1024;
    entry_t *ebuf = malloc(sizeof(entry_t) * 1024);
    int64_t ebuf_cap = 1024;
    v3 S[MAXV], C[MAXV], out[MAXV], tmp[MAXV];
    int64_t local_cand = 0, local_fl = 0, local_fl_kept = 0;
    g_fl = g_count_flops ? &local_fl : NULL;
    for (;;) {
        int64_t r0 = atomic_fetch_add(&c->next_row, ROW_CHUNK);
        if (r0 >= c->row_end) break;
        int64_t r1 = r0 + ROW_CHUNK < c->row_end ? r0 + ROW_CHUNK : c->row_end;
        for (int64_t d = r0; d < r1; ++d) {
            cap_t qc = c->dst_caps[d];
            double reach = qc.r + g->rmax;
            int64_t ncand = 0;
            int64_t ix0 = bucket_coord(qc.cx - reach, g->h, g->nb), ix1 = bucket_coord(qc.cx + reach, g->h, g->nb);
            int64_t iy0 = bucket_coord(qc.cy - reach, g->h, g->nb), iy1 = bucket_coord(qc.cy + reach, g->h, g->nb);
            int64_t iz0 = bucket_coord(qc.cz - reach, g->h, g->nb), iz1 = bucket_coord(qc.cz + reach, g->h, g->nb);
            for (int64_t ix = ix0; ix <= ix1; ++ix)
                for (int64_t iy = iy0; iy <= iy1; ++iy) {
                    int64_t kbeg = (ix * g->nb + iy) * g->nb + iz0;
                    int64_t kend = (ix * g->nb + iy) * g->nb + iz1;
                    for (int64_t p = lower_bound_key(g->sorted, g->n, kbeg); p < g->n && g->sorted[p].key <= kend; ++p) {
                        int64_t s = g->sorted[p].idx;
                        /* zero-area cells (padding ghosts) are not in the
                         * reference's search trees: PaddedTreeWrapper,
                         * ext/ConservativeRegriddingOceananigansExt.jl:65-115 */
                        if (c->a_src[s] == 0.0 || c->a_dst[d] == 0.0) continue;
                        cap_t sc = c->src_caps[s];
                        double dx = sc.cx - qc.cx, dy = sc.cy - qc.cy, dz = sc.cz - qc.cz;
                        double rr = sc.r + qc.r;
                        if (dx * dx + dy * dy + dz * dz > rr * rr) continue;
                        if (ncand == cand_cap) { cand_cap *= 2; cand = realloc(cand, sizeof(int64_t) * cand_cap); }
                        cand[ncand++] = s;
                    }
                }
            local_cand += ncand;
            if (ncand > ebuf_cap) { ebuf_cap = ncand; ebuf = realloc(ebuf, sizeof(entry_t) * ebuf_cap); }
            const double *dp = c->dst_xyz + 3 * (int64_t)c->dst_nv * d;
            int64_t ne = 0;
            for (int64_t k = 0; k < ncand; ++k) {
                int64_t s = cand[k];
                const double *sp = c->src_xyz + 3 * (int64_t)c->src_nv * s;
                const double *subj = c->clip_is_dst ? sp : dp;
                const double *clip = c->clip_is_dst ? dp : sp;
                int ns = c->clip_is_dst ? c->src_nv : c->dst_nv;
                int nc = c->clip_is_dst ? c->dst_nv : c->src_nv;
                for (int i = 0; i < ns; ++i) S[i] = v3_load(subj + 3 * i);
                for (int i = 0; i < nc; ++i) C[i] = v3_load(clip + 3 * i);
                const int64_t fl0 = local_fl;
                int m = sh_clip(S, ns, C, nc, out, tmp);
                if (m < 0) { atomic_store(&c->failed, 1); continue; }
                double a = m >= 3 ? ring_area(out, m) : 0.0;
                double amin = c->a_src[s] < c->a_dst[d] ? c->a_src[s] : c->a_dst[d];
                if (a > c->threshold_rel * amin) {
                    ebuf[ne].col = (int32_t)s;
                    ebuf[ne].val = a;
                    ne++;
                    local_fl_kept += local_fl - fl0;
                }
            }
            qsort(ebuf, ne, sizeof(entry_t), cmp_entry);
            c->row_count[d] = (int32_t)ne;
            if (ne) {
                c->row_entries[d] = malloc(sizeof(entry_t) * ne);
                memcpy(c->row_entries[d], ebuf, sizeof(entry_t) * ne);
            }
        }
    }
    atomic_fetch_add(&c->n_candidates, local_cand);
    atomic_fetch_add(&c->clip_flops, local_fl);
    atomic_fetch_add(&c->clip_flops_kept, local_fl_kept);
    g_fl = NULL;
    free(cand);
    free(ebuf);
    return NULL;
}

typedef struct {
    const double *xyz; int32_t nv; int64_t n; double *area; cap_t *caps;
    _Atomic int64_t next;
} area_ctx_t;

static void *area_worker(void *arg) {
    area_ctx_t *c = arg;
    v3 p[MAXV];
    for (;;) {
        int64_t i0 = atomic_fetch_add(&c->next, 4096);
        if (i0 >= c->n) break;
        int64_t i1 = i0 + 4096 < c->n ? i0 + 4096 : c->n;
        for (int64_t i = i0; i < i1; ++i) {
            const double *q = c->xyz + 3 * (int64_t)c->nv * i;
            for (int k = 0; k < c->nv; ++k) p[k] = v3_load(q + 3 * k);
            c->area[i] = ring_area(p, c->nv);
            c->caps[i] = cell_cap(q, c->nv);
        }
    }
    return NULL;
}

static void run_threads(void *(*fn)(void *), void *ctx, int nthreads) {
    if (nthreads <= 1) { fn(ctx); return; }
    pthread_t *th = malloc(sizeof(pthread_t) * nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, fn, ctx);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
}

/* Cell areas (steradians, signed as given) for an (ncells x nv) polygon table. */
void orc_cell_areas(const double *xyz, int64_t ncells, int32_t nv, double *out, int32_t nthreads) {
    area_ctx_t c = {xyz, nv, ncells, out, malloc(sizeof(cap_t) * (ncells ? ncells : 1)), 0};
    run_threads(area_worker, &c, nthreads);
    free(c.caps);
}

/*
 * Build A (n_dst x n_src, CSR with columns sorted per row) for unit-sphere
 * polygon tables.  clip_is_dst selects which side is the (convex) clip polygon.
 * Only rows in [row_begin, row_end) are computed; the rest stay empty (bounded
 * CPU-baseline samples).  Entries with area <= threshold_rel * min(a_src, a_dst)
 * are dropped (threshold_rel = 0 keeps exactly the pairs with area > 0).
 * A cell of zero area is in no pair.
 */
orc_result *orc_build(const double *dst_xyz, int64_t n_dst, int32_t dst_nv,
                      const double *src_xyz, int64_t n_src, int32_t src_nv,
                      int32_t clip_is_dst, double threshold_rel, int32_t nthreads,
                      int64_t row_begin, int64_t row_end) {
    double t0 = now_s();
    orc_result *R = calloc(1, sizeof(orc_result));
    R->n_dst = n_dst; R->n_src = n_src;
    if (dst_nv > MAXV / 4 || src_nv > MAXV / 4) { R->status = 1; return R; }
    if (row_begin < 0) row_begin = 0;
    if (row_end > n_dst || row_end < 0) row_end = n_dst;
    R->a_dst = malloc(sizeof(double) * (n_dst ? n_dst : 1));
    R->a_src = malloc(sizeof(double) * (n_src ? n_src : 1));
    cap_t *dcaps = malloc(sizeof(cap_t) * (n_dst ? n_dst : 1));
    cap_t *scaps = malloc(sizeof(cap_t) * (n_src ? n_src : 1));
    area_ctx_t ad = {dst_xyz, dst_nv, n_dst, R->a_dst, dcaps, 0};
    run_threads(area_worker, &ad, nthreads);
    area_ctx_t as = {src_xyz, src_nv, n_src, R->a_src, scaps, 0};
    run_threads(area_worker, &as, nthreads);
    double t1 = now_s();
    R->t_areas = t1 - t0;

    bucket_grid_t g;
    bucket_grid_build(&g, scaps, n_src);

    build_ctx_t c;
    memset(&c, 0, sizeof(c));
    c.dst_xyz = dst_xyz; c.src_xyz = src_xyz; c.dst_nv = dst_nv; c.src_nv = src_nv;
    c.clip_is_dst = clip_is_dst; c.threshold_rel = threshold_rel;
    c.dst_caps = dcaps; c.src_caps = scaps; c.grid = &g; c.a_dst = R->a_dst; c.a_src = R->a_src;
    c.row_begin = row_begin; c.row_end = row_end;
    c.row_entries = calloc(n_dst ? n_dst : 1, sizeof(entry_t *));
    c.row_count = calloc(n_dst ? n_dst : 1, sizeof(int32_t));
    atomic_store(&c.next_row, row_begin);
    run_threads(build_worker, &c, nthreads);

    R->rowptr = malloc(sizeof(int64_t) * (n_dst + 1));
    R->rowptr[0] = 0;
    for (int64_t d = 0; d < n_dst; ++d) R->rowptr[d + 1] = R->rowptr[d] + c.row_count[d];
    R->nnz = R->rowptr[n_dst];
    R->col = malloc(sizeof(int32_t) * (R->nnz ? R->nnz : 1));
    R->val = malloc(sizeof(double) * (R->nnz ? R->nnz : 1));
    for (int64_t d = 0; d < n_dst; ++d) {
        int64_t o = R->rowptr[d];
        for (int32_t k = 0; k < c.row_count[d]; ++k) {
            R->col[o + k] = c.row_entries[d][k].col;
            R->val[o + k] = c.row_entries[d][k].val;
        }
        free(c.row_entries[d]);
    }
    R->n_candidates = atomic_load(&c.n_candidates);
    R->clip_flops = atomic_load(&c.clip_flops);
    R->clip_flops_kept = atomic_load(&c.clip_flops_kept);
    R->status = atomic_load(&c.failed) ? 2 : 0;
    free(c.row_entries); free(c.row_count); free(g.sorted); free(dcaps); free(scaps);
    double t2 = now_s();
    R->t_pairs = t2 - t1;
    R->t_total = t2 - t0;
    return R;
}

int64_t orc_result_nnz(const orc_result *R) { return R->nnz; }
int64_t orc_result_candidates(const orc_result *R) { return R->n_candidates; }
int64_t orc_result_clip_flops(const orc_result *R) { return R->clip_flops; }
int64_t orc_result_clip_flops_kept(const orc_result *R) { return R->clip_flops_kept; }
/* Count the FP64 operations of later builds' clips (on: 1) or not (0, the default). */
void orc_set_flop_counting(int32_t on) { g_count_flops = on != 0; }
/* FP64 operations of orc_clip_area on one pair (the same count a build adds per candidate). */
int64_t orc_clip_area_flops(const double *subj, int32_t ns, const double *clip, int32_t nc) {
    int64_t fl = 0;
    int64_t *saved = g_fl;
    g_fl = &fl;
    (void)orc_clip_area(subj, ns, clip, nc);
    g_fl = saved;
    return fl;
}
int32_t orc_result_status(const orc_result *R) { return R->status; }
void orc_result_times(const orc_result *R, double *out3) {
    out3[0] = R->t_total; out3[1] = R->t_areas; out3[2] = R->t_pairs;
}
void orc_result_copy(const orc_result *R, int64_t *rowptr, int32_t *col, double *val,
                     double *a_dst, double *a_src) {
    if (rowptr) memcpy(rowptr, R->rowptr, sizeof(int64_t) * (R->n_dst + 1));
    if (col) memcpy(col, R->col, sizeof(int32_t) * R->nnz);
    if (val) memcpy(val, R->val, sizeof(double) * R->nnz);
    if (a_dst) memcpy(a_dst, R->a_dst, sizeof(double) * R->n_dst);
    if (a_src) memcpy(a_src, R->a_src, sizeof(double) * R->n_src);
}
void orc_result_free(orc_result *R) {
    if (!R) return;
    free(R->rowptr); free(R->col); free(R->val); free(R->a_dst); free(R->a_src); free(R);
}

/* ---------------------------------------------------------------- apply -- */

typedef struct {
    const int64_t *rowptr; const int32_t *col; const double *val;
    int64_t nrows; const double *x; double *y; const double *area;
    _Atomic int64_t next;
} spmv_ctx_t;

static void *spmv_worker(void *arg) {
    spmv_ctx_t *c = arg;
    for (;;) {
        int64_t r0 = atomic_fetch_add(&c->next, 8192);
        if (r0 >= c->nrows) break;
        int64_t r1 = r0 + 8192 < c->nrows ? r0 + 8192 : c->nrows;
        for (int64_t i = r0; i < r1; ++i) {
            double s = 0.0;
            for (int64_t k = c->rowptr[i]; k < c->rowptr[i + 1]; ++k) s += c->val[k] * c->x[c->col[k]];
            c->y[i] = c->area ? s / c->area[i] : s;
        }
    }
    return NULL;
}

/* y = (A x) ./ area  (area == NULL: no normalisation).  regrid.jl:95-118. */
void orc_spmv(const int64_t *rowptr, const int32_t *col, const double *val, int64_t nrows,
              const double *x, double *y, const double *area, int32_t nthreads) {
    spmv_ctx_t c = {rowptr, col, val, nrows, x, y, area, 0};
    run_threads(spmv_worker, &c, nthreads);
}

/* CSR(A) -> CSR(A^T) (rows of the result sorted by column of A, stable). */
void orc_transpose(const int64_t *rowptr, const int32_t *col, const double *val,
                   int64_t nrows, int64_t ncols,
                   int64_t *t_rowptr, int32_t *t_col, double *t_val) {
    int64_t nnz = rowptr[nrows];
    memset(t_rowptr, 0, sizeof(int64_t) * (ncols + 1));
    for (int64_t k = 0; k < nnz; ++k) t_rowptr[col[k] + 1]++;
    for (int64_t j = 0; j < ncols; ++j) t_rowptr[j + 1] += t_rowptr[j];
    int64_t *pos = malloc(sizeof(int64_t) * (ncols ? ncols : 1));
    memcpy(pos, t_rowptr, sizeof(int64_t) * ncols);
    for (int64_t i = 0; i < nrows; ++i)
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            int64_t p = pos[col[k]]++;
            t_col[p] = (int32_t)i;
            t_val[p] = val[k];
        }
    free(pos);
}
