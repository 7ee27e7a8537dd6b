// se.hip -- spectral-element weights (the ClimaCore extension's regridders).
//
// Reference: ext/ConservativeRegriddingClimaCoreExt.jl
//   accumulate_principled_b (:248-298)  per (element, cell) pair, the Nq x Nq
//       moments B[i,j] = sum over fan triangles (centroid, v_k, v_k+1) of the
//       intersection polygon, Dunavant rule points x = normalize(l1 c + l2 v1 +
//       l3 v2) with weight 2 A_t w, of phi_i(xi(x)) phi_j(eta(x)); repeated
//       vertices of the clipped polygon are merged first (merge_repeats)
//   se_to_fv_principled (:424-478)       SE -> FV: rows = cells, columns = nodes
//       e*Nq^2 + i + j*Nq, values B; dst areas = cell areas, src areas = node
//       weights W_{e,i,j}
//   fv_to_se_l2_projection (:562-676)    FV -> SE: per element, the local mass
//       matrix (Jacobian interpolated from W / (w_i w_j), GLL rule with
//       ceil((3Nq-2)/2)+1 points), rows rescaled to the moments' row sums,
//       restricted to covered nodes, inverted, applied to each cell's moments;
//       dst areas = 1, src areas = cell areas
//   inverse_element_map (:188-202), element_jacobian_at (:218-233)
//   reference_rule (src/regridder/triangle_quadrature.jl:10-156)
//
// The element map is ClimaCore's NormalizedBilinearMap: x(xi, eta) = p / |p|
// with p the bilinear interpolation of the corners c1 (-1,-1), c2 (1,-1),
// c3 (1,1), c4 (-1,1), here the element cell's four vertices in ring order.
// Its inverse is Newton's method on the two tangent-plane components of
// p(xi, eta) x x = 0 (bilinear, so it converges in a few steps from the centre).
//
// The moments (one thread per intersecting pair, FP64) run on the device; the
// FV -> SE per-element dense solves (Nq^2 x Nq^2) run on host threads.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>
#include <thread>
#include <vector>

#include "regridder.hpp"
#include "sphgeom.hpp"
#include "trirule.hpp"

namespace cr {

constexpr int kSeMaxNq = 8;
constexpr int kSeMaxCellV = 32;            // cells of at most this many vertices
constexpr int kSeMaxPoly = kSeMaxCellV + 8;  // clipped polygon buffer

struct NodeSet {
    int n;
    double x[kSeMaxNq];
    double w[kSeMaxNq];
};

// ---------------------------------------------------------------- rules --
namespace {

void add3(TriRule &r, double a, double b, double w) {
    const double p[3][3] = {{a, b, b}, {b, a, b}, {b, b, a}};
    for (auto &q : p) {
        r.l1[r.n] = q[0]; r.l2[r.n] = q[1]; r.l3[r.n] = q[2]; r.w[r.n] = w;
        ++r.n;
    }
}

void add6(TriRule &r, double a, double b, double c, double w) {
    const double p[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
    for (auto &q : p) {
        r.l1[r.n] = q[0]; r.l2[r.n] = q[1]; r.l3[r.n] = q[2]; r.w[r.n] = w;
        ++r.n;
    }
}

void add1(TriRule &r, double w) {
    r.l1[r.n] = r.l2[r.n] = r.l3[r.n] = 1.0 / 3.0;
    r.w[r.n] = w;
    ++r.n;
}

// Dunavant (1985) symmetric rules on the triangle of area 1/2, orbit by orbit
// in the point order of triangle_quadrature.jl's table.
bool dunavant(int degree, TriRule &r) {
    r.n = 0;
    switch (degree) {
    case 1: add1(r, 0.5); return true;
    case 2: add3(r, 2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0); return true;
    case 3: add1(r, -27.0 / 96.0); add3(r, 0.6, 0.2, 25.0 / 96.0); return true;
    case 4: {
        const double a = 0.445948490915965, b = 0.091576213509771;
        add3(r, 1 - 2 * a, a, 0.111690794839005);
        add3(r, 1 - 2 * b, b, 0.054975871827661);
        return true;
    }
    case 5: {
        const double s15 = std::sqrt(15.0), a = (6 + s15) / 21, b = (6 - s15) / 21;
        add1(r, 9.0 / 80.0);
        add3(r, 1 - 2 * a, a, (155 + s15) / 2400);
        add3(r, 1 - 2 * b, b, (155 - s15) / 2400);
        return true;
    }
    case 6:
        add3(r, 0.873821971016996, 0.063089014491502, 0.025422453185103);
        add3(r, 0.501426509658179, 0.249286745170910, 0.058393137863189);
        add6(r, 0.636502499121399, 0.310352451033785, 0.053145049844816, 0.041425537809187);
        return true;
    case 7:
        add1(r, -0.074785022233835);
        add3(r, 0.479308067841920, 0.260345966079040, 0.087807628716602);
        add3(r, 0.869739794195568, 0.065130102902216, 0.026673617804419);
        add6(r, 0.048690315425316, 0.312865496004875, 0.638444188569809, 0.038556880445129);
        return true;
    case 8:
        add1(r, 0.072157803838894);
        add3(r, 0.081414823414554, 0.459292588292723, 0.047545817133643);
        add3(r, 0.658861384496480, 0.170569307751760, 0.051608685267359);
        add3(r, 0.898905543365938, 0.050547228317031, 0.016229248811599);
        add6(r, 0.008394777409958, 0.263112829634638, 0.728492392955404, 0.013615157087218);
        return true;
    default: return false;
    }
}

}  // namespace

// reference_rule: the rule of this degree, or of the next degree that exists.
TriRule tri_rule(int degree) {
    TriRule r{};
    for (int d = std::max(degree, 1); d <= 8; ++d)
        if (dunavant(d, r)) return r;
    throw Error{CR_UNSUPPORTED, "triangle quadrature degree " + std::to_string(degree) +
                                    " not available (max is 8)"};
}

namespace {

// Gauss-Lobatto-Legendre nodes (ascending) and weights: -1, the roots of
// P'_{n-1} (Newton from the Chebyshev-Lobatto points), 1; w = 2 / (n(n-1) P_{n-1}^2).
void gll(int n, double *x, double *w) {
    const int N = n - 1;
    auto legendre = [N](double t, double &pn, double &pn1) {
        double p0 = 1.0, p1 = t;
        for (int k = 2; k <= N; ++k) {
            const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        pn = N == 0 ? 1.0 : p1;
        pn1 = N == 0 ? 0.0 : p0;
    };
    for (int i = 0; i < n; ++i) {
        double t = -std::cos(M_PI * i / N);
        if (i > 0 && i < N)
            for (int it = 0; it < 100; ++it) {
                double pn, pn1;
                legendre(t, pn, pn1);
                const double d1 = N * (pn1 - t * pn) / (1.0 - t * t);
                const double d2 = (2.0 * t * d1 - N * (N + 1.0) * pn) / (1.0 - t * t);
                const double step = d1 / d2;
                t -= step;
                if (std::fabs(step) < 1e-16) break;
            }
        double pn, pn1;
        legendre(t, pn, pn1);
        x[i] = t;
        w[i] = 2.0 / (N * (N + 1.0) * pn * pn);
    }
    for (int i = 0; i < n / 2; ++i) {  // exact symmetry
        const double a = 0.5 * (x[n - 1 - i] - x[i]), b = 0.5 * (w[i] + w[n - 1 - i]);
        x[i] = -a;
        x[n - 1 - i] = a;
        w[i] = w[n - 1 - i] = b;
    }
    if (n % 2) x[n / 2] = 0.0;
}

NodeSet node_set(int nq) {
    CR_REQUIRE(nq >= 2 && nq <= kSeMaxNq, CR_UNSUPPORTED,
               "spectral elements need 2 <= nq <= " + std::to_string(kSeMaxNq));
    NodeSet s{};
    s.n = nq;
    gll(nq, s.x, s.w);
    return s;
}

}  // namespace

// ------------------------------------------------------- element map --
struct Bilinear {
    d3 a, b, c, d;  // p(xi, eta) = a + b xi + c eta + d xi eta
    CR_HD void set(const double *cn) {
        const d3 c1 = load3(cn), c2 = load3(cn + 3), c3 = load3(cn + 6), c4 = load3(cn + 9);
        a = 0.25 * ((c1 + c2) + (c3 + c4));
        b = 0.25 * ((c2 + c3) - (c1 + c4));
        c = 0.25 * ((c3 + c4) - (c1 + c2));
        d = 0.25 * ((c1 + c3) - (c2 + c4));
    }
    CR_HD d3 p(double xi, double eta) const { return (a + xi * b) + (eta * c + (xi * eta) * d); }
    // |dx/dxi x dx/deta| of x = p/|p|
    CR_HD double jacobian(double xi, double eta) const {
        const d3 q = p(xi, eta);
        const double inv = 1.0 / sqrt(dot(q, q));
        const d3 x = inv * q;
        const d3 pxi = b + eta * d, peta = c + xi * d;
        const d3 txi = inv * (pxi - dot(x, pxi) * x), teta = inv * (peta - dot(x, peta) * x);
        const d3 n = cross(txi, teta);
        return sqrt(dot(n, n));
    }
    // (xi, eta) with p(xi, eta) parallel to x: Newton on the components of p
    // along two tangent directions of x.
    CR_HD void inverse(const d3 &x, double &xi, double &eta) const {
        const d3 e = fabs(x.x) < 0.6 ? mk3(1, 0, 0) : (fabs(x.y) < 0.6 ? mk3(0, 1, 0) : mk3(0, 0, 1));
        const d3 t1 = normalized(cross(x, e)), t2 = cross(x, t1);
        xi = 0.0;
        eta = 0.0;
        for (int it = 0; it < 30; ++it) {
            const d3 q = p(xi, eta), pxi = b + eta * d, peta = c + xi * d;
            const double f1 = dot(t1, q), f2 = dot(t2, q);
            const double j11 = dot(t1, pxi), j12 = dot(t1, peta), j21 = dot(t2, pxi), j22 = dot(t2, peta);
            const double det = j11 * j22 - j12 * j21;
            const double dx = (f1 * j22 - f2 * j12) / det, de = (j11 * f2 - j21 * f1) / det;
            xi -= dx;
            eta -= de;
            if (fabs(dx) + fabs(de) < 1e-15) break;
        }
    }
};

template <int NQ>
CR_HD void lagrange(const NodeSet &ns, double t, double *phi) {
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        double v = 1.0;
#pragma unroll
        for (int k = 0; k < NQ; ++k)
            if (k != i) v *= (t - ns.x[k]) / (ns.x[i] - ns.x[k]);
        phi[i] = v;
    }
}

// Sutherland-Hodgman of a subject ring by a convex ring into out[] (tmp is
// scratch of the same size); -1 if the buffer would overflow.
CR_HD int clip_polygon(const double *clip, int nc, const double *subj, int ns, d3 *out, d3 *tmp) {
    if (ns > kSeMaxPoly) return -1;
    d3 *cur = tmp, *nxt = out;
    for (int i = 0; i < ns; ++i) cur[i] = load3(subj + 3 * i);
    int m = ns;
    for (int k = 0; k < nc && m > 0; ++k) {
        const d3 a = load3(clip + 3 * k), bb = load3(clip + 3 * ((k + 1) % nc));
        const d3 n = cross(a, bb - a);
        if (dot(n, n) < kDegenerateEdge2) continue;
        int w = 0;
        d3 p = cur[m - 1];
        double sp = dot(n, p - a);
        for (int i = 0; i < m; ++i) {
            const d3 q = cur[i];
            const double sq = dot(n, q - a);
            if (w > kSeMaxPoly - 2) return -1;
            if (sq >= 0.0) {
                if (sp < 0.0) nxt[w++] = crossing(p, sp, q, sq);
                nxt[w++] = q;
            } else if (sp >= 0.0) {
                nxt[w++] = crossing(p, sp, q, sq);
            }
            p = q;
            sp = sq;
        }
        m = w;
        d3 *t = cur;
        cur = nxt;
        nxt = t;
    }
    if (cur != out)
        for (int i = 0; i < m; ++i) out[i] = cur[i];
    return m;
}

// Merge consecutive (cyclically) vertices closer than the degenerate-edge chord:
// a clip repeats a vertex where a cell corner lies on a clip edge (the pole
// corners of lon-lat cells), how many times depends on rounding, and the fan
// centre is the vertex mean.  Returns the new count (m < 0 passes through).
CR_HD int merge_repeats(d3 *poly, int m) {
    if (m <= 1) return m;
    int w = 1;
    for (int v = 1; v < m; ++v) {
        const d3 d = poly[v] - poly[w - 1];
        if (dot(d, d) >= kDegenerateEdge2) poly[w++] = poly[v];
    }
    while (w > 1) {
        const d3 d = poly[w - 1] - poly[0];
        if (dot(d, d) >= kDegenerateEdge2) break;
        --w;
    }
    return w;
}

// Entry k of CSR rows -> its row.
__global__ void k_entry_rows(const int64_t *__restrict__ rowptr, int64_t nrows, int32_t *__restrict__ rowid) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) rowid[k] = (int32_t)r;
}

// Moments of every non-zero (cell row, element column) of CSR(A); B[k*NQ^2 +
// i + j*NQ], in area units (scale = radius^2).  err is set if a clip buffer
// overflows (the pair then holds zeros).
template <int NQ>
__global__ void __launch_bounds__(64) k_se_moments(const int32_t *__restrict__ rowid,
                                                   const int32_t *__restrict__ col, int64_t nnz,
                                                   const double *__restrict__ cells, int cell_nv, int cell_convex,
                                                   const double *__restrict__ elems, TriRule rule, NodeSet ns,
                                                   double scale, double *__restrict__ B, int *__restrict__ err) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const double *cp = cells + 3 * (int64_t)cell_nv * rowid[k];
    const double *ep = elems + 12 * (int64_t)col[k];
    d3 poly[kSeMaxPoly], tmp[kSeMaxPoly];
    // the element is the subject clipped by a convex cell (GO.intersection(elem, cell));
    // a non-convex cell is clipped by the (convex) element instead: the same region
    const int m0 = cell_convex ? clip_polygon(cp, cell_nv, ep, 4, poly, tmp) : clip_polygon(ep, 4, cp, cell_nv, poly, tmp);
    const int m = merge_repeats(poly, m0);
    double acc[NQ * NQ];
#pragma unroll
    for (int q = 0; q < NQ * NQ; ++q) acc[q] = 0.0;
    if (m < 0) atomicExch(err, 1);
    if (m >= 3) {
        Bilinear E;
        E.set(ep);
        d3 cen = mk3(0, 0, 0);
        for (int v = 0; v < m; ++v) cen = cen + poly[v];
        cen = normalized((1.0 / m) * cen);
        for (int v = 0; v < m; ++v) {
            const d3 v1 = poly[v], v2 = poly[v + 1 < m ? v + 1 : 0];
            AreaAcc t;
            t.init();
            t.push(cen);
            t.push(v1);
            t.push(v2);
            const double at = t.area();
            if (at <= 0.0) continue;
            for (int q = 0; q < rule.n; ++q) {
                const d3 x = normalized((rule.l1[q] * cen + rule.l2[q] * v1) + rule.l3[q] * v2);
                double xi, eta, px[NQ], py[NQ];
                E.inverse(x, xi, eta);
                lagrange<NQ>(ns, xi, px);
                lagrange<NQ>(ns, eta, py);
                const double wa = rule.w[q] * 2.0 * at;
#pragma unroll
                for (int j = 0; j < NQ; ++j)
#pragma unroll
                    for (int i = 0; i < NQ; ++i) acc[i + j * NQ] += wa * px[i] * py[j];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NQ * NQ; ++q) B[k * (NQ * NQ) + q] = acc[q] * scale;
}

// Node positions and weights W = w_i w_j J(xi_i, eta_j) r^2 of every element.
__global__ void k_se_nodes(const double *__restrict__ elems, int64_t nelem, NodeSet ns, double r2,
                           double *__restrict__ xyz, double *__restrict__ wj) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Q = ns.n * ns.n;
    if (t >= nelem * Q) return;
    const int64_t e = t / Q;
    const int q = (int)(t - e * Q), i = q % ns.n, j = q / ns.n;
    Bilinear E;
    E.set(elems + 12 * e);
    if (xyz) {
        const d3 x = normalized(E.p(ns.x[i], ns.x[j]));
        xyz[3 * t] = x.x;
        xyz[3 * t + 1] = x.y;
        xyz[3 * t + 2] = x.z;
    }
    if (wj) wj[t] = ns.w[i] * ns.w[j] * E.jacobian(ns.x[i], ns.x[j]) * r2;
}

// SE -> FV matrix from the pair CSR: entry (k, q) -> row of k, column
// col[k]*Q + q, value B[k*Q + q] (rows keep their pairs' order).
__global__ void k_se_expand(const int32_t *__restrict__ rowid, const int32_t *__restrict__ col, int64_t nnz, int Q,
                            int32_t *__restrict__ ocol, int32_t *__restrict__ orow) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnz * Q) return;
    const int64_t k = t / Q;
    ocol[t] = col[k] * Q + (int32_t)(t - k * Q);
    orow[t] = rowid[k];
}

__global__ void k_scale_rowptr(const int64_t *__restrict__ rp, int64_t n, int64_t Q, int64_t *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= n) out[r] = rp[r] * Q;
}

namespace {

void check_elements(const cr_grid *elems) {
    CR_REQUIRE(!elems->planar, CR_UNSUPPORTED, "spectral elements live on the sphere, not on a planar grid");
    CR_REQUIRE(elems->nv == 4, CR_INVALID_ARGUMENT, "spectral elements must be quadrilaterals");
    CR_REQUIRE(elems->flipped == 0, CR_INVALID_ARGUMENT,
               "element corners must run counter-clockwise from (-1,-1): (-1,-1), (1,-1), (1,1), (-1,1)");
    CR_REQUIRE(elems->ncells * kSeMaxNq * kSeMaxNq < (int64_t(1) << 31), CR_UNSUPPORTED,
               "too many spectral-element nodes");
}

// The pair regridder: rows = cells (destination), columns = elements (source).
cr_regridder *pair_regridder(const cr_grid *cells, const cr_grid *elems, const cr_build_opts *opts, hipStream_t s) {
    cr_build_opts o;
    if (opts) o = *opts; else cr_build_opts_default(&o);
    CR_REQUIRE(o.radius > 0.0 && o.threshold_rel >= 0.0, CR_INVALID_ARGUMENT, "bad build options");
    o.normalize = 0;
    auto *r = new cr_regridder();
    try {
        build_regridder(r, cells, elems, o, s);
    } catch (...) {
        delete r;
        throw;
    }
    return r;
}

template <int NQ>
void launch_moments(const cr_regridder *R, const cr_grid *cells, const cr_grid *elems, const TriRule &rule,
                    const NodeSet &ns, double scale, double *B, int32_t *pair_cell, int32_t *pair_elem,
                    hipStream_t s) {
    const Csr &F = R->fwd;
    if (F.nnz == 0) return;
    DevBuf<int32_t> rowid(F.nnz, s);
    DevBuf<int> err(1, s);
    CR_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_entry_rows, dim3(grid_for(F.nrows, 256)), dim3(256), 0, s, F.rowptr.p, F.nrows, rowid.p);
    if (pair_cell)
        CR_HIP(hipMemcpyAsync(pair_cell, rowid.p, F.nnz * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    if (pair_elem)
        CR_HIP(hipMemcpyAsync(pair_elem, F.col.p, F.nnz * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_se_moments<NQ>, dim3(grid_for(F.nnz, 64)), dim3(64), 0, s, rowid.p, F.col.p, F.nnz,
                       cells->cells.p, cells->nv, cells->convex, elems->cells.p, rule, ns, scale, B, err.p);
    CR_HIP(hipGetLastError());
    int h = 0;
    CR_HIP(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    CR_HIP(hipStreamSynchronize(s));
    CR_REQUIRE(h == 0, CR_INTERNAL, "spectral-element moments: clipped polygon buffer overflow");
}

void moments(const cr_regridder *R, const cr_grid *cells, const cr_grid *elems, int nq, int degree, double scale,
             double *B, hipStream_t s, int32_t *pair_cell = nullptr, int32_t *pair_elem = nullptr) {
    check_elements(elems);
    CR_REQUIRE(!cells->planar, CR_UNSUPPORTED, "spectral-element weights need a spherical cell grid");
    CR_REQUIRE(cells->nv <= kSeMaxCellV, CR_UNSUPPORTED,
               "spectral-element weights need cells of at most " + std::to_string(kSeMaxCellV) + " vertices");
    CR_REQUIRE(R->fwd.nrows == cells->ncells && R->fwd.ncols == elems->ncells, CR_DIMENSION_MISMATCH,
               "the pair regridder must map the elements (source) to the cells (destination)");
    const NodeSet ns = node_set(nq);
    const TriRule rule = tri_rule(degree > 0 ? degree : 2 * (nq - 1));
    switch (nq) {
    case 2: launch_moments<2>(R, cells, elems, rule, ns, scale, B, pair_cell, pair_elem, s); break;
    case 3: launch_moments<3>(R, cells, elems, rule, ns, scale, B, pair_cell, pair_elem, s); break;
    case 4: launch_moments<4>(R, cells, elems, rule, ns, scale, B, pair_cell, pair_elem, s); break;
    case 5: launch_moments<5>(R, cells, elems, rule, ns, scale, B, pair_cell, pair_elem, s); break;
    case 6: launch_moments<6>(R, cells, elems, rule, ns, scale, B, pair_cell, pair_elem, s); break;
    case 7: launch_moments<7>(R, cells, elems, rule, ns, scale, B, pair_cell, pair_elem, s); break;
    default: launch_moments<8>(R, cells, elems, rule, ns, scale, B, pair_cell, pair_elem, s); break;
    }
}

// Node weights of the elements (caller's, or the map's own).
void node_weights(const cr_grid *elems, const NodeSet &ns, double r2, const double *user, DevBuf<double> &wj,
                  hipStream_t s) {
    const int64_t n = elems->ncells * ns.n * ns.n;
    wj.alloc_kept(n, s);
    if (user) {
        CR_HIP(hipMemcpyAsync(wj.p, user, n * sizeof(double), hipMemcpyHostToDevice, s));
        CR_HIP(hipStreamSynchronize(s));
    } else {
        hipLaunchKernelGGL(k_se_nodes, dim3(grid_for(n, 256)), dim3(256), 0, s, elems->cells.p, elems->ncells, ns,
                           r2, nullptr, wj.p);
        CR_HIP(hipGetLastError());
    }
}

// Gauss-Jordan inverse with partial pivoting of an n x n row-major matrix.
bool invert(std::vector<double> &a, int n) {
    std::vector<double> inv(n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
        if (a[p * n + c] == 0.0) return false;
        if (p != c)
            for (int k = 0; k < n; ++k) {
                std::swap(a[p * n + k], a[c * n + k]);
                std::swap(inv[p * n + k], inv[c * n + k]);
            }
        const double d = 1.0 / a[c * n + c];
        for (int k = 0; k < n; ++k) {
            a[c * n + k] *= d;
            inv[c * n + k] *= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                a[r * n + k] -= f * a[c * n + k];
                inv[r * n + k] -= f * inv[c * n + k];
            }
        }
    }
    a.swap(inv);
    return true;
}

void lagrange_host(const double *x, int n, double t, double *phi) {
    for (int i = 0; i < n; ++i) {
        double v = 1.0;
        for (int k = 0; k < n; ++k)
            if (k != i) v *= (t - x[k]) / (x[i] - x[k]);
        phi[i] = v;
    }
}

}  // namespace
}  // namespace cr

using namespace cr;

extern "C" {

cr_status cr_se_tri_rule(int32_t degree, double *bary, double *w, int32_t *npts) {
    return guard([&] {
        CR_REQUIRE(bary && w && npts, CR_INVALID_ARGUMENT, "cr_se_tri_rule: null argument");
        const TriRule r = tri_rule(degree);
        for (int q = 0; q < r.n; ++q) {
            bary[3 * q] = r.l1[q];
            bary[3 * q + 1] = r.l2[q];
            bary[3 * q + 2] = r.l3[q];
            w[q] = r.w[q];
        }
        *npts = r.n;
    });
}

cr_status cr_se_gll(int32_t n, double *x, double *w) {
    return guard([&] {
        CR_REQUIRE(x && w && n >= 2 && n <= 64, CR_INVALID_ARGUMENT, "cr_se_gll: need 2 <= n <= 64");
        gll(n, x, w);
    });
}

cr_status cr_se_nodes(const cr_grid *elems, int32_t nq, double radius, double *xyz, double *weights) {
    return guard([&] {
        CR_REQUIRE(elems && radius > 0.0 && (xyz || weights), CR_INVALID_ARGUMENT, "cr_se_nodes: bad arguments");
        check_elements(elems);
        const NodeSet ns = node_set(nq);
        const int64_t n = elems->ncells * nq * nq;
        DevBuf<double> dx(xyz ? 3 * n : 0), dw(weights ? n : 0);
        hipLaunchKernelGGL(k_se_nodes, dim3(grid_for(n, 256)), dim3(256), 0, 0, elems->cells.p, elems->ncells, ns,
                           radius * radius, xyz ? dx.p : nullptr, weights ? dw.p : nullptr);
        CR_HIP(hipGetLastError());
        if (xyz) CR_HIP(hipMemcpy(xyz, dx.p, dx.bytes(), hipMemcpyDeviceToHost));
        if (weights) CR_HIP(hipMemcpy(weights, dw.p, dw.bytes(), hipMemcpyDeviceToHost));
    });
}

cr_status cr_se_moments(const cr_regridder *R, const cr_grid *cells, const cr_grid *elems, int32_t nq,
                        int32_t tri_degree, double *B, int32_t *pair_cell, int32_t *pair_elem, void *stream) {
    return guard([&] {
        CR_REQUIRE(R && cells && elems && (B || R->fwd.nnz == 0), CR_INVALID_ARGUMENT, "cr_se_moments: null argument");
        CR_REQUIRE(R->area_scale > 0.0, CR_UNSUPPORTED, "cr_se_moments needs a regridder built from grids");
        moments(R, cells, elems, nq, tri_degree, R->area_scale, B, as_stream(stream), pair_cell, pair_elem);
    });
}

cr_status cr_se_regridder_to_fv(cr_regridder **out, const cr_grid *cells, const cr_grid *elems, int32_t nq,
                                int32_t tri_degree, const double *node_weights_host, const cr_build_opts *opts,
                                void *stream) {
    return guard([&] {
        CR_REQUIRE(out && cells && elems, CR_INVALID_ARGUMENT, "cr_se_regridder_to_fv: null argument");
        check_device();
        check_elements(elems);
        hipStream_t s = as_stream(stream);
        std::unique_ptr<cr_regridder> P(pair_regridder(cells, elems, opts, s));
        const int Q = nq * nq;
        const int64_t nnz = P->fwd.nnz * Q, nodes = elems->ncells * Q;
        CR_REQUIRE(nnz < (int64_t(1) << 31), CR_UNSUPPORTED, "more than 2^31 spectral-element weights");
        auto R = std::make_unique<cr_regridder>();
        Csr &F = R->fwd;
        F.nrows = cells->ncells;
        F.ncols = nodes;
        F.nnz = nnz;
        F.rowptr.alloc_kept(F.nrows + 1, s);
        F.col.alloc_kept(nnz, s);
        F.val.alloc_kept(nnz, s);
        moments(P.get(), cells, elems, nq, tri_degree, P->area_scale, F.val.p, s);
        {
            DevBuf<int32_t> rowid(P->fwd.nnz, s), orow(nnz, s);
            hipLaunchKernelGGL(k_entry_rows, dim3(grid_for(F.nrows, 256)), dim3(256), 0, s, P->fwd.rowptr.p,
                               F.nrows, rowid.p);
            if (nnz)
                hipLaunchKernelGGL(k_se_expand, dim3(grid_for(nnz, 256)), dim3(256), 0, s, rowid.p, P->fwd.col.p,
                                   P->fwd.nnz, Q, F.col.p, orow.p);
            hipLaunchKernelGGL(k_scale_rowptr, dim3(grid_for(F.nrows + 1, 256)), dim3(256), 0, s, P->fwd.rowptr.p,
                               F.nrows, (int64_t)Q, F.rowptr.p);
            CR_HIP(hipGetLastError());
            Temp T(s);
            csr_transpose(F, orow.p, R->bwd, T, s);
        }
        R->a_dst = std::move(P->a_dst);
        node_weights(elems, node_set(nq), P->area_scale, node_weights_host, R->a_src, s);
        finalize_regridder(R.get(), s);
        CR_HIP(hipStreamSynchronize(s));
        R->stats = P->stats;
        R->stats.n_src = nodes;
        R->stats.nnz = nnz;
        *out = R.release();
    });
}

cr_status cr_se_regridder_from_fv(cr_regridder **out, const cr_grid *elems, const cr_grid *cells, int32_t nq,
                                  int32_t tri_degree, const double *node_weights_host, const cr_build_opts *opts,
                                  void *stream) {
    return guard([&] {
        CR_REQUIRE(out && cells && elems, CR_INVALID_ARGUMENT, "cr_se_regridder_from_fv: null argument");
        check_device();
        check_elements(elems);
        hipStream_t s = as_stream(stream);
        std::unique_ptr<cr_regridder> P(pair_regridder(cells, elems, opts, s));
        const NodeSet ns = node_set(nq);
        const int Q = nq * nq;
        const int64_t ne = elems->ncells, nc = cells->ncells, np = P->fwd.nnz;
        // moments and pair structure on the host
        std::vector<double> B((size_t)np * Q), wj((size_t)ne * Q), a_cell(nc);
        std::vector<int64_t> rp(nc + 1);
        std::vector<int32_t> col(np);
        if (np) {
            DevBuf<double> dB((size_t)np * Q, s);
            moments(P.get(), cells, elems, nq, tri_degree, P->area_scale, dB.p, s);
            CR_HIP(hipMemcpyAsync(B.data(), dB.p, dB.bytes(), hipMemcpyDeviceToHost, s));
            CR_HIP(hipMemcpyAsync(col.data(), P->fwd.col.p, np * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        }
        CR_HIP(hipMemcpyAsync(rp.data(), P->fwd.rowptr.p, (nc + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        CR_HIP(hipMemcpyAsync(a_cell.data(), P->a_dst.p, nc * sizeof(double), hipMemcpyDeviceToHost, s));
        {
            DevBuf<double> dw;
            node_weights(elems, ns, P->area_scale, node_weights_host, dw, s);
            CR_HIP(hipMemcpyAsync(wj.data(), dw.p, dw.bytes(), hipMemcpyDeviceToHost, s));
            CR_HIP(hipStreamSynchronize(s));
        }
        // pairs of each element, cells ascending
        std::vector<int64_t> eoff(ne + 1, 0);
        for (int64_t k = 0; k < np; ++k) eoff[col[k] + 1]++;
        for (int64_t e = 0; e < ne; ++e) eoff[e + 1] += eoff[e];
        std::vector<int64_t> epair(np), ecell(np);
        {
            std::vector<int64_t> pos(eoff.begin(), eoff.end() - 1);
            for (int64_t c = 0; c < nc; ++c)
                for (int64_t k = rp[c]; k < rp[c + 1]; ++k) {
                    const int64_t p = pos[col[k]]++;
                    epair[p] = k;
                    ecell[p] = c;
                }
        }
        // mass-matrix quadrature: GLL with ceil((3Nq-2)/2)+1 points
        const int nm = (3 * nq - 2 + 1) / 2 + 1;
        std::vector<double> xm(nm), wm(nm), phm((size_t)nm * nq);
        gll(nm, xm.data(), wm.data());
        for (int p = 0; p < nm; ++p) lagrange_host(ns.x, nq, xm[p], &phm[(size_t)p * nq]);
        // per element: entries (node row, cell, value), cells ascending
        std::vector<std::vector<std::pair<int32_t, double>>> rows((size_t)ne * Q);
        std::atomic<bool> singular{false};
        auto work = [&](int64_t e0, int64_t e1) {
            std::vector<double> M((size_t)Q * Q), bfull(Q);
            std::vector<int> cov;
            for (int64_t e = e0; e < e1; ++e) {
                if (eoff[e] == eoff[e + 1]) continue;
                std::fill(bfull.begin(), bfull.end(), 0.0);
                for (int64_t p = eoff[e]; p < eoff[e + 1]; ++p)
                    for (int q = 0; q < Q; ++q) bfull[q] += B[(size_t)epair[p] * Q + q];
                cov.clear();
                for (int q = 0; q < Q; ++q)
                    if (bfull[q] != 0.0) cov.push_back(q);
                if (cov.empty()) continue;
                // M[(a,b),(i,j)] = sum_pq wm_p wm_q J(x_p, x_q) phi_a phi_i (x_p) phi_b phi_j (x_q)
                std::fill(M.begin(), M.end(), 0.0);
                for (int qq = 0; qq < nm; ++qq)
                    for (int pp = 0; pp < nm; ++pp) {
                        const double *fx = &phm[(size_t)pp * nq], *fy = &phm[(size_t)qq * nq];
                        double J = 0.0;
                        for (int b = 0; b < nq; ++b)
                            for (int a = 0; a < nq; ++a)
                                J += wj[(size_t)e * Q + a + b * nq] / (ns.w[a] * ns.w[b]) * fx[a] * fy[b];
                        const double wJ = wm[pp] * wm[qq] * J;
                        for (int jb = 0; jb < nq; ++jb)
                            for (int ja = 0; ja < nq; ++ja) {
                                const int r = ja + jb * nq;
                                const double fr = wJ * fx[ja] * fy[jb];
                                for (int jj = 0; jj < nq; ++jj)
                                    for (int ii = 0; ii < nq; ++ii) M[(size_t)r * Q + ii + jj * nq] += fr * fx[ii] * fy[jj];
                            }
                    }
                const int nc2 = (int)cov.size();
                std::vector<double> Mc((size_t)nc2 * nc2);
                for (int a = 0; a < nc2; ++a)
                    for (int b = 0; b < nc2; ++b) Mc[(size_t)a * nc2 + b] = M[(size_t)cov[a] * Q + cov[b]];
                for (int a = 0; a < nc2; ++a) {  // row sums -> the moments' sums
                    double rs = 0.0;
                    for (int b = 0; b < nc2; ++b) rs += Mc[(size_t)a * nc2 + b];
                    if (rs == 0.0) continue;
                    const double sc = bfull[cov[a]] / rs;
                    for (int b = 0; b < nc2; ++b) Mc[(size_t)a * nc2 + b] *= sc;
                }
                if (!invert(Mc, nc2)) {
                    singular = true;
                    continue;
                }
                for (int64_t p = eoff[e]; p < eoff[e + 1]; ++p) {
                    const double *b = &B[(size_t)epair[p] * Q];
                    for (int a = 0; a < nc2; ++a) {
                        double v = 0.0;
                        for (int c2 = 0; c2 < nc2; ++c2) v += Mc[(size_t)a * nc2 + c2] * b[cov[c2]];
                        if (v != 0.0) rows[(size_t)e * Q + cov[a]].push_back({(int32_t)ecell[p], v});
                    }
                }
            }
        };
        {
            const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> th;
            const int64_t chunk = (ne + nt - 1) / nt;
            for (unsigned t = 0; t < nt; ++t) {
                const int64_t a = t * chunk, b = std::min<int64_t>(ne, a + chunk);
                if (a < b) th.emplace_back(work, a, b);
            }
            for (auto &t : th) t.join();
        }
        CR_REQUIRE(!singular.load(), CR_DEGENERATE_GEOMETRY, "a spectral element's mass matrix is singular");
        const int64_t nodes = ne * Q;
        std::vector<int64_t> orp(nodes + 1, 0);
        for (int64_t r = 0; r < nodes; ++r) orp[r + 1] = orp[r] + (int64_t)rows[r].size();
        std::vector<int32_t> ocol(orp[nodes]);
        std::vector<double> oval(orp[nodes]);
        for (int64_t r = 0; r < nodes; ++r)
            for (size_t k = 0; k < rows[r].size(); ++k) {
                ocol[orp[r] + k] = rows[r][k].first;
                oval[orp[r] + k] = rows[r][k].second;
            }
        std::vector<double> ones(nodes, 1.0);
        cr_regridder *R = nullptr;
        CR_REQUIRE(cr_regridder_from_csr(&R, nodes, nc, orp[nodes], orp.data(), ocol.data(), oval.data(), ones.data(),
                                         a_cell.data(), stream) == CR_OK,
                   CR_INTERNAL, cr_last_error());
        R->stats.n_candidates = P->stats.n_candidates;
        *out = R;
    });
}

}  // extern "C"
