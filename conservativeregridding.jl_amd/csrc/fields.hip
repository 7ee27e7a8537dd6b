// fields.hip -- cell averages of the reference's example fields on the device.
//
// The reference's sweat tests fill a grid's field with the cell averages of an
// analytic function, int f dOmega / int dOmega over each cell, computed by
// SphericalPolygonIntegrator: fan triangles from the first vertex, each mapped
// from the reference triangle by x = p/|p|, p = l1 A + l2 B + l3 C, with
// Jacobian |det(A, B, C)| / |p|^3, summed with the degree-8 symmetric rule
// (lib/ConservativeRegriddingTestHelpers/src/ConservativeRegriddingTestHelpers.jl:84-128,
// set_field_values!).  The functions are those of src/utils/example_data.jl,
// evaluated at (lon, lat) in degrees, lon = atan2(y, x), lat = asin(z).  One
// thread per cell, FP64.
#include <cmath>

#include "regridder.hpp"
#include "sphgeom.hpp"
#include "trirule.hpp"

namespace cr {
namespace {

constexpr int kMaxFieldParams = 9;

struct FieldSpec {
    int kind;
    double p[kMaxFieldParams];
};

constexpr double kPi = 3.14159265358979323846;
constexpr double kDeg = kPi / 180.0;

// OASIS "fcos" term shared by the sinusoid and Gulf Stream fields.
__device__ double fcos(double coef, double dp_length, double lon, double lat) {
    return coef - cos(kPi * acos(cos(lon * kDeg) * cos(lat * kDeg)) / dp_length);
}

__device__ double example_field(const FieldSpec &F, double lon, double lat) {
    const double *p = F.p;
    switch (F.kind) {
    case CR_FIELD_LONGITUDE: return lon;
    case CR_FIELD_SINUSOID:  // dp_length, coef, coefmult
        return p[2] * fcos(p[1], p[0], lon, lat);
    case CR_FIELD_HARMONIC: {
        const double s = sin(2.0 * lat * kDeg), s2 = s * s, s4 = s2 * s2, s8 = s4 * s4;
        return 2.0 + s8 * s8 * cos(16.0 * lon * kDeg);
    }
    case CR_FIELD_GULF_STREAM: {  // dp_length, coef, gf_coef, origin, end, damping (lon, lat)
        const double dr0 = sqrt(((p[5] - p[3]) * kDeg) * ((p[5] - p[3]) * kDeg) +
                                ((p[6] - p[4]) * kDeg) * ((p[6] - p[4]) * kDeg));
        const double dr1 = sqrt(((p[7] - p[3]) * kDeg) * ((p[7] - p[3]) * kDeg) +
                                ((p[8] - p[4]) * kDeg) * ((p[8] - p[4]) * kDeg));
        double r = fcos(p[1], p[0], lon, lat);
        double plon = lon;
        if (plon > 180.0) plon -= 360.0;
        else if (plon < -180.0) plon += 360.0;
        const double dx = (plon - p[3]) * kDeg, dy = (lat - p[4]) * kDeg;
        const double dr = sqrt(dx * dx + dy * dy), dth = atan2(dy, dx);
        double dc = 1.3 * p[2];
        if (dr > dr0) dc = 0.0;
        else if (dr > dr1) dc *= cos(kPi * 0.5 * (dr - dr1) / (dr0 - dr1));
        r += (fmax(1000.0 * sin(0.4 * (0.5 * dr + dth) + 0.007 * cos(50.0 * dth) + 0.37 * kPi), 999.0) - 999.0) * dc;
        return r;
    }
    default: {  // CR_FIELD_VORTEX: lon0_rad, lat0_rad, r0, d, t
        const double sc = sin(p[1]), cc = cos(p[1]);
        const double sl = sin(lat * kDeg), cl = cos(lat * kDeg);
        const double trm = cl * cos(lon * kDeg - p[0]);
        const double X = sc * trm - cc * sl, Y = cl * sin(lon * kDeg - p[0]), Z = sc * sl + cc * trm;
        double dlon = atan2(Y, X);
        if (dlon < 0.0) dlon += 2.0 * kPi;
        const double dlat = asin(fmin(1.0, fmax(-1.0, Z)));
        const double rho = p[2] * cos(dlat);
        const double ch = cosh(rho);
        const double vt = 3.0 * sqrt(3.0) / 2.0 / (ch * ch) * tanh(rho);
        const double omega = rho == 0.0 ? 0.0 : vt / rho;
        return 2.0 * (1.0 + tanh(rho / p[3] * sin(dlon - omega * p[4])));
    }
    }
}

__global__ void k_cell_averages(const double *__restrict__ cells, int nv, int64_t ncells, TriRule rule, FieldSpec F,
                                double *__restrict__ out) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const double *v = cells + 3 * (int64_t)nv * c;
    const d3 A = load3(v);
    double total = 0.0, area = 0.0;
    for (int i = 1; i + 1 < nv; ++i) {
        const d3 B = load3(v + 3 * i), C = load3(v + 3 * (i + 1));
        const double det = fabs(dot(A, cross(B, C)));
        if (det == 0.0) continue;  // repeated padding vertex: an empty triangle
        for (int k = 0; k < rule.n; ++k) {
            const d3 q = rule.l1[k] * A + rule.l2[k] * B + rule.l3[k] * C;
            const double nq = sqrt(dot(q, q));
            const d3 s = (1.0 / nq) * q;
            const double J = det / (nq * nq * nq);
            const double lon = atan2(s.y, s.x) / kDeg, lat = asin(fmin(1.0, fmax(-1.0, s.z))) / kDeg;
            total += rule.w[k] * example_field(F, lon, lat) * J;
            area += rule.w[k] * J;
        }
    }
    out[c] = area == 0.0 ? 0.0 : total / area;  // zero-area cells (ghosts) get 0, not 0/0
}

// The reference's defaults, in its keyword order.
int default_params(int kind, double *p) {
    switch (kind) {
    case CR_FIELD_SINUSOID:
        p[0] = 1.2 * kPi; p[1] = 2.0; p[2] = 1.0;
        return 3;
    case CR_FIELD_GULF_STREAM:
        p[0] = 1.2 * kPi; p[1] = 2.0; p[2] = 1.0; p[3] = -80.0; p[4] = 25.0; p[5] = -1.8; p[6] = 50.0;
        p[7] = -25.5; p[8] = 55.5;
        return 9;
    case CR_FIELD_VORTEX:
        p[0] = 5.5; p[1] = 0.2; p[2] = 3.0; p[3] = 5.0; p[4] = 6.0;
        return 5;
    default: return 0;
    }
}

}  // namespace
}  // namespace cr

using namespace cr;

extern "C" {

cr_status cr_cell_averages(const cr_grid *grid, int32_t field, const double *params, int32_t nparams,
                           int32_t degree, double *out, void *stream) {
    return guard([&] {
        CR_REQUIRE(grid && out, CR_INVALID_ARGUMENT, "cr_cell_averages: null argument");
        CR_REQUIRE(!grid->planar, CR_UNSUPPORTED, "cr_cell_averages: the example fields are functions on the sphere");
        CR_REQUIRE(field >= CR_FIELD_LONGITUDE && field <= CR_FIELD_VORTEX, CR_INVALID_ARGUMENT,
                   "cr_cell_averages: unknown example field " + std::to_string(field));
        FieldSpec F{};
        F.kind = field;
        const int want = default_params(field, F.p);
        if (params) {
            CR_REQUIRE(nparams == want, CR_INVALID_ARGUMENT,
                       "cr_cell_averages: this field takes " + std::to_string(want) + " parameters, got " +
                           std::to_string(nparams));
            for (int i = 0; i < want; ++i) F.p[i] = params[i];
        }
        const TriRule rule = tri_rule(degree);
        hipStream_t s = as_stream(stream);
        if (grid->ncells)
            hipLaunchKernelGGL(k_cell_averages, dim3(grid_for(grid->ncells, 128)), dim3(128), 0, s, grid->cells.p,
                               grid->nv, grid->ncells, rule, F, out);
        CR_HIP(hipGetLastError());
    });
}

}  // extern "C"
