// gridgen.hpp -- grid vertex generators, compiled for host and device.
//
// They stand in for the grid packages the reference's adapters call:
//   lon-lat point matrices   Matrix{UnitSphericalPoint} -> CellBasedGrid
//                            (src/trees/interfaces.jl:125-126, grids.jl:57-85)
//   equiangular cubed sphere ext/ConservativeRegriddingClimaCoreExt.jl:29-94
//   HEALPix pixel boundaries ext/ConservativeRegriddingHealpixExt.jl:76-90, 154-169
//                            (Healpix.boundariesRing / nest2ring / ring2nest)
// HEALPix geometry follows Gorski et al. 2005, ApJ 622:759 (the projection of
// the 12 base faces, ring and nested numbering).
#pragma once

#include <cmath>
#include <cstdint>

#include "common.hpp"

namespace cr {

// sin and cos of an angle in degrees; multiples of 90 degrees are exact.  The
// quadrant is removed in degrees, the remainder converted with a two-term pi/180.
CR_HD void sincosd(double deg, double *s, double *c) {
    double q = rint(deg / 90.0);
    double r = deg - 90.0 * q;  // |r| <= 45
    const double k_hi = 1.7453292519943295e-02;  // pi/180 rounded
    const double k_lo = 2.9486522708701687e-19;  // pi/180 - k_hi
    double rad = r * k_hi + r * k_lo;
    double sr = sin(rad), cr_ = cos(rad);
    if (r == 0.0) { sr = 0.0; cr_ = 1.0; }
    long long qi = (long long)q;
    int m = (int)(((qi % 4) + 4) % 4);
    switch (m) {
        case 0: *s = sr; *c = cr_; break;
        case 1: *s = cr_; *c = -sr; break;
        case 2: *s = -sr; *c = -cr_; break;
        default: *s = -cr_; *c = sr; break;
    }
}

CR_HD void lonlat_to_xyz(double lon_deg, double lat_deg, double *p) {
    double sl, cl, sp, cp;
    sincosd(lon_deg, &sl, &cl);
    sincosd(lat_deg, &sp, &cp);
    p[0] = cp * cl;
    p[1] = cp * sl;
    p[2] = sp;
}

// ------------------------------------------------------------ cubed sphere --
// Face frames (normal n, along-i e1, along-j e2) with e1 x e2 = n, so cells
// (i,j),(i+1,j),(i+1,j+1),(i,j+1) run counter-clockwise seen from outside.
// Shared face edges come out bitwise identical (exact +-1 end tangents).
CR_HD void cube_face_frame(int f, double n[3], double e1[3], double e2[3]) {
    const double F[6][9] = {
        {1, 0, 0, 0, 1, 0, 0, 0, 1},    // +x
        {0, 1, 0, -1, 0, 0, 0, 0, 1},   // +y
        {-1, 0, 0, 0, -1, 0, 0, 0, 1},  // -x
        {0, -1, 0, 1, 0, 0, 0, 0, 1},   // -y
        {0, 0, 1, 0, 1, 0, -1, 0, 0},   // +z
        {0, 0, -1, 0, 1, 0, 1, 0, 0},   // -z
    };
    for (int k = 0; k < 3; ++k) { n[k] = F[f][k]; e1[k] = F[f][3 + k]; e2[k] = F[f][6 + k]; }
}

// tan of the equiangular coordinate -pi/4 + i*(pi/2)/ne, antisymmetric in i and
// exactly +-1 at the face edges.
CR_HD double equiangular_tan(int64_t i, int64_t ne) {
    if (2 * i == ne) return 0.0;
    if (i == 0) return -1.0;
    if (i == ne) return 1.0;
    bool neg = 2 * i < ne;
    int64_t k = neg ? ne - i : i;  // k > ne/2
    double a = (double)(2 * k - ne) / (double)ne * (M_PI / 4.0);
    double t = tan(a);
    return neg ? -t : t;
}

CR_HD void cube_point(int f, int64_t i, int64_t j, int64_t ne, double *p) {
    double n[3], e1[3], e2[3];
    cube_face_frame(f, n, e1, e2);
    double a = equiangular_tan(i, ne), b = equiangular_tan(j, ne);
    double v[3];
    for (int k = 0; k < 3; ++k) v[k] = n[k] + a * e1[k] + b * e2[k];
    double inv = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) p[k] = v[k] * inv;
}

// ------------------------------------------------------------------ HEALPix --
// Face f: row f/4 (north, equatorial, south), "ring offset" jr0 in {2,3,4} and
// longitude index jp in units of pi/4.
CR_HD int hpx_face_jr(int f) { return 2 + f / 4; }
CR_HD int hpx_face_jp(int f) {
    int row = f / 4, col = f % 4;
    return row == 1 ? 2 * col : 2 * col + 1;
}

CR_HD int64_t hpx_compact_bits(int64_t v) {  // even bits -> contiguous
    uint64_t x = (uint64_t)v & 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x >> 4)) & 0x00ff00ff00ff00ffull;
    x = (x | (x >> 8)) & 0x0000ffff0000ffffull;
    x = (x | (x >> 16)) & 0x00000000ffffffffull;
    return (int64_t)x;
}
CR_HD int64_t hpx_spread_bits(int64_t v) {  // contiguous -> even bits
    uint64_t x = (uint64_t)v & 0x00000000ffffffffull;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return (int64_t)x;
}

CR_HD void hpx_nest2fxy(int64_t nside, int64_t p, int *f, int64_t *ix, int64_t *iy) {
    int64_t n2 = nside * nside;
    *f = (int)(p / n2);
    int64_t q = p - (int64_t)(*f) * n2;
    *ix = hpx_compact_bits(q);
    *iy = hpx_compact_bits(q >> 1);
}
CR_HD int64_t hpx_fxy2nest(int64_t nside, int f, int64_t ix, int64_t iy) {
    return (int64_t)f * nside * nside + hpx_spread_bits(ix) + (hpx_spread_bits(iy) << 1);
}

CR_HD int64_t hpx_isqrt(int64_t v) {
    int64_t r = (int64_t)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// Ring index -> (face, ix, iy).  Ring i (1-based, north to south) holds pixels
// whose face coordinates satisfy ix + iy = jr0*nside - 1 - i; within a ring,
// pixels run east with growing ix - iy, starting at longitude 0.
CR_HD void hpx_ring2fxy(int64_t nside, int64_t r, int *f, int64_t *ix, int64_t *iy) {
    int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    if (r < ncap) {  // north polar cap: ring i has 4i pixels, starts at 2i(i-1)
        int64_t i = (1 + hpx_isqrt(1 + 2 * r)) / 2;
        int64_t j = r - 2 * i * (i - 1);
        *f = (int)(j / i);
        int64_t k = j % i;
        *ix = nside - i + k;
        *iy = nside - 1 - k;
        return;
    }
    if (r >= npix - ncap) {  // south polar cap, mirrored
        int64_t rs = npix - 1 - r;
        int64_t i = (1 + hpx_isqrt(1 + 2 * rs)) / 2;  // ring counted from the south pole
        int64_t j = r - (npix - 2 * i * (i + 1));
        *f = 8 + (int)(j / i);
        int64_t k = j % i;
        *ix = k;
        *iy = i - 1 - k;
        return;
    }
    // equatorial belt: rings nside..3nside, 4*nside pixels each
    int64_t rp = r - ncap;
    int64_t i = rp / (4 * nside) + nside;
    int64_t j = rp % (4 * nside);
    int64_t s = (i - nside + 1) & 1;  // 1: ring starts half a pixel east of 0
    // pixel centre longitude in units of pi/(2 nside): (jp*nside + ix - iy)/2 = j + 1 - s/2
    int64_t b0 = 2 * (j + 1) - s;  // = jp*nside + ix - iy (mod 8 nside)
    for (int ff = 0; ff < 12; ++ff) {
        int64_t a = hpx_face_jr(ff) * nside - 1 - i;  // ix + iy
        if (a < 0 || a > 2 * (nside - 1)) continue;
        for (int w = 0; w < 2; ++w) {
            int64_t b = b0 - hpx_face_jp(ff) * nside - (w ? 8 * nside : 0);
            if (((a + b) & 1) != 0) continue;
            int64_t x = (a + b) / 2, y = (a - b) / 2;
            if (x >= 0 && x < nside && y >= 0 && y < nside) {
                *f = ff; *ix = x; *iy = y;
                return;
            }
        }
    }
    *f = -1; *ix = -1; *iy = -1;  // unreachable for valid r
}

CR_HD int64_t hpx_fxy2ring(int64_t nside, int f, int64_t ix, int64_t iy) {
    int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    int64_t i = hpx_face_jr(f) * nside - 1 - ix - iy;  // ring number, 1-based
    if (i < nside) {  // north cap
        int64_t k = ix - (nside - i);
        return 2 * i * (i - 1) + (int64_t)f * i + k;
    }
    if (i > 3 * nside) {  // south cap
        int64_t is = 4 * nside - i;
        int64_t k = ix;
        return npix - 2 * is * (is + 1) + (int64_t)(f - 8) * is + k;
    }
    int64_t s = (i - nside + 1) & 1;
    int64_t b = hpx_face_jp(f) * nside + ix - iy;
    int64_t j = (b + s) / 2 - 1;
    int64_t m = 4 * nside;
    j = ((j % m) + m) % m;
    return ncap + (i - nside) * m + j;
}

// Unit vector of the face lattice point (u, v) / M, M = nside*step, on face f.
// z and phi come from integer ratios, so lattice points shared by pixels of one
// face are bitwise identical.
CR_HD void hpx_lattice_point(int f, int64_t u, int64_t v, int64_t M, double *p) {
    int jr0 = hpx_face_jr(f), jp = hpx_face_jp(f);
    int64_t K = (int64_t)jr0 * M - u - v;  // ring coordinate jr scaled by M
    double z, sth, phi4;  // phi4 = phi / (pi/4)
    if (K < M) {  // north polar cap: nr = K/M
        if (K == 0) { p[0] = 0.0; p[1] = 0.0; p[2] = 1.0; return; }
        double t = (double)(K * K) / (3.0 * (double)M * (double)M);  // 1 - z
        z = 1.0 - t;
        sth = sqrt(t * (2.0 - t));
        phi4 = (double)jp + (double)(u - v) / (double)K;
    } else if (K > 3 * M) {  // south polar cap: nr = (4M - K)/M
        int64_t L = 4 * M - K;
        if (L == 0) { p[0] = 0.0; p[1] = 0.0; p[2] = -1.0; return; }
        double t = (double)(L * L) / (3.0 * (double)M * (double)M);  // 1 + z
        z = t - 1.0;
        sth = sqrt(t * (2.0 - t));
        phi4 = (double)jp + (double)(u - v) / (double)L;
    } else {
        z = (double)(2 * (2 * M - K)) / (3.0 * (double)M);
        sth = sqrt((1.0 - z) * (1.0 + z));
        phi4 = (double)(jp * M + u - v) / (double)M;
    }
    if (phi4 < 0.0) phi4 += 8.0;
    if (phi4 >= 8.0) phi4 -= 8.0;
    double phi = phi4 * (M_PI / 4.0);
    p[0] = sth * cos(phi);
    p[1] = sth * sin(phi);
    p[2] = z;
}

// Boundary of pixel (f, ix, iy) with `step` points per side, counter-clockwise
// from the north corner: N -> W -> S -> E.
CR_HD void hpx_pixel_vertex(int f, int64_t ix, int64_t iy, int64_t nside, int step, int k,
                            double *p) {
    int64_t M = nside * step;
    int side = k / step, o = k % step;
    int64_t u, v;
    switch (side) {
        case 0: u = (ix + 1) * step - o; v = (iy + 1) * step; break;  // N -> W
        case 1: u = ix * step; v = (iy + 1) * step - o; break;        // W -> S
        case 2: u = ix * step + o; v = iy * step; break;              // S -> E
        default: u = (ix + 1) * step; v = iy * step + o; break;       // E -> N
    }
    hpx_lattice_point(f, u, v, M, p);
}

CR_HD void hpx_cell_fxy(int64_t nside, int order, int64_t idx, int *f, int64_t *ix, int64_t *iy) {
    if (order == CR_HEALPIX_RING) hpx_ring2fxy(nside, idx, f, ix, iy);
    else hpx_nest2fxy(nside, idx, f, ix, iy);
}

}  // namespace cr
