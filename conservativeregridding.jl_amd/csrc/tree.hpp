// tree.hpp -- bounding caps and the implicit source tree of the candidate search
// (host + device; the kernels in build.hip are thin wrappers).
//
// Replaces the reference's dual-tree DFS over spherical caps
// (src/utils/MultithreadedDualDepthFirstSearch.jl:12-65,
//  src/trees/grids.jl:194-287 caps, src/trees/quadtree_cursors.jl:214-316).
// A cap is a ball (centre c, chord radius r) in R^3: for a cell inside a
// hemisphere, the ball around the normalised vertex mean that holds every
// vertex holds the whole cell (the spherical cap it cuts is geodesically
// convex).  Two cells can only overlap if |c1 - c2| <= r1 + r2.
#pragma once

#include "sphgeom.hpp"

namespace cr {

constexpr int kTreeFanout = 8;
constexpr int kMaxLevels = 24;
constexpr float kCapAll = 4.0f;  // radius that overlaps everything (chords are <= 2)

struct Levels {
    int64_t off[kMaxLevels];
    int64_t size[kMaxLevels];
    int top;
};

// FP32 copy of a cap, padded outward to cover rounding of the centre and radius.
CR_HD float4 pad_cap(const d3 &m, double r) {
    float rf = r >= 1.4 ? kCapAll : (float)(r * (1.0 + 1e-6) + 2e-6);
    return make_float4((float)m.x, (float)m.y, (float)m.z, rf);
}

CR_HD float4 cell_cap(const double *p, int nv) {
    d3 s = mk3(0, 0, 0);
    for (int k = 0; k < nv; ++k) s = s + load3(p + 3 * k);
    double L = sqrt(dot(s, s));
    if (L < 1e-12) return make_float4(0.f, 0.f, 1.f, kCapAll);
    d3 m = (1.0 / L) * s;
    double r = 0.0;
    for (int k = 0; k < nv; ++k) {
        d3 d = load3(p + 3 * k) - m;
        r = fmax(r, sqrt(dot(d, d)));
    }
    return pad_cap(m, r);
}

// Space-filling-curve key of a cap centre: the cube face the centre projects
// to (3 bits) and the Hilbert index of its gnomonic (u, v) on that face at
// 2^kSfcOrder x 2^kSfcOrder resolution (28 bits).  Runs of consecutive keys
// are compact patches of the sphere (no long jumps inside a face), which keeps
// tree nodes of 8 consecutive cells and search clusters of 64 cells small.
constexpr int kSfcOrder = 14;

CR_HD uint32_t hilbert_d(uint32_t x, uint32_t y) {
    uint32_t d = 0;
    for (uint32_t s = 1u << (kSfcOrder - 1); s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {  // rotate the quadrant
            if (rx == 1) {
                x = s - 1 - (x & (s - 1)) + (x & ~(s - 1));
                y = s - 1 - (y & (s - 1)) + (y & ~(s - 1));
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

CR_HD uint32_t sfc_key(const float4 &q) {
    const float ax = fabsf(q.x), ay = fabsf(q.y), az = fabsf(q.z);
    int f;
    float m, u, v;
    if (ax >= ay && ax >= az) { f = 0; m = q.x; u = q.y; v = q.z; }
    else if (ay >= az) { f = 1; m = q.y; u = q.z; v = q.x; }
    else { f = 2; m = q.z; u = q.x; v = q.y; }
    const uint32_t face = (uint32_t)(2 * f + (m < 0.0f ? 1 : 0));
    const float am = fabsf(m) > 0.0f ? fabsf(m) : 1.0f;
    const float n = (float)(1u << kSfcOrder);
    int iu = (int)((u / am + 1.0f) * 0.5f * n), iv = (int)((v / am + 1.0f) * 0.5f * n);
    const int top = (1 << kSfcOrder) - 1;
    iu = iu < 0 ? 0 : (iu > top ? top : iu);
    iv = iv < 0 ? 0 : (iv > top ? top : iv);
    return (face << (2 * kSfcOrder)) | hilbert_d((uint32_t)iu, (uint32_t)iv);
}

// Parent cap of children [c0, c1): ball around the normalised mean of the child
// centres holding every child ball (triangle inequality on chords).
CR_HD float4 parent_cap(const float4 *child, int64_t c0, int64_t c1) {
    d3 s = mk3(0, 0, 0);
    bool all = false;
    for (int64_t c = c0; c < c1; ++c) {
        float4 q = child[c];
        s = s + mk3(q.x, q.y, q.z);
        all |= q.w >= 2.0f;
    }
    double L = sqrt(dot(s, s));
    if (all || L < 1e-9) return make_float4(0.f, 0.f, 1.f, kCapAll);
    d3 m = (1.0 / L) * s;
    double r = 0.0;
    for (int64_t c = c0; c < c1; ++c) {
        float4 q = child[c];
        d3 d = mk3(q.x, q.y, q.z) - m;
        r = fmax(r, sqrt(dot(d, d)) + (double)q.w);
    }
    return pad_cap(m, r >= 2.0 ? 4.0 : r);
}

CR_HD bool caps_overlap(const float4 &a, const float4 &b) {
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    float rr = a.w + b.w;
    return dx * dx + dy * dy + dz * dz <= rr * rr;
}

// Level sizes of the implicit kTreeFanout-ary tree over n leaves (level 0).
inline Levels make_levels(int64_t n) {
    Levels lv{};
    int64_t off = 0;
    int L = 0;
    for (;;) {
        lv.off[L] = off;
        lv.size[L] = n;
        off += n;
        if (n <= 1 || L + 1 == kMaxLevels) break;
        n = (n + kTreeFanout - 1) / kTreeFanout;
        ++L;
    }
    lv.top = L;
    return lv;
}

// Stackless depth-first walk: visit every leaf whose cap overlaps q, calling
// emit(leaf index).  Node (L, i) has children [i*F, min((i+1)*F, size[L-1])) on
// level L-1; after a node, go to its next sibling or climb to the first
// ancestor that has one.
template <class Emit>
CR_HD void tree_query(const float4 &q, const float4 *nodes, const Levels &lv, Emit &&emit) {
    int L = lv.top;
    int64_t i = 0;
    for (;;) {
        if (caps_overlap(q, nodes[lv.off[L] + i])) {
            if (L == 0) {
                emit(i);
            } else {
                --L;
                i *= kTreeFanout;
                continue;
            }
        }
        for (;;) {
            if (L == lv.top) return;
            int64_t nxt = i + 1;
            if ((nxt % kTreeFanout) != 0 && nxt < lv.size[L]) {
                i = nxt;
                break;
            }
            i /= kTreeFanout;
            ++L;
        }
    }
}

}  // namespace cr
