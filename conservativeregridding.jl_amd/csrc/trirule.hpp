// Symmetric quadrature rules on the reference triangle of area 1/2 (Dunavant
// 1985, degrees 1-8), as barycentric points (l1, l2, l3) and weights summing to
// 1/2: the rules of TriangleQuadrature.reference_rule
// (src/regridder/triangle_quadrature.jl:139-156).  Defined in se.hip.
#pragma once

#include "common.hpp"

namespace cr {

constexpr int kSeMaxRule = 16;

struct TriRule {
    int n;
    double l1[kSeMaxRule], l2[kSeMaxRule], l3[kSeMaxRule], w[kSeMaxRule];
};

// The rule of this degree, or of the next degree that exists (CR_UNSUPPORTED past 8).
TriRule tri_rule(int degree);

}  // namespace cr
