// regridder.hpp -- the device-resident Regridder handle.
//
// Mirrors the fields of the reference struct (src/regridder/regridder.jl:25-36):
//   intersections  -> CSR(A) (rows = dst cells) and CSR(A^T) (rows = src cells);
//                     the transpose is materialised once so transpose(R) stays a
//                     flag and both directions are gather SpMVs without atomics
//   dst_areas      -> a_dst
//   src_areas      -> a_src
//   dst_temp/src_temp (work arrays for strided fields) have no device
//   counterpart: cr_regrid takes strides.
#pragma once

#include "common.hpp"

namespace cr {

struct Csr {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    DevBuf<int64_t> rowptr;
    DevBuf<int32_t> col;
    DevBuf<double> val;
    DevBuf<double> inv_area;   // 1 / area of the row cells (FP32 batched path)
    DevBuf<int32_t> long_rows; // rows with more than kLongRow non-zeros (own kernel)
    int64_t nlong = 0;
    int64_t max_row_nnz = 0;
    int tuned_variant = -1;    // FP32 field-block width chosen by cr_regrid_tune; -1: default
};

}  // namespace cr

struct cr_regridder {
    cr::Csr fwd;  // A:   n_dst x n_src
    cr::Csr bwd;  // A^T: n_src x n_dst
    cr::DevBuf<double> a_dst, a_src;
    cr_build_stats stats{};
    // stored areas = area_scale * unit-sphere areas (radius^2, divided by max(A)
    // after normalize!); 0 when built from arrays (no geometry: no vertex gradients)
    double area_scale = 0.0;
};

namespace cr {
// Finish a regridder whose CSR arrays and areas are filled: inverse areas and
// apply plans.
void finalize_regridder(cr_regridder *R, hipStream_t s);
}  // namespace cr
