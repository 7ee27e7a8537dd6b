// apply.hip -- regrid!(dst, R, src) and regrid!(s, transpose(R), d) on the device.
//
// Reference: perform_regridding! = mul!(dst, R.intersections, src) followed by
// finalize_regridding! (dst ./= R.dst_areas), src/regridder/regrid.jl:95-118;
// N-D fields loop 1-D slices (NDSliceLoop, regrid.jl:303-318).  Here a single
// launch handles every field.  A is sparse (about 5 non-zeros per row), so the
// apply is bound by HBM bandwidth, not by arithmetic: no MFMA.
//
// Kernels
//   k_spmv_f64<G>         FP64 fields, G lanes per row, row gather (CSR-vector)
//   k_spmm_f32_cells<G,R> FP32 fields with cells fastest ([n, k] column-major,
//                         what regrid! receives with dims=1): G lanes per row keep
//                         the row's (col, val) in registers and sweep the fields
//   k_spmm_f32_fields     FP32 fields with fields fastest: one wave per row,
//                         lanes across fields, fully coalesced loads
// FP32 kernels form each product in FP64 (the FP64 A times the widened field
// value), round it to FP32 and accumulate with compensated (Kahan-Babuska)
// summation; the row is scaled by the FP64 reciprocal area.
#include <algorithm>

#include "regridder.hpp"

namespace cr {

// ---- FP32 compensated sums ------------------------------------------------
struct CompF {
    float s, c;
};
__device__ __forceinline__ void comp_add(CompF &a, float p) {
    float t = a.s + p;
    float bp = t - a.s;
    float e = (a.s - (t - bp)) + (p - bp);
    a.s = t;
    a.c += e;
}
__device__ __forceinline__ CompF comp_merge(CompF a, CompF b) {
    comp_add(a, b.s);
    a.c += b.c;
    return a;
}

template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
    return v;
}

template <int G>
__device__ __forceinline__ CompF group_sum(CompF v) {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) {
        CompF o{__shfl_xor(v.s, m, G), __shfl_xor(v.c, m, G)};
        v = comp_merge(v, o);
    }
    return v;
}

// ---- FP64 ------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(256) k_spmv_f64(const int64_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ col,
                                                  const double *__restrict__ val, int64_t nrows,
                                                  const double *__restrict__ x, int64_t xc, int64_t xf,
                                                  double *__restrict__ y, int64_t yc, int64_t yf,
                                                  int64_t nfields, const double *__restrict__ area) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t row = t / G;
    int lane = (int)(t % G);
    if (row >= nrows) return;  // whole groups exit together (G divides 64)
    int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
    for (int64_t f = 0; f < nfields; ++f) {
        const double *xs = x + f * xf;
        double s = 0.0;
        for (int64_t k = k0 + lane; k < k1; k += G) s += val[k] * xs[(int64_t)col[k] * xc];
        s = group_sum<G>(s);
        if (lane == 0) y[row * yc + f * yf] = area ? s / area[row] : s;
    }
}

// ---- FP32, cells fastest -----------------------------------------------------
// G lanes per row (G = 1: one thread per row, no cross-lane reduction).  Rows
// with at most G*R non-zeros keep their (offset, value) pairs in registers and
// sweep the fields FB at a time, so FB*R independent gathers are in flight per
// lane; longer rows stream (col, val) from cache for every field batch.
template <int G, int R, int FB>
__global__ void __launch_bounds__(256) k_spmm_f32_cells(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ val, int64_t nrows, const float *__restrict__ x, int64_t xc,
    int64_t xf, float *__restrict__ y, int64_t yc, int64_t yf, int64_t nfields,
    const double *__restrict__ inv_area) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t row = t / G;
    int lane = (int)(t % G);
    if (row >= nrows) return;
    int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
    double scale = inv_area ? inv_area[row] : 1.0;
    float *yr = y + row * yc;
    if (k1 - k0 <= (int64_t)G * R) {
        int64_t off[R];
        double v[R];
        const int nmine = (int)((k1 - k0 - lane + G - 1) / G);  // entries this lane holds
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int64_t k = k0 + lane + (int64_t)r * G;
            bool ok = r < nmine;
            off[r] = ok ? (int64_t)col[k] * xc : 0;
            v[r] = ok ? val[k] : 0.0;
        }
        for (int64_t f = 0; f < nfields; f += FB) {
            CompF acc[FB];
#pragma unroll
            for (int b = 0; b < FB; ++b) acc[b] = CompF{0.f, 0.f};
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r < nmine) {
#pragma unroll
                    for (int b = 0; b < FB; ++b)
                        if (FB == 1 || f + b < nfields)
                            comp_add(acc[b], (float)(v[r] * (double)x[off[r] + (f + b) * xf]));
                }
            }
#pragma unroll
            for (int b = 0; b < FB; ++b) {
                CompF a = group_sum<G>(acc[b]);
                if (lane == 0 && (FB == 1 || f + b < nfields))
                    yr[(f + b) * yf] = (float)(((double)a.s + (double)a.c) * scale);
            }
        }
    } else {
        for (int64_t f = 0; f < nfields; ++f) {
            const float *xs = x + f * xf;
            CompF acc{0.f, 0.f};
            for (int64_t k = k0 + lane; k < k1; k += G)
                comp_add(acc, (float)(val[k] * (double)xs[(int64_t)col[k] * xc]));
            acc = group_sum<G>(acc);
            if (lane == 0) yr[f * yf] = (float)(((double)acc.s + (double)acc.c) * scale);
        }
    }
}

// ---- FP32, fields fastest (field stride 1) -----------------------------------
// One wave per row; lane l owns fields l, l+64, l+128, ... (up to kFieldRegs).
constexpr int kFieldRegs = 8;

__global__ void __launch_bounds__(256) k_spmm_f32_fields(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ val, int64_t nrows, const float *__restrict__ x, int64_t xc,
    float *__restrict__ y, int64_t yc, int64_t nfields, const double *__restrict__ inv_area) {
    int64_t row = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (row >= nrows) return;
    int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
    double scale = inv_area ? inv_area[row] : 1.0;
    for (int64_t f0 = 0; f0 < nfields; f0 += 64 * kFieldRegs) {
        CompF acc[kFieldRegs];
#pragma unroll
        for (int j = 0; j < kFieldRegs; ++j) acc[j] = CompF{0.f, 0.f};
        for (int64_t k = k0; k < k1; ++k) {
            const float *xr = x + (int64_t)col[k] * xc;
            double v = val[k];
#pragma unroll
            for (int j = 0; j < kFieldRegs; ++j) {
                int64_t f = f0 + lane + 64 * j;
                if (f < nfields) comp_add(acc[j], (float)(v * (double)xr[f]));
            }
        }
#pragma unroll
        for (int j = 0; j < kFieldRegs; ++j) {
            int64_t f = f0 + lane + 64 * j;
            if (f < nfields) y[row * yc + f] = (float)(((double)acc[j].s + (double)acc[j].c) * scale);
        }
    }
}

__global__ void k_inv(const double *__restrict__ a, int64_t n, double *__restrict__ inv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[i] = 1.0 / a[i];
}

void finalize_regridder(cr_regridder *R, hipStream_t s) {
    R->fwd.inv_area.alloc(R->fwd.nrows);
    R->bwd.inv_area.alloc(R->bwd.nrows);
    hipLaunchKernelGGL(k_inv, dim3(grid_for(R->fwd.nrows, 256)), dim3(256), 0, s, R->a_dst.p,
                       R->fwd.nrows, R->fwd.inv_area.p);
    hipLaunchKernelGGL(k_inv, dim3(grid_for(R->bwd.nrows, 256)), dim3(256), 0, s, R->a_src.p,
                       R->bwd.nrows, R->bwd.inv_area.p);
    CR_HIP(hipGetLastError());
}

// Apply variants of the FP32 cells-fastest kernel (the autotuner's search space).
struct CellsVariant {
    int G, R, FB;
};
constexpr CellsVariant kCellsVariants[] = {{1, 8, 1}, {1, 8, 2}, {1, 8, 4}, {2, 4, 2}, {4, 4, 1}, {4, 4, 4}};
constexpr int kNumCellsVariants = sizeof(kCellsVariants) / sizeof(kCellsVariants[0]);
constexpr int kDefaultCellsVariant = 1;

template <int G, int R, int FB>
void launch_cells(const Csr &A, const float *x, int64_t sc, int64_t sf, float *y, int64_t dc,
                  int64_t df, int64_t nf, const double *inv, hipStream_t s) {
    const int B = 256;
    hipLaunchKernelGGL((k_spmm_f32_cells<G, R, FB>), dim3(grid_for(A.nrows * G, B)), dim3(B), 0, s,
                       A.rowptr.p, A.col.p, A.val.p, A.nrows, x, sc, sf, y, dc, df, nf, inv);
}

void launch_cells_variant(int v, const Csr &A, const float *x, int64_t sc, int64_t sf, float *y,
                          int64_t dc, int64_t df, int64_t nf, const double *inv, hipStream_t s) {
    switch (v) {
        case 0: launch_cells<1, 8, 1>(A, x, sc, sf, y, dc, df, nf, inv, s); break;
        case 1: launch_cells<1, 8, 2>(A, x, sc, sf, y, dc, df, nf, inv, s); break;
        case 2: launch_cells<1, 8, 4>(A, x, sc, sf, y, dc, df, nf, inv, s); break;
        case 3: launch_cells<2, 4, 2>(A, x, sc, sf, y, dc, df, nf, inv, s); break;
        case 4: launch_cells<4, 4, 1>(A, x, sc, sf, y, dc, df, nf, inv, s); break;
        default: launch_cells<4, 4, 4>(A, x, sc, sf, y, dc, df, nf, inv, s); break;
    }
}

// Kernel choice for a call: FP64 -> CSR-vector; FP32 with unit field strides
// and several fields -> fields-fastest kernel; otherwise the cells-fastest
// variant the autotuner picked for this direction (or the default).
void apply(const cr_regridder *R, bool transpose, int dtype, int64_t nf, const void *src,
           int64_t sc, int64_t sf, void *dst, int64_t dc, int64_t df, bool normalize, hipStream_t s,
           int force_variant = -1) {
    const Csr &A = transpose ? R->bwd : R->fwd;
    const double *area = transpose ? R->a_src.p : R->a_dst.p;
    if (A.nrows == 0 || nf == 0) return;
    const int B = 256;
    if (dtype == CR_F64) {
        constexpr int G = 4;
        hipLaunchKernelGGL(k_spmv_f64<G>, dim3(grid_for(A.nrows * G, B)), dim3(B), 0, s, A.rowptr.p,
                           A.col.p, A.val.p, A.nrows, (const double *)src, sc, sf, (double *)dst, dc, df,
                           nf, normalize ? area : nullptr);
    } else if (sf == 1 && df == 1 && nf > 1) {
        hipLaunchKernelGGL(k_spmm_f32_fields, dim3(grid_for(A.nrows, B / 64)), dim3(B), 0, s, A.rowptr.p,
                           A.col.p, A.val.p, A.nrows, (const float *)src, sc, (float *)dst, dc, nf,
                           normalize ? A.inv_area.p : nullptr);
    } else {
        int v = force_variant >= 0 ? force_variant : A.tuned_cells_variant;
        if (v < 0 || v >= kNumCellsVariants) v = kDefaultCellsVariant;
        launch_cells_variant(v, A, (const float *)src, sc, sf, (float *)dst, dc, df, nf,
                             normalize ? A.inv_area.p : nullptr, s);
    }
    CR_HIP(hipGetLastError());
}

}  // namespace cr

using namespace cr;

extern "C" {

cr_status cr_regrid(const cr_regridder *R, int32_t transpose, int32_t dtype, int64_t nfields,
                    const void *src, int64_t src_cell_stride, int64_t src_field_stride, void *dst,
                    int64_t dst_cell_stride, int64_t dst_field_stride, int32_t normalize,
                    void *stream) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "cr_regrid: null regridder");
        CR_REQUIRE(dtype == CR_F64 || dtype == CR_F32, CR_UNSUPPORTED, "cr_regrid: dtype must be F64 or F32");
        CR_REQUIRE(nfields >= 0, CR_INVALID_ARGUMENT, "cr_regrid: nfields < 0");
        CR_REQUIRE(src_cell_stride > 0 && dst_cell_stride > 0 && src_field_stride >= 0 &&
                       dst_field_stride >= 0,
                   CR_INVALID_ARGUMENT, "cr_regrid: strides must be positive");
        CR_REQUIRE(nfields <= 1 || (src_field_stride > 0 && dst_field_stride > 0), CR_INVALID_ARGUMENT,
                   "cr_regrid: field strides must be positive for several fields");
        CR_REQUIRE(nfields == 0 || (src && dst), CR_INVALID_ARGUMENT, "cr_regrid: null field pointer");
        apply(R, transpose != 0, dtype, nfields, src, src_cell_stride, src_field_stride, dst,
              dst_cell_stride, dst_field_stride, normalize != 0, as_stream(stream));
    });
}

cr_status cr_regrid_tune(cr_regridder *R, int32_t transpose, int64_t nfields, const void *src,
                         int64_t src_cell_stride, int64_t src_field_stride, void *dst,
                         int64_t dst_cell_stride, int64_t dst_field_stride, int32_t *chosen,
                         float *best_ms, void *stream) {
    return guard([&] {
        CR_REQUIRE(R && src && dst && nfields > 0, CR_INVALID_ARGUMENT, "cr_regrid_tune: bad arguments");
        CR_REQUIRE(src_cell_stride > 0 && dst_cell_stride > 0 && src_field_stride > 0 &&
                       dst_field_stride > 0,
                   CR_INVALID_ARGUMENT, "cr_regrid_tune: strides must be positive");
        hipStream_t s = as_stream(stream);
        Csr &A = transpose ? R->bwd : R->fwd;
        hipEvent_t e0, e1;
        CR_HIP(hipEventCreate(&e0));
        CR_HIP(hipEventCreate(&e1));
        int best = -1;
        float best_t = 1e30f;
        for (int v = 0; v < kNumCellsVariants; ++v) {
            // one warm-up launch, then the median of three timed launches
            apply(R, transpose != 0, CR_F32, nfields, src, src_cell_stride, src_field_stride, dst,
                  dst_cell_stride, dst_field_stride, true, s, v);
            float t[3];
            for (int rep = 0; rep < 3; ++rep) {
                CR_HIP(hipEventRecord(e0, s));
                apply(R, transpose != 0, CR_F32, nfields, src, src_cell_stride, src_field_stride, dst,
                      dst_cell_stride, dst_field_stride, true, s, v);
                CR_HIP(hipEventRecord(e1, s));
                CR_HIP(hipEventSynchronize(e1));
                CR_HIP(hipEventElapsedTime(&t[rep], e0, e1));
            }
            std::sort(t, t + 3);
            if (t[1] < best_t) { best_t = t[1]; best = v; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        A.tuned_cells_variant = best;
        if (chosen) *chosen = best;
        if (best_ms) *best_ms = best_t;
    });
}

}  // extern "C"
