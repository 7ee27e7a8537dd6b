// apply.hip -- regrid!(dst, R, src) and regrid!(s, transpose(R), d) on the device.
//
// Reference: perform_regridding! = mul!(dst, R.intersections, src) followed by
// finalize_regridding! (dst ./= R.dst_areas), src/regridder/regrid.jl:95-118;
// N-D fields loop 1-D slices (NDSliceLoop, regrid.jl:303-318).  Here a single
// launch handles every field.  A is sparse (about 5 non-zeros per row, ~0.25
// flop per byte), so the apply is bound by HBM bandwidth: no MFMA.
//
// Kernels
//   k_spmm_tiles<T,F>   (tiles.hip) staged merge-path SpMM over row tiles:
//                       distinct source cells of a tile gathered once per field
//                       into LDS, non-zeros split evenly over threads
//   k_apply_rows<T,FB>  one thread per row; blockIdx.y = block of FB fields,
//                       the slowest launch dimension; XCD-contiguous row tiles
//   k_apply_long<T,FB>  rows with more than kLongRow non-zeros, one wave each
//   k_apply_fieldsfast  FP32 fields with field stride 1: one wave per row,
//                       lanes across fields, fully coalesced
// FP64 fields accumulate in FP64 (fused multiply-add) and divide by the area
// like finalize_regridding!.  FP32 fields use the FP64 A rounded to FP32 once
// per non-zero and field block, FP32 products, FP32 Kahan-compensated row
// sums, and a final scale by the FP64 reciprocal area.
#include <algorithm>
#include <vector>

#include "regridder.hpp"

namespace cr {

constexpr int kApplyBlock = 256;
constexpr int kLongRow = 32;

// ---- accumulators ------------------------------------------------------------
template <class T>
struct Acc;

template <>
struct Acc<double> {
    double s;
    __device__ void init() { s = 0.0; }
    __device__ void add(double v, double x) { s = fma(v, x, s); }
    __device__ void merge_from_lane(int m) { s += __shfl_xor(s, m, 64); }
    __device__ double finish(const double *area, const double *, int64_t row) const {
        return area ? s / area[row] : s;
    }
};

template <>
struct Acc<float> {
    float s, c;  // Kahan: the exact sum is s - c
    __device__ void init() {
        s = 0.f;
        c = 0.f;
    }
    __device__ void add(float v, float x) {
        float yk = v * x - c;
        float t = s + yk;
        c = (t - s) - yk;
        s = t;
    }
    __device__ void merge_from_lane(int m) {
        float os = __shfl_xor(s, m, 64), oc = __shfl_xor(c, m, 64);
        // (s - c) + (os - oc) with the rounding error of s + os kept in c
        float t = s + os;
        float bp = t - s;
        float e = (s - (t - bp)) + (os - bp);
        s = t;
        c = (c + oc) - e;
    }
    __device__ float finish(const double *, const double *inv, int64_t row) const {
        double v = (double)s - (double)c;
        return (float)(inv ? v * inv[row] : v);
    }
};

template <class T>
__device__ __forceinline__ T a_value(double v);
template <>
__device__ __forceinline__ double a_value<double>(double v) { return v; }
template <>
__device__ __forceinline__ float a_value<float>(double v) { return (float)v; }

// ---- row kernel ------------------------------------------------------------------
// One thread per row; blockIdx.y selects FB fields and is the slowest launch
// dimension, so the whole GPU sweeps the same few fields at a time (their
// slices of x stay in L2/MALL, few pages are live) while A is re-read from
// cache once per field block.  Workgroups go round-robin to the 8 XCDs, so
// tile = (bx % 8) * tiles_per_xcd + bx / 8 gives each XCD's L2 a contiguous
// range of rows, whose source cells overlap.  Rows longer than kLongRow are
// skipped here and done by k_apply_long.
template <class T, int FB>
__global__ void __launch_bounds__(kApplyBlock) k_apply_rows(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ val, int64_t nrows, const T *__restrict__ x, int64_t xc, int64_t xf,
    T *__restrict__ y, int64_t yc, int64_t yf, int64_t nfields, const double *__restrict__ area,
    const double *__restrict__ inv_area, int64_t tiles_per_xcd) {
    const int64_t bx = blockIdx.x;
    const int64_t tile = (bx & 7) * tiles_per_xcd + (bx >> 3);
    const int64_t row = tile * kApplyBlock + threadIdx.x;
    if (row >= nrows) return;
    const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
    if (k1 - k0 > kLongRow) return;
    const int64_t f0 = (int64_t)blockIdx.y * FB;
    const int nb = (int)min((int64_t)FB, nfields - f0);
    Acc<T> acc[FB];
#pragma unroll
    for (int b = 0; b < FB; ++b) acc[b].init();
    const T *xb = x + f0 * xf;
    for (int64_t k = k0; k < k1; ++k) {
        const T *xp = xb + (int64_t)col[k] * xc;
        const T v = a_value<T>(val[k]);
#pragma unroll
        for (int b = 0; b < FB; ++b)
            if (FB == 1 || b < nb) acc[b].add(v, xp[b * xf]);
    }
    T *yp = y + row * yc + f0 * yf;
#pragma unroll
    for (int b = 0; b < FB; ++b)
        if (FB == 1 || b < nb) yp[b * yf] = acc[b].finish(area, inv_area, row);
}

// Rows longer than kLongRow: one wave per (row, field block); lanes split the
// row; fixed-order butterfly reduction (deterministic).
template <class T, int FB>
__global__ void __launch_bounds__(64) k_apply_long(
    const int32_t *__restrict__ long_rows, const int64_t *__restrict__ rowptr,
    const int32_t *__restrict__ col, const double *__restrict__ val, const T *__restrict__ x,
    int64_t xc, int64_t xf, T *__restrict__ y, int64_t yc, int64_t yf, int64_t nfields,
    const double *__restrict__ area, const double *__restrict__ inv_area) {
    const int64_t row = long_rows[blockIdx.x];
    const int lane = threadIdx.x;
    const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
    const int64_t f0 = (int64_t)blockIdx.y * FB;
    const int nb = (int)min((int64_t)FB, nfields - f0);
    Acc<T> acc[FB];
#pragma unroll
    for (int b = 0; b < FB; ++b) acc[b].init();
    const T *xb = x + f0 * xf;
    for (int64_t k = k0 + lane; k < k1; k += 64) {
        const T *xp = xb + (int64_t)col[k] * xc;
        const T v = a_value<T>(val[k]);
#pragma unroll
        for (int b = 0; b < FB; ++b)
            if (FB == 1 || b < nb) acc[b].add(v, xp[b * xf]);
    }
#pragma unroll
    for (int b = 0; b < FB; ++b)
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) acc[b].merge_from_lane(m);
    if (lane == 0) {
        T *yp = y + row * yc + f0 * yf;
#pragma unroll
        for (int b = 0; b < FB; ++b)
            if (FB == 1 || b < nb) yp[b * yf] = acc[b].finish(area, inv_area, row);
    }
}

// ---- FP32, fields fastest (field stride 1) ----------------------------------------
// One wave per row; lane l owns fields l, l+64, ... (up to kFieldRegs per lane).
constexpr int kFieldRegs = 4;

__global__ void __launch_bounds__(256) k_apply_fieldsfast(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ val, int64_t nrows, const float *__restrict__ x, int64_t xc,
    float *__restrict__ y, int64_t yc, int64_t nfields, const double *__restrict__ inv_area) {
    int64_t row = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (row >= nrows) return;
    int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
    for (int64_t f0 = 0; f0 < nfields; f0 += 64 * kFieldRegs) {
        Acc<float> acc[kFieldRegs];
#pragma unroll
        for (int j = 0; j < kFieldRegs; ++j) acc[j].init();
        for (int64_t k = k0; k < k1; ++k) {
            const float *xr = x + (int64_t)col[k] * xc;
            float v = (float)val[k];
#pragma unroll
            for (int j = 0; j < kFieldRegs; ++j) {
                int64_t f = f0 + lane + 64 * j;
                if (f < nfields) acc[j].add(v, xr[f]);
            }
        }
#pragma unroll
        for (int j = 0; j < kFieldRegs; ++j) {
            int64_t f = f0 + lane + 64 * j;
            if (f < nfields) y[row * yc + f] = acc[j].finish(nullptr, inv_area, row);
        }
    }
}

__global__ void k_inv(const double *__restrict__ a, int64_t n, double *__restrict__ inv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[i] = 1.0 / a[i];
}

// Rows with more than kLongRow non-zeros, appended in any order (each long row
// is reduced by its own wave, so the list order does not affect results).
__global__ void k_long_rows(const int64_t *__restrict__ rowptr, int64_t n, int32_t *__restrict__ list,
                            unsigned long long *__restrict__ count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (rowptr[i + 1] - rowptr[i] > kLongRow) list[atomicAdd(count, 1ull)] = (int32_t)i;
}

// The long-row list of A, made on the first row-kernel launch (synchronises s once).
static void ensure_long_rows(const Csr &A, hipStream_t s) {
    std::lock_guard<std::mutex> lock(A.plan_mu);
    if (A.nlong >= 0) return;
    CR_REQUIRE_NOT_CAPTURING(s, "the long-row list of this direction");
    int64_t nlong = 0;
    if (A.nrows) {
        DevBuf<int32_t> list(A.nrows, s);
        DevBuf<unsigned long long> count(1, s);
        CR_HIP(hipMemsetAsync(count.p, 0, count.bytes(), s));
        hipLaunchKernelGGL(k_long_rows, dim3(grid_for(A.nrows, 256)), dim3(256), 0, s, A.rowptr.p, A.nrows,
                           list.p, count.p);
        CR_HIP(hipGetLastError());
        unsigned long long h = 0;
        CR_HIP(hipMemcpyAsync(&h, count.p, sizeof(h), hipMemcpyDeviceToHost, s));
        CR_HIP(hipStreamSynchronize(s));
        nlong = (int64_t)h;
        if (nlong) {
            A.long_rows.alloc_kept(nlong, s);
            CR_HIP(hipMemcpyAsync(A.long_rows.p, list.p, nlong * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        }
    }
    A.nlong = nlong;
}

void finalize_regridder(cr_regridder *R, hipStream_t s) {
    R->fwd.inv_area.alloc_kept(R->fwd.nrows, s);
    R->bwd.inv_area.alloc_kept(R->bwd.nrows, s);
    hipLaunchKernelGGL(k_inv, dim3(grid_for(R->fwd.nrows, 256)), dim3(256), 0, s, R->a_dst.p,
                       R->fwd.nrows, R->fwd.inv_area.p);
    hipLaunchKernelGGL(k_inv, dim3(grid_for(R->bwd.nrows, 256)), dim3(256), 0, s, R->a_src.p,
                       R->bwd.nrows, R->bwd.inv_area.p);
    CR_HIP(hipGetLastError());
}

// ---- launch ------------------------------------------------------------------------
template <class T, int FB>
static void launch_rows(const Csr &A, const T *x, int64_t sc, int64_t sf, T *y, int64_t dc,
                        int64_t df, int64_t nf, const double *area, const double *inv, hipStream_t s) {
    int64_t ntiles = (A.nrows + kApplyBlock - 1) / kApplyBlock;
    int64_t tiles_per_xcd = (ntiles + 7) / 8;
    int64_t gy = (nf + FB - 1) / FB;
    CR_REQUIRE(gy <= 65535, CR_UNSUPPORTED, "too many fields for one launch");
    ensure_long_rows(A, s);
    hipLaunchKernelGGL((k_apply_rows<T, FB>), dim3((unsigned)(tiles_per_xcd * 8), (unsigned)gy),
                       dim3(kApplyBlock), 0, s, A.rowptr.p, A.col.p, A.val.p, A.nrows, x, sc, sf, y, dc,
                       df, nf, area, inv, tiles_per_xcd);
    if (A.nlong)
        hipLaunchKernelGGL((k_apply_long<T, FB>), dim3((unsigned)A.nlong, (unsigned)gy), dim3(64), 0, s,
                           A.long_rows.p, A.rowptr.p, A.col.p, A.val.p, x, sc, sf, y, dc, df, nf, area,
                           inv);
}

// Staged tile kernel (tiles.hip) followed by the rows it leaves to the long-row
// kernel.  scale: 1/area for FP32 fields, area for FP64 fields, or nullptr.
template <class T>
void launch_tiles(const TilePlan &P, int F, int acc, int nchunks, int chunk_major, const T *x, int64_t sc,
                  int64_t sf, T *y, int64_t dc, int64_t df, int64_t nf, const double *scale, int64_t nrows,
                  int64_t ncols, hipStream_t s);

template <class T>
static void launch_tiled(const Csr &A, int geom, int F, int acc, int nchunks, int chunk_major, const T *x,
                         int64_t sc, int64_t sf, T *y, int64_t dc, int64_t df, int64_t nf, const double *area,
                         const double *inv, hipStream_t s, bool blocks = false) {
    constexpr bool f32 = sizeof(T) == 4;
    const TilePlan &P = tile_plan(A, f32 ? kRounds : kEven, geom, s, blocks);
    launch_tiles<T>(P, F, acc, nchunks, chunk_major, x, sc, sf, y, dc, df, nf, f32 ? inv : area, A.nrows,
                    A.ncols, s);
    if (P.nhuge) {
        constexpr int FB = f32 ? 8 : 1;
        const int64_t gy = (nf + FB - 1) / FB;
        CR_REQUIRE(gy <= 65535, CR_UNSUPPORTED, "too many fields for one launch");
        hipLaunchKernelGGL((k_apply_long<T, FB>), dim3((unsigned)P.nhuge, (unsigned)gy), dim3(64), 0, s,
                           P.huge.p, A.rowptr.p, A.col.p, A.val.p, x, sc, sf, y, dc, df, nf, area, inv);
    }
}

// FP32 kernel variants the autotuner tries (index = variant): the row kernel
// with field-block widths 1..16 (Kahan), the staged tile kernel with 2, 4 or 8
// fields per LDS stage and Kahan sums, then the tile kernel with FP64 sums.
struct F32Variant {
    bool tiled;
    int width;        // FB (row kernel) or F (tile kernel)
    int acc;          // tile kernel: 0 Kahan FP32 sums, 1 FP64 sums, 2 FP64 sums of FP32 values
    int chunks;       // tile kernel: field chunks per launch (0: enough workgroups)
    int chunk_major;  // tile kernel: sweep one field chunk at a time
    int geom;         // tile kernel: TileGeomId
    int blocks = 0;   // tile kernel: plan over 32-row blocks sorted by their smallest column
};
constexpr F32Variant kF32Variants[] = {
    {false, 1, false, 0, 0, 0},  {false, 2, false, 0, 0, 0},  {false, 4, false, 0, 0, 0},
    {false, 8, false, 0, 0, 0},  {false, 16, false, 0, 0, 0}, {true, 2, false, 0, 0, 0},
    {true, 4, false, 0, 0, 0},   {true, 8, false, 0, 0, 0},   {true, 2, true, 0, 0, 0},
    {true, 4, true, 0, 0, 0},    {true, 8, true, 0, 0, 0},    {true, 4, true, 4, 0, 0},
    {true, 4, true, 8, 1, 0},    {true, 4, true, 16, 1, 0},   {true, 4, true, 1, 0, 0},
    {true, 4, true, 0, 0, 1},    {true, 2, true, 0, 0, 1},    {true, 4, true, 4, 0, 1},
    {true, 4, false, 0, 0, 1},   {true, 4, false, 4, 0, 1},   {true, 4, 2, 0, 0, 1},
    {true, 4, 2, 4, 0, 1},       {true, 4, 2, 0, 0, 1, 1},    {true, 4, 2, 4, 0, 1, 1},
    {true, 4, 2, 4, 1, 1},       {true, 4, 1, 4, 1, 1},       {true, 4, 2, 2, 1, 1},
    {true, 4, 2, 4, 1, 1, 1},    {true, 4, 1, 0, 0, 1, 1}};
constexpr int kNumVariants = sizeof(kF32Variants) / sizeof(kF32Variants[0]);
// 1024-row tiles, 4 fields per stage, products and sums in FP64 from the FP64 A,
// field chunks by default_chunks
constexpr int kDefaultF32Variant = 15;
constexpr int kBlockF32Variant = 28;  // the same on the block-ordered plan

// The default FP32 variant of a matrix, chosen once: the block-ordered plan
// when the plain plan's tiles fetch more than 3 source lines (32 columns each)
// per 32 distinct columns and the block-ordered one fetches fewer (rows that
// are strips of a row-major grid over sources numbered in another order).
static int default_f32_variant(const Csr &A, hipStream_t s) {
    int v = A.default_variant.load();
    if (v >= 0) return v;
    CR_REQUIRE_NOT_CAPTURING(s, "the default FP32 kernel of this direction");
    v = kDefaultF32Variant;
    const TilePlan &P = tile_plan(A, kRounds, kGeomL, s, false);
    if (P.tile_cols > 0 && P.tile_lines * 32 > 3 * P.tile_cols) {
        const TilePlan &Q = tile_plan(A, kRounds, kGeomL, s, true);
        if (Q.tile_cols > 0 && Q.tile_lines * P.tile_cols < P.tile_lines * Q.tile_cols) v = kBlockF32Variant;
    }
    A.default_variant.store(v);
    return v;
}

static void launch_f32_variant(int v, const Csr &A, const float *x, int64_t sc, int64_t sf, float *y,
                               int64_t dc, int64_t df, int64_t nf, const double *inv, hipStream_t s) {
    const F32Variant &V = kF32Variants[v];
    if (V.tiled) {
        launch_tiled<float>(A, V.geom, V.width, V.acc, V.chunks, V.chunk_major, x, sc, sf, y, dc, df, nf,
                            nullptr, inv, s, V.blocks != 0);
        return;
    }
    switch (V.width) {
        case 1: launch_rows<float, 1>(A, x, sc, sf, y, dc, df, nf, nullptr, inv, s); break;
        case 2: launch_rows<float, 2>(A, x, sc, sf, y, dc, df, nf, nullptr, inv, s); break;
        case 4: launch_rows<float, 4>(A, x, sc, sf, y, dc, df, nf, nullptr, inv, s); break;
        case 8: launch_rows<float, 8>(A, x, sc, sf, y, dc, df, nf, nullptr, inv, s); break;
        default: launch_rows<float, 16>(A, x, sc, sf, y, dc, df, nf, nullptr, inv, s); break;
    }
}

constexpr int kNumF64Variants = 2;  // 0 row kernel, 1 tile kernel
constexpr int kDefaultF64Variant = 1;

// Kernel choice: FP64 -> the tile kernel (1 or 2 fields per stage) unless the
// row kernel is pinned; FP32 with unit field strides and many fields ->
// fields-fastest kernel; other FP32 layouts -> the autotuned (or pinned, or
// default: 1024-row tiles, 4 fields per stage, FP64 sums) variant.
void apply(const cr_regridder *R, bool transpose, int dtype, int64_t nf, const void *src,
           int64_t sc, int64_t sf, void *dst, int64_t dc, int64_t df, bool normalize, hipStream_t s,
           int force_variant = -1) {
    const Csr &A = transpose ? R->bwd : R->fwd;
    const double *area = transpose ? R->a_src.p : R->a_dst.p;
    if (A.nrows == 0 || nf == 0) return;
    if (dtype == CR_F64) {
        const double *x = (const double *)src;
        double *y = (double *)dst;
        const double *ar = normalize ? area : nullptr;
        const int v = A.f64_variant >= 0 ? A.f64_variant : kDefaultF64Variant;
        if (v == 1)
            launch_tiled<double>(A, kGeomS, nf >= 2 ? 2 : 1, false, 0, 0, x, sc, sf, y, dc, df, nf, ar, nullptr,
                                 s);
        else if (nf == 1) launch_rows<double, 1>(A, x, sc, sf, y, dc, df, nf, ar, nullptr, s);
        else launch_rows<double, 8>(A, x, sc, sf, y, dc, df, nf, ar, nullptr, s);
    } else if (sf == 1 && df == 1 && nf >= 64 && force_variant < 0) {
        hipLaunchKernelGGL(k_apply_fieldsfast, dim3(grid_for(A.nrows, 4)), dim3(256), 0, s, A.rowptr.p,
                           A.col.p, A.val.p, A.nrows, (const float *)src, sc, (float *)dst, dc, nf,
                           normalize ? A.inv_area.p : nullptr);
    } else {
        int v = force_variant >= 0 ? force_variant : A.tuned_variant;
        if (v < 0 || v >= kNumVariants) v = default_f32_variant(A, s);
        launch_f32_variant(v, A, (const float *)src, sc, sf, (float *)dst, dc, df, nf,
                           normalize ? A.inv_area.p : nullptr, s);
    }
    CR_HIP(hipGetLastError());
}

// True when the apply above, called with these arguments, would make no
// per-direction set-up (nothing to allocate, nothing to synchronise).
bool apply_ready(const cr_regridder *R, bool transpose, int dtype, int64_t nf, int64_t sf, int64_t df) {
    const Csr &A = transpose ? R->bwd : R->fwd;
    if (A.nrows == 0 || nf == 0) return true;
    auto long_rows_ready = [&] {
        std::lock_guard<std::mutex> lock(A.plan_mu);
        return A.nlong >= 0;
    };
    if (dtype == CR_F64) {
        const int v = A.f64_variant >= 0 ? A.f64_variant : kDefaultF64Variant;
        return v == 1 ? plan_ready(A, kEven, kGeomS, false) : long_rows_ready();
    }
    if (sf == 1 && df == 1 && nf >= 64) return true;
    int v = A.tuned_variant;
    if (v < 0 || v >= kNumVariants) {
        v = A.default_variant.load();
        if (v < 0) return false;
    }
    const F32Variant &V = kF32Variants[v];
    return V.tiled ? plan_ready(A, kRounds, V.geom, V.blocks != 0) : long_rows_ready();
}

// The per-direction set-up the apply above makes on its first call with these
// arguments (default variant, tile plan or long-row list), without launching
// the apply: afterwards such calls allocate nothing and never synchronise, so
// they can be captured into a HIP graph.
void prepare(const cr_regridder *R, bool transpose, int dtype, int64_t nf, int64_t sf, int64_t df,
             hipStream_t s) {
    const Csr &A = transpose ? R->bwd : R->fwd;
    if (A.nrows == 0 || nf == 0) return;
    if (dtype == CR_F64) {
        const int v = A.f64_variant >= 0 ? A.f64_variant : kDefaultF64Variant;
        if (v == 1) {
            const TilePlan &P = tile_plan(A, kEven, kGeomS, s, false);
            (void)P;
        } else {
            ensure_long_rows(A, s);
        }
    } else if (!(sf == 1 && df == 1 && nf >= 64)) {
        int v = A.tuned_variant;
        if (v < 0 || v >= kNumVariants) v = default_f32_variant(A, s);
        const F32Variant &V = kF32Variants[v];
        if (V.tiled) {
            const TilePlan &P = tile_plan(A, kRounds, V.geom, s, V.blocks != 0);
            (void)P;
        } else {
            ensure_long_rows(A, s);
        }
    }
    CR_HIP(hipStreamSynchronize(s));
}

// Fold partners: y[partner] = y[primary] for every field (the mirror kernel of
// ext/ConservativeRegriddingOceananigansExt.jl:223-240), and its adjoint
// g[primary] += g[partner], g[partner] = 0 (each partner has one primary).
template <class T>
__global__ void k_mirror(T *__restrict__ y, int64_t yc, int64_t yf, int64_t nf, const int32_t *__restrict__ pairs,
                         int64_t npairs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * nf) return;
    const int64_t f = t / npairs, k = t - f * npairs;
    y[pairs[2 * k + 1] * yc + f * yf] = y[pairs[2 * k] * yc + f * yf];
}

__global__ void k_mirror_adjoint(double *__restrict__ g, int64_t gc, int64_t gf, int64_t nf,
                                 const int32_t *__restrict__ pairs, int64_t npairs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * nf) return;
    const int64_t f = t / npairs, k = t - f * npairs;
    double &gp = g[pairs[2 * k + 1] * gc + f * gf];
    g[pairs[2 * k] * gc + f * gf] += gp;
    gp = 0.0;
}

}  // namespace cr

using namespace cr;

extern "C" {

cr_status cr_mirror_fold_partners(const cr_regridder *R, int32_t transpose, int32_t dtype, int64_t nfields,
                                  void *dst, int64_t dst_cell_stride, int64_t dst_field_stride, void *stream) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "cr_mirror_fold_partners: null regridder");
        CR_REQUIRE(dtype == CR_F64 || dtype == CR_F32, CR_UNSUPPORTED, "dtype must be F64 or F32");
        CR_REQUIRE(nfields >= 0 && dst_cell_stride > 0 && (nfields <= 1 || dst_field_stride > 0),
                   CR_INVALID_ARGUMENT, "cr_mirror_fold_partners: bad sizes or strides");
        const int64_t n = transpose ? R->nfold_src : R->nfold_dst;
        if (n == 0 || nfields == 0) return;
        CR_REQUIRE(dst, CR_INVALID_ARGUMENT, "cr_mirror_fold_partners: null field pointer");
        const int32_t *pairs = transpose ? R->fold_src.p : R->fold_dst.p;
        hipStream_t s = as_stream(stream);
        const unsigned g = grid_for(n * nfields, 256);
        if (dtype == CR_F64)
            hipLaunchKernelGGL(k_mirror<double>, dim3(g), dim3(256), 0, s, (double *)dst, dst_cell_stride,
                               dst_field_stride, nfields, pairs, n);
        else
            hipLaunchKernelGGL(k_mirror<float>, dim3(g), dim3(256), 0, s, (float *)dst, dst_cell_stride,
                               dst_field_stride, nfields, pairs, n);
        CR_HIP(hipGetLastError());
    });
}

cr_status cr_mirror_fold_partners_adjoint(const cr_regridder *R, int32_t transpose, int64_t nfields, double *g,
                                          int64_t cell_stride, int64_t field_stride, void *stream) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "cr_mirror_fold_partners_adjoint: null regridder");
        CR_REQUIRE(nfields >= 0 && cell_stride > 0 && (nfields <= 1 || field_stride > 0), CR_INVALID_ARGUMENT,
                   "cr_mirror_fold_partners_adjoint: bad sizes or strides");
        const int64_t n = transpose ? R->nfold_src : R->nfold_dst;
        if (n == 0 || nfields == 0) return;
        CR_REQUIRE(g, CR_INVALID_ARGUMENT, "cr_mirror_fold_partners_adjoint: null gradient pointer");
        hipLaunchKernelGGL(k_mirror_adjoint, dim3(grid_for(n * nfields, 256)), dim3(256), 0, as_stream(stream), g,
                           cell_stride, field_stride, nfields, transpose ? R->fold_src.p : R->fold_dst.p, n);
        CR_HIP(hipGetLastError());
    });
}

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

cr_status cr_regrid_prepare(const cr_regridder *R, int32_t transpose, int32_t dtype, int64_t nfields,
                            int64_t src_field_stride, int64_t dst_field_stride, void *stream) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "cr_regrid_prepare: null regridder");
        CR_REQUIRE(dtype == CR_F64 || dtype == CR_F32, CR_UNSUPPORTED, "cr_regrid_prepare: dtype must be F64 or F32");
        CR_REQUIRE(nfields >= 0 && src_field_stride >= 0 && dst_field_stride >= 0, CR_INVALID_ARGUMENT,
                   "cr_regrid_prepare: bad sizes or strides");
        prepare(R, transpose != 0, dtype, nfields, src_field_stride, dst_field_stride, as_stream(stream));
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
        for (int v = 0; v < kNumVariants; ++v) {
            // the variants with the default's numerics: products and sums in FP64 from the FP64 A
            if (!(kF32Variants[v].tiled && kF32Variants[v].acc == 1)) continue;
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
            if (t[1] < best_t) {
                best_t = t[1];
                best = v;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        A.tuned_variant = best;
        if (chosen) *chosen = best;
        if (best_ms) *best_ms = best_t;
    });
}

int32_t cr_regrid_variant_count(int32_t dtype) {
    return dtype == CR_F32 ? kNumVariants : dtype == CR_F64 ? kNumF64Variants : 0;
}

cr_status cr_regrid_set_variant(cr_regridder *R, int32_t transpose, int32_t dtype, int32_t variant) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "cr_regrid_set_variant: null regridder");
        CR_REQUIRE(dtype == CR_F64 || dtype == CR_F32, CR_UNSUPPORTED, "cr_regrid_set_variant: dtype must be F64 or F32");
        const int n = dtype == CR_F32 ? kNumVariants : kNumF64Variants;
        CR_REQUIRE(variant >= -1 && variant < n, CR_INVALID_ARGUMENT, "cr_regrid_set_variant: no such variant");
        Csr &A = transpose ? R->bwd : R->fwd;
        if (dtype == CR_F32) A.tuned_variant = variant;
        else A.f64_variant = variant;
    });
}

}  // extern "C"
