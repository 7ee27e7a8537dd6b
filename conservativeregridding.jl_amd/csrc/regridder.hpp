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

#include <atomic>
#include <memory>
#include <mutex>

#include "common.hpp"
#include "tileplan.hpp"

namespace cr {

struct TilePlan {
    int geom = kGeomS;
    int64_t ntiles = 0;
    DevBuf<TileDesc> tiles;
    DevBuf<uint32_t> key;    // per slot: local column (low 16 bits) | segment id (high 16)
    DevBuf<float> val32;     // per slot: A value rounded to FP32 (0 in padding slots)
    DevBuf<double> val64;    // per slot: A value
    DevBuf<int32_t> cols;    // distinct columns of each tile, ascending
    DevBuf<uint16_t> rseg;   // per tile, nrows + 1 entries: first segment of each row
    DevBuf<int32_t> huge;    // rows too long for a tile (done by the long-row kernel)
    // block-ordered plans: plan row v is matrix row rowmap[v] (-1: none); the
    // plan's rows are 32-row blocks of the matrix in an order of their own
    DevBuf<int32_t> rowmap;
    int64_t nhuge = 0;
    int64_t nslots = 0;
    // the tiles' distinct columns and the 32-column lines (128 bytes of an FP32
    // field) they touch, summed over tiles: lines * 32 / cols is the plan's
    // source-line fetch per source value used
    int64_t tile_cols = 0, tile_lines = 0;
};

struct Csr {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    DevBuf<int64_t> rowptr;
    DevBuf<int32_t> col;
    DevBuf<double> val;
    DevBuf<double> inv_area;   // 1 / area of the row cells (FP32 batched path)
    int tuned_variant = -1;    // FP32 kernel variant (cr_regrid_tune / set_variant); -1: default
    mutable std::atomic<int> default_variant{-1};  // the default FP32 variant chosen for this matrix; -1: not yet
    int f64_variant = -1;      // FP64 kernel variant; -1: default
    // Kernel plans, made on first use (guarded: handles are shared by callers):
    // the staged-apply tiles, and the rows too long for the row kernels.
    mutable std::shared_ptr<TilePlan> tiles[2][2][2];  // per TileStyle, TileGeomId, block order
    mutable DevBuf<int32_t> long_rows;
    mutable int64_t nlong = -1;  // -1: not planned yet
    mutable std::mutex plan_mu;
};

// The tile plan of A in the given TileStyle and TileGeomId, made on the device
// on the first call (synchronises stream s).  blocks: plan the rows as 32-row
// blocks sorted by their smallest column (compact source patches for tiles
// whose rows are strips of a row-major grid).
const TilePlan &tile_plan(const Csr &A, int style, int geom, hipStream_t s, bool blocks = false);
// Forget the plan (after A's values change).
void drop_tile_plan(const Csr &A);
// Whether that plan is made already (making it allocates and synchronises).
bool plan_ready(const Csr &A, int style, int geom, bool blocks);

}  // namespace cr

struct cr_regridder {
    cr::Csr fwd;  // A:   n_dst x n_src
    cr::Csr bwd;  // A^T: n_src x n_dst
    cr::DevBuf<double> a_dst, a_src;
    cr_build_stats stats{};
    // stored areas = area_scale * unit-sphere areas (radius^2, divided by max(A)
    // after normalize!); 0 when built from arrays (no geometry: no vertex gradients)
    double area_scale = 0.0;
    // fold partners of the destination / source grid, (primary, partner) pairs:
    // cr_mirror_fold_partners copies primary -> partner after an apply into that grid
    cr::DevBuf<int32_t> fold_dst, fold_src;
    int64_t nfold_dst = 0, nfold_src = 0;
};

namespace cr {
// Finish a regridder whose CSR arrays and areas are filled: inverse areas and
// apply plans.
void finalize_regridder(cr_regridder *R, hipStream_t s);
// The device build of Regridder(dst, src) into R (no normalize!).
void build_regridder(cr_regridder *R, const cr_grid *dst, const cr_grid *src, const cr_build_opts &o,
                     hipStream_t s);
// Planar cell areas of a planar grid (grids.hip), on stream s.
void launch_planar_areas(const cr_grid *g, double *area, hipStream_t s);
// Temporary storage of the device-wide rocPRIM algorithms, grown on demand.
struct Temp {
    hipStream_t s;
    DevBuf<unsigned char> buf;
    explicit Temp(hipStream_t st) : s(st) {}
    void *get(size_t bytes) {
        if (buf.n < bytes) buf.alloc_async(bytes > 256 ? bytes : (size_t)256, s);
        return buf.p;
    }
};
// Device-wide rocPRIM exclusive sum and stable radix sort of (key, value) pairs
// on the bits [0, end_bit), with T's temporary storage (build.hip).
template <class In, class Out>
void exclusive_sum(Temp &T, In in, Out out, int64_t n, hipStream_t s);
template <class K, class V>
void sort_pairs(Temp &T, const K *kin, K *kout, const V *vin, V *vout, int64_t n, int end_bit, hipStream_t s);
// CSR(A^T) of the device CSR(A) F whose entry k lies in row coo_row[k].
void csr_transpose(const Csr &F, const int32_t *coo_row, Csr &Bk, Temp &T, hipStream_t s);
}  // namespace cr
