// planblocks.hip -- tile plans over 32-row blocks in an order of their own
// (tileplan.hpp, plan.hip): for a matrix whose rows are strips of a row-major
// grid over sources numbered in another order (HEALPix nested sources under
// lon-lat rows), a tile of contiguous rows gathers source cells from all over
// the source numbering and uses a small part of each fetched line.  Sorting
// the 32-row blocks by their smallest column and planning the reordered copy
// gives tiles whose source cells are close in the source numbering; each
// block is still one 128-byte line of every FP32 output field, and the tile
// kernel writes plan row v to matrix row rowmap[v].  Its own file: its kernels
// and sorts load only when a block-ordered plan is first made.
#include "regridder.hpp"

namespace cr {

template <class G, int STYLE>
void make_plan(const Csr &A, TilePlan &P, hipStream_t s);

namespace {

// Smallest column of each 32-row block (UINT32_MAX for a block with no non-zeros).
__global__ void k_block_keys(int64_t nrows, int64_t nb, const int64_t *__restrict__ rp,
                             const int32_t *__restrict__ col, uint32_t *__restrict__ key,
                             int32_t *__restrict__ id) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int64_t k0 = rp[b * kRowAlign], k1 = rp[min(nrows, (b + 1) * kRowAlign)];
    uint32_t m = 0xffffffffu;
    for (int64_t k = k0; k < k1; ++k) m = min(m, (uint32_t)col[k]);
    key[b] = m;
    id[b] = (int32_t)b;
}

// Rows of the block-ordered matrix: plan row v = 32 * i + j is matrix row
// 32 * order[i] + j, or no row (past the last matrix row).
__global__ void k_block_rows(int64_t nrows, int64_t nv, const int32_t *__restrict__ order,
                             const int64_t *__restrict__ rp, int32_t *__restrict__ rowmap,
                             int64_t *__restrict__ n) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int64_t r = (int64_t)order[v / kRowAlign] * kRowAlign + v % kRowAlign;
    const bool ok = r < nrows;
    rowmap[v] = ok ? (int32_t)r : -1;
    n[v] = ok ? rp[r + 1] - rp[r] : 0;
}

__global__ void k_block_gather(int64_t nv, const int32_t *__restrict__ rowmap, const int64_t *__restrict__ rp,
                               const int32_t *__restrict__ col, const double *__restrict__ val,
                               const int64_t *__restrict__ rp2, int32_t *__restrict__ col2,
                               double *__restrict__ val2) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int32_t r = rowmap[v];
    if (r < 0) return;
    const int64_t k0 = rp[r], o = rp2[v], n = rp[r + 1] - k0;
    for (int64_t k = 0; k < n; ++k) {
        col2[o + k] = col[k0 + k];
        val2[o + k] = val[k0 + k];
    }
}

__global__ void k_map_rows(int32_t *__restrict__ rows, int64_t n, const int32_t *__restrict__ rowmap) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) rows[t] = rowmap[rows[t]];
}

}  // namespace

// The plan of A's rows taken as 32-row blocks in the order of their smallest
// column: the block-ordered copy of A is planned as usual, its row map kept.
template <class G, int STYLE>
void make_block_plan(const Csr &A, TilePlan &P, hipStream_t s) {
    const int64_t nrows = A.nrows, nnz = A.nnz;
    if (nrows == 0) return;
    const int64_t nb = (nrows + kRowAlign - 1) / kRowAlign, nv = nb * kRowAlign;
    CR_REQUIRE(nv < ((int64_t)1 << 31) - kRowAlign, CR_UNSUPPORTED, "tile plans take fewer than 2^31 rows");
    Temp T(s);
    Csr B;  // the block-ordered copy
    B.nrows = nv;
    B.ncols = A.ncols;
    B.nnz = nnz;
    DevBuf<int32_t> rowmap;
    rowmap.alloc_kept(nv, s);
    {
        DevBuf<uint32_t> key(nb, s), skey(nb, s);
        DevBuf<int32_t> id(nb, s), order(nb, s);
        hipLaunchKernelGGL(k_block_keys, dim3(grid_for(nb, 256)), dim3(256), 0, s, nrows, nb, A.rowptr.p,
                           A.col.p, key.p, id.p);
        sort_pairs<uint32_t, int32_t>(T, key.p, skey.p, id.p, order.p, nb, 32, s);
        DevBuf<int64_t> n(nv + 1, s);
        CR_HIP(hipMemsetAsync(n.p + nv, 0, sizeof(int64_t), s));
        hipLaunchKernelGGL(k_block_rows, dim3(grid_for(nv, 256)), dim3(256), 0, s, nrows, nv, order.p, A.rowptr.p,
                           rowmap.p, n.p);
        B.rowptr.alloc_async(nv + 1, s);
        exclusive_sum<int64_t *, int64_t *>(T, n.p, B.rowptr.p, nv + 1, s);
    }
    B.col.alloc_async(nnz > 0 ? nnz : 1, s);
    B.val.alloc_async(nnz > 0 ? nnz : 1, s);
    hipLaunchKernelGGL(k_block_gather, dim3(grid_for(nv, 256)), dim3(256), 0, s, nv, rowmap.p, A.rowptr.p,
                       A.col.p, A.val.p, B.rowptr.p, B.col.p, B.val.p);
    CR_HIP(hipGetLastError());
    make_plan<G, STYLE>(B, P, s);
    if (P.nhuge)
        hipLaunchKernelGGL(k_map_rows, dim3(grid_for(P.nhuge, 256)), dim3(256), 0, s, P.huge.p, P.nhuge,
                           rowmap.p);
    CR_HIP(hipGetLastError());
    CR_HIP(hipStreamSynchronize(s));  // B's buffers are released after this
    P.rowmap = std::move(rowmap);
}


template void make_block_plan<GeomS, kRounds>(const Csr &, TilePlan &, hipStream_t);
template void make_block_plan<GeomS, kEven>(const Csr &, TilePlan &, hipStream_t);
template void make_block_plan<GeomL, kRounds>(const Csr &, TilePlan &, hipStream_t);
template void make_block_plan<GeomL, kEven>(const Csr &, TilePlan &, hipStream_t);

}  // namespace cr
