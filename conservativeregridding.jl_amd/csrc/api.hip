// api.hip -- library-level entry points: errors, Regridder construction from
// CSR arrays, queries and host copies.
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "regridder.hpp"

namespace cr {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

// Keep freed pool memory cached (default threshold 0 returns it at every sync).
void configure_pool() {
    static thread_local int done_for_device = -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev == done_for_device) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    done_for_device = dev;
}

// CSR(A) -> CSR(A^T) on the host (used by cr_regridder_from_csr only).
static void host_transpose(int64_t nrows, int64_t ncols, const int64_t *rp, const int32_t *col,
                           const double *val, std::vector<int64_t> &trp, std::vector<int32_t> &tcol,
                           std::vector<double> &tval) {
    int64_t nnz = rp[nrows];
    trp.assign(ncols + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) trp[col[k] + 1]++;
    for (int64_t j = 0; j < ncols; ++j) trp[j + 1] += trp[j];
    std::vector<int64_t> pos(trp.begin(), trp.end() - 1);
    tcol.resize(nnz);
    tval.resize(nnz);
    for (int64_t i = 0; i < nrows; ++i)
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
            int64_t p = pos[col[k]]++;
            tcol[p] = (int32_t)i;
            tval[p] = val[k];
        }
}

template <class T>
static void upload(DevBuf<T> &b, const T *h, size_t n, hipStream_t s) {
    b.alloc(n);
    if (n) CR_HIP(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}

}  // namespace cr

using namespace cr;

extern "C" {

const char *cr_last_error(void) { return g_last_error.c_str(); }

const char *cr_version(void) { return "crhip 0.1.0 (gfx950)"; }

int32_t cr_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

cr_status cr_regridder_from_csr(cr_regridder **R, int64_t n_dst, int64_t n_src, int64_t nnz,
                                const int64_t *rowptr, const int32_t *col, const double *val,
                                const double *dst_areas, const double *src_areas, void *stream) {
    return guard([&] {
        CR_REQUIRE(R && rowptr && dst_areas && src_areas && n_dst > 0 && n_src > 0 && nnz >= 0,
                   CR_INVALID_ARGUMENT, "cr_regridder_from_csr: bad arguments");
        CR_REQUIRE(nnz == 0 || (col && val), CR_INVALID_ARGUMENT, "cr_regridder_from_csr: null col/val");
        CR_REQUIRE(rowptr[0] == 0 && rowptr[n_dst] == nnz, CR_DIMENSION_MISMATCH,
                   "cr_regridder_from_csr: rowptr does not span nnz");
        for (int64_t i = 0; i < n_dst; ++i)
            CR_REQUIRE(rowptr[i + 1] >= rowptr[i], CR_INVALID_ARGUMENT, "rowptr not monotone");
        for (int64_t k = 0; k < nnz; ++k)
            CR_REQUIRE(col[k] >= 0 && col[k] < n_src, CR_DIMENSION_MISMATCH, "column index out of range");
        check_device();
        hipStream_t s = as_stream(stream);
        std::vector<int64_t> trp;
        std::vector<int32_t> tcol;
        std::vector<double> tval;
        host_transpose(n_dst, n_src, rowptr, col, val, trp, tcol, tval);
        auto *r = new cr_regridder();
        try {
            r->fwd.nrows = n_dst; r->fwd.ncols = n_src; r->fwd.nnz = nnz;
            r->bwd.nrows = n_src; r->bwd.ncols = n_dst; r->bwd.nnz = nnz;
            upload(r->fwd.rowptr, rowptr, n_dst + 1, s);
            upload(r->fwd.col, col, nnz, s);
            upload(r->fwd.val, val, nnz, s);
            upload(r->bwd.rowptr, trp.data(), n_src + 1, s);
            upload(r->bwd.col, tcol.data(), nnz, s);
            upload(r->bwd.val, tval.data(), nnz, s);
            upload(r->a_dst, dst_areas, n_dst, s);
            upload(r->a_src, src_areas, n_src, s);
            finalize_regridder(r, s);
            CR_HIP(hipStreamSynchronize(s));
            r->stats.n_dst = n_dst;
            r->stats.n_src = n_src;
            r->stats.nnz = nnz;
        } catch (...) {
            delete r;
            throw;
        }
        *R = r;
    });
}

cr_status cr_regridder_info(const cr_regridder *R, int64_t *n_dst, int64_t *n_src, int64_t *nnz) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "null regridder");
        if (n_dst) *n_dst = R->fwd.nrows;
        if (n_src) *n_src = R->fwd.ncols;
        if (nnz) *nnz = R->fwd.nnz;
    });
}

cr_status cr_regridder_stats(const cr_regridder *R, cr_build_stats *st) {
    return guard([&] {
        CR_REQUIRE(R && st, CR_INVALID_ARGUMENT, "null argument");
        *st = R->stats;
    });
}

// Host copy with the columns of every row ascending (the order sparse() gives
// the reference's matrix); the device keeps each row in candidate order.
cr_status cr_regridder_copy_csr(const cr_regridder *R, int32_t transpose, int64_t *rowptr,
                                int32_t *col, double *val) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "null regridder");
        const Csr &A = transpose ? R->bwd : R->fwd;
        std::vector<int64_t> rp(A.nrows + 1);
        std::vector<int32_t> c(A.nnz);
        std::vector<double> v(A.nnz);
        CR_HIP(hipMemcpy(rp.data(), A.rowptr.p, A.rowptr.bytes(), hipMemcpyDeviceToHost));
        if (A.nnz) {
            CR_HIP(hipMemcpy(c.data(), A.col.p, A.col.bytes(), hipMemcpyDeviceToHost));
            CR_HIP(hipMemcpy(v.data(), A.val.p, A.val.bytes(), hipMemcpyDeviceToHost));
        }
        auto sort_rows = [&](int64_t i0, int64_t i1) {
            std::vector<int32_t> order;
            for (int64_t i = i0; i < i1; ++i) {
                const int64_t k0 = rp[i], n = rp[i + 1] - rp[i];
                bool sorted = true;
                for (int64_t k = 1; k < n && sorted; ++k) sorted = c[k0 + k - 1] < c[k0 + k];
                if (sorted) {
                    if (col) std::copy(c.begin() + k0, c.begin() + k0 + n, col + k0);
                    if (val) std::copy(v.begin() + k0, v.begin() + k0 + n, val + k0);
                    continue;
                }
                order.resize(n);
                for (int64_t k = 0; k < n; ++k) order[k] = (int32_t)k;
                std::sort(order.begin(), order.end(),
                          [&](int32_t a, int32_t b) { return c[k0 + a] < c[k0 + b]; });
                for (int64_t k = 0; k < n; ++k) {
                    if (col) col[k0 + k] = c[k0 + order[k]];
                    if (val) val[k0 + k] = v[k0 + order[k]];
                }
            }
        };
        const int nth = (int)std::max<int64_t>(1, std::min<int64_t>(16, A.nrows / 65536));
        std::vector<std::thread> th;
        const int64_t per = (A.nrows + nth - 1) / nth;
        for (int t = 0; t < nth; ++t) {
            const int64_t i0 = t * per, i1 = std::min(A.nrows, i0 + per);
            if (i0 < i1) th.emplace_back(sort_rows, i0, i1);
        }
        for (auto &x : th) x.join();
        if (rowptr) std::copy(rp.begin(), rp.end(), rowptr);
    });
}

cr_status cr_regridder_copy_areas(const cr_regridder *R, double *dst_areas, double *src_areas) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "null regridder");
        if (dst_areas) CR_HIP(hipMemcpy(dst_areas, R->a_dst.p, R->a_dst.bytes(), hipMemcpyDeviceToHost));
        if (src_areas) CR_HIP(hipMemcpy(src_areas, R->a_src.p, R->a_src.bytes(), hipMemcpyDeviceToHost));
    });
}

void cr_regridder_destroy(cr_regridder *R) { delete R; }

}  // extern "C"
