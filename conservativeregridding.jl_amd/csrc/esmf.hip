// esmf.hip -- ESMF offline-weights files (host code).
//
// Reference: save_esmf_weights(path, regridder; src_grid_name, dst_grid_name,
// src_shape, dst_shape, created_at), ext/ConservativeRegriddingNCDatasetsExt.jl:15-59:
// dimensions n_s (non-zeros), n_a (source cells), n_b (destination cells);
// variables S = A ./ dst_areas[row], row, col (1-based, in the column-major
// order findnz gives a SparseMatrixCSC), frac_a = column sums ./ src_areas,
// frac_b = row sums ./ dst_areas, area_a, area_b; global attributes title,
// created_by, source_grid, destination_grid, normalization = "destarea",
// map_method, and the optional created_at / grid shapes.
//
// No netCDF library is available, so the file is written directly in the
// netCDF classic format with 64-bit offsets (CDF-2; published format
// specification "The NetCDF Classic Format Specification"): big-endian
// header of dimension, attribute and variable lists, then each variable's
// data at its recorded offset.  Integer variables are NC_INT (32-bit): the
// classic format has no 64-bit integer type.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "regridder.hpp"

namespace cr {
namespace {

constexpr uint32_t kNcChar = 2, kNcInt = 4, kNcDouble = 6;
constexpr uint32_t kNcDimension = 0x0A, kNcVariable = 0x0B, kNcAttribute = 0x0C;

struct Header {
    std::vector<uint8_t> b;
    void u32(uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void u64(uint64_t v) {
        for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void pad() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string &s) {
        u32((uint32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
    void text_attr(const std::string &n, const std::string &v) {
        name(n);
        u32(kNcChar);
        u32((uint32_t)v.size());
        b.insert(b.end(), v.begin(), v.end());
        pad();
    }
    void int_attr(const std::string &n, const std::vector<int64_t> &v) {
        name(n);
        u32(kNcInt);
        u32((uint32_t)v.size());
        for (int64_t x : v) u32((uint32_t)(int32_t)x);
    }
};

struct Var {
    std::string name, long_name;
    uint32_t dim;   // dimension id
    uint32_t type;  // kNcInt or kNcDouble
    int64_t n;
    const void *data;
    uint64_t bytes() const { return (uint64_t)n * (type == kNcDouble ? 8 : 4); }
};

// Big-endian copy of n values of `size` bytes, in blocks, to the file.
void write_be(FILE *f, const void *src, int64_t n, int size) {
    const unsigned char *p = (const unsigned char *)src;
    std::vector<unsigned char> blk(std::min<int64_t>(n, 1 << 20) * size);
    for (int64_t i0 = 0; i0 < n; i0 += 1 << 20) {
        const int64_t m = std::min<int64_t>(n - i0, 1 << 20);
        for (int64_t i = 0; i < m; ++i)
            for (int k = 0; k < size; ++k) blk[i * size + k] = p[(i0 + i) * size + (size - 1 - k)];
        CR_REQUIRE(fwrite(blk.data(), 1, (size_t)(m * size), f) == (size_t)(m * size), CR_INTERNAL,
                   "esmf: short write");
    }
}

}  // namespace

void write_esmf_file(const char *path, int64_t n_s, const double *S, const int32_t *row, const int32_t *col,
                     int64_t n_a, const double *frac_a, const double *area_a, int64_t n_b,
                     const double *frac_b, const double *area_b, const cr_esmf_attrs &at) {
    CR_REQUIRE(path && n_s >= 0 && n_a > 0 && n_b > 0, CR_INVALID_ARGUMENT, "esmf: bad sizes");
    CR_REQUIRE(n_s == 0 || (S && row && col), CR_INVALID_ARGUMENT, "esmf: null weights");
    CR_REQUIRE(frac_a && area_a && frac_b && area_b, CR_INVALID_ARGUMENT, "esmf: null cell arrays");
    CR_REQUIRE(n_s < ((int64_t)1 << 31) && n_a < ((int64_t)1 << 31) && n_b < ((int64_t)1 << 31),
               CR_UNSUPPORTED, "esmf: dimensions must fit 32 bits");
    const std::vector<Var> vars = {
        {"S", "weight value", 0, kNcDouble, n_s, S},
        {"row", "destination cell index (1-based)", 0, kNcInt, n_s, row},
        {"col", "source cell index (1-based)", 0, kNcInt, n_s, col},
        {"frac_a", "source cell fraction covered", 1, kNcDouble, n_a, frac_a},
        {"frac_b", "destination cell fraction covered", 2, kNcDouble, n_b, frac_b},
        {"area_a", "source cell areas", 1, kNcDouble, n_a, area_a},
        {"area_b", "destination cell areas", 2, kNcDouble, n_b, area_b},
    };
    // header with placeholder offsets, then the real ones (the header size
    // does not depend on the offset values)
    auto build = [&](const std::vector<uint64_t> &begin) {
        Header h;
        h.b = {'C', 'D', 'F', 2};
        h.u32(0);  // numrecs: no record dimension
        h.u32(kNcDimension);
        h.u32(3);
        h.name("n_s");
        h.u32((uint32_t)n_s);
        h.name("n_a");
        h.u32((uint32_t)n_a);
        h.name("n_b");
        h.u32((uint32_t)n_b);
        std::vector<std::pair<std::string, std::string>> text = {
            {"title", "ConservativeRegridding weights (ESMF format)"},
            {"created_by", "crhip cr_regridder_write_esmf"},
            {"source_grid", at.src_grid_name ? at.src_grid_name : "source"},
            {"destination_grid", at.dst_grid_name ? at.dst_grid_name : "destination"},
            {"normalization", "destarea"},
            {"map_method", "Conservative remapping"}};
        if (at.created_at) text.push_back({"created_at", at.created_at});
        const uint32_t natt = (uint32_t)text.size() + (at.src_ndim > 0 ? 1 : 0) + (at.dst_ndim > 0 ? 1 : 0);
        h.u32(kNcAttribute);
        h.u32(natt);
        for (auto &kv : text) h.text_attr(kv.first, kv.second);
        if (at.src_ndim > 0) h.int_attr("source_grid_shape", std::vector<int64_t>(at.src_shape, at.src_shape + at.src_ndim));
        if (at.dst_ndim > 0)
            h.int_attr("destination_grid_shape", std::vector<int64_t>(at.dst_shape, at.dst_shape + at.dst_ndim));
        h.u32(kNcVariable);
        h.u32((uint32_t)vars.size());
        for (size_t v = 0; v < vars.size(); ++v) {
            h.name(vars[v].name);
            h.u32(1);
            h.u32(vars[v].dim);
            h.u32(kNcAttribute);
            h.u32(1);
            h.text_attr("long_name", vars[v].long_name);
            h.u32(vars[v].type);
            const uint64_t vs = (vars[v].bytes() + 3) / 4 * 4;
            h.u32(vs > 0xffffffffull ? 0xffffffffu : (uint32_t)vs);
            h.u64(begin[v]);
        }
        return h;
    };
    std::vector<uint64_t> begin(vars.size(), 0);
    uint64_t off = build(begin).b.size();
    for (size_t v = 0; v < vars.size(); ++v) {
        begin[v] = off;
        off += (vars[v].bytes() + 3) / 4 * 4;
    }
    const Header h = build(begin);
    FILE *f = fopen(path, "wb");
    CR_REQUIRE(f, CR_INVALID_ARGUMENT, std::string("esmf: cannot open ") + path);
    try {
        CR_REQUIRE(fwrite(h.b.data(), 1, h.b.size(), f) == h.b.size(), CR_INTERNAL, "esmf: short write");
        for (const Var &v : vars) {
            write_be(f, v.data, v.n, v.type == kNcDouble ? 8 : 4);
            const uint64_t padn = (4 - v.bytes() % 4) % 4;
            const unsigned char z[4] = {0, 0, 0, 0};
            if (padn) CR_REQUIRE(fwrite(z, 1, padn, f) == padn, CR_INTERNAL, "esmf: short write");
        }
    } catch (...) {
        fclose(f);
        throw;
    }
    CR_REQUIRE(fclose(f) == 0, CR_INTERNAL, "esmf: close failed");
}

}  // namespace cr

using namespace cr;

extern "C" {

cr_status cr_esmf_write(const char *path, int64_t n_s, const double *S, const int32_t *row, const int32_t *col,
                        int64_t n_a, const double *frac_a, const double *area_a, int64_t n_b,
                        const double *frac_b, const double *area_b, const cr_esmf_attrs *attrs) {
    return guard([&] {
        cr_esmf_attrs at{};
        if (attrs) at = *attrs;
        write_esmf_file(path, n_s, S, row, col, n_a, frac_a, area_a, n_b, frac_b, area_b, at);
    });
}

cr_status cr_regridder_write_esmf(const cr_regridder *R, int32_t transpose, const char *path,
                                  const cr_esmf_attrs *attrs) {
    return guard([&] {
        CR_REQUIRE(R && path, CR_INVALID_ARGUMENT, "cr_regridder_write_esmf: null argument");
        // this direction's A (rows = its destination cells) and areas
        const int64_t n_b = transpose ? R->fwd.ncols : R->fwd.nrows;
        const int64_t n_a = transpose ? R->fwd.nrows : R->fwd.ncols;
        const int64_t nnz = R->fwd.nnz;
        std::vector<double> a_dst(R->fwd.nrows), a_src(R->fwd.ncols);
        CR_HIP(hipMemcpy(a_dst.data(), R->a_dst.p, a_dst.size() * sizeof(double), hipMemcpyDeviceToHost));
        CR_HIP(hipMemcpy(a_src.data(), R->a_src.p, a_src.size() * sizeof(double), hipMemcpyDeviceToHost));
        const std::vector<double> &area_b = transpose ? a_src : a_dst, &area_a = transpose ? a_dst : a_src;
        // column-major order of this direction's A = rows of its transpose,
        // with ascending row indices inside each column
        std::vector<int64_t> rp(n_a + 1);
        std::vector<int32_t> ci(nnz);
        std::vector<double> cv(nnz);
        CR_REQUIRE(cr_regridder_copy_csr(R, transpose ? 0 : 1, rp.data(), ci.data(), cv.data()) == CR_OK,
                   CR_INTERNAL, cr_last_error());
        std::vector<double> S(nnz), frac_a(n_a, 0.0), frac_b(n_b, 0.0);
        std::vector<int32_t> row(nnz), col(nnz);
        for (int64_t j = 0; j < n_a; ++j)
            for (int64_t k = rp[j]; k < rp[j + 1]; ++k) {
                const int32_t i = ci[k];
                S[k] = cv[k] / area_b[i];
                row[k] = i + 1;
                col[k] = (int32_t)j + 1;
                frac_a[j] += cv[k];
                frac_b[i] += cv[k];
            }
        for (int64_t j = 0; j < n_a; ++j) frac_a[j] /= area_a[j];
        for (int64_t i = 0; i < n_b; ++i) frac_b[i] /= area_b[i];
        cr_esmf_attrs at{};
        if (attrs) at = *attrs;
        write_esmf_file(path, nnz, S.data(), row.data(), col.data(), n_a, frac_a.data(), area_a.data(), n_b,
                        frac_b.data(), area_b.data(), at);
    });
}

}  // extern "C"
