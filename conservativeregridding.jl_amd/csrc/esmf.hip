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
// netCDF classic format (published format specification "The NetCDF Classic
// Format Specification"): big-endian header of dimension, attribute and
// variable lists, then each variable's data at its recorded offset.  The
// default is CDF-5 (64-bit data), whose NC_INT64 row/col are the reference's
// Int64 indices (findnz of a SparseMatrixCSC{Float64,Int64}); CDF-2 (64-bit
// offsets, NC_INT row/col) is what netCDF-3 readers that predate CDF-5 take.
// The reader parses CDF-1, CDF-2 and CDF-5 headers and turns a weights file
// back into a device regridder: A = S .* area_b[row], rows = destination
// cells, so a build is reused without rebuilding.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "regridder.hpp"

namespace cr {
namespace {

constexpr uint32_t kNcByte = 1, kNcChar = 2, kNcShort = 3, kNcInt = 4, kNcFloat = 5, kNcDouble = 6,
                   kNcUbyte = 7, kNcUshort = 8, kNcUint = 9, kNcInt64 = 10, kNcUint64 = 11;
constexpr uint32_t kNcDimension = 0x0A, kNcVariable = 0x0B, kNcAttribute = 0x0C;

int type_size(uint32_t t) {
    switch (t) {
        case kNcByte: case kNcChar: case kNcUbyte: return 1;
        case kNcShort: case kNcUshort: return 2;
        case kNcInt: case kNcFloat: case kNcUint: return 4;
        case kNcDouble: case kNcInt64: case kNcUint64: return 8;
        default: return 0;
    }
}

// Header writer.  Counts, lengths and sizes are 32-bit in CDF-1/2 and 64-bit
// in CDF-5; data offsets are 64-bit in CDF-2 and CDF-5.
struct Header {
    int version = 5;
    std::vector<uint8_t> b;
    void u32(uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void u64(uint64_t v) {
        for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void count(uint64_t v) {
        if (version == 5) u64(v);
        else u32((uint32_t)v);
    }
    void pad() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string &s) {
        count(s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
    void text_attr(const std::string &n, const std::string &v) {
        name(n);
        u32(kNcChar);
        count(v.size());
        b.insert(b.end(), v.begin(), v.end());
        pad();
    }
    void int_attr(const std::string &n, const std::vector<int64_t> &v) {
        name(n);
        u32(version == 5 ? kNcInt64 : kNcInt);
        count(v.size());
        for (int64_t x : v) {
            if (version == 5) u64((uint64_t)x);
            else u32((uint32_t)(int32_t)x);
        }
    }
};

struct Var {
    std::string name, long_name;
    uint32_t dim;   // dimension id
    uint32_t type;  // kNcInt, kNcInt64 or kNcDouble
    int64_t n;
    const void *data;
    uint64_t bytes() const { return (uint64_t)n * type_size(type); }
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

void write_esmf_file(const char *path, int64_t n_s, const double *S, const int64_t *row, const int64_t *col,
                     int64_t n_a, const double *frac_a, const double *area_a, int64_t n_b,
                     const double *frac_b, const double *area_b, const cr_esmf_attrs &at) {
    const int version = at.cdf_version == 0 ? 5 : at.cdf_version;
    CR_REQUIRE(version == 2 || version == 5, CR_INVALID_ARGUMENT, "esmf: cdf_version must be 0, 2 or 5");
    CR_REQUIRE(path && n_s >= 0 && n_a > 0 && n_b > 0, CR_INVALID_ARGUMENT, "esmf: bad sizes");
    CR_REQUIRE(n_s == 0 || (S && row && col), CR_INVALID_ARGUMENT, "esmf: null weights");
    CR_REQUIRE(frac_a && area_a && frac_b && area_b, CR_INVALID_ARGUMENT, "esmf: null cell arrays");
    if (version == 2)
        CR_REQUIRE(n_s < ((int64_t)1 << 31) && n_a < ((int64_t)1 << 31) && n_b < ((int64_t)1 << 31),
                   CR_UNSUPPORTED, "esmf: CDF-2 dimensions and indices must fit 32 bits (use CDF-5)");
    // CDF-2 stores the indices as NC_INT
    std::vector<int32_t> row32, col32;
    if (version == 2) {
        for (int64_t k = 0; k < n_s; ++k)
            CR_REQUIRE(row[k] >= INT32_MIN && row[k] <= INT32_MAX && col[k] >= INT32_MIN && col[k] <= INT32_MAX,
                       CR_UNSUPPORTED, "esmf: a CDF-2 file stores row/col as 32-bit integers (use CDF-5)");
        row32.assign(row, row + n_s);
        col32.assign(col, col + n_s);
    }
    const uint32_t itype = version == 5 ? kNcInt64 : kNcInt;
    const void *rowp = version == 5 ? (const void *)row : (const void *)row32.data();
    const void *colp = version == 5 ? (const void *)col : (const void *)col32.data();
    const std::vector<Var> vars = {
        {"S", "weight value", 0, kNcDouble, n_s, S},
        {"row", "destination cell index (1-based)", 0, itype, n_s, rowp},
        {"col", "source cell index (1-based)", 0, itype, n_s, colp},
        {"frac_a", "source cell fraction covered", 1, kNcDouble, n_a, frac_a},
        {"frac_b", "destination cell fraction covered", 2, kNcDouble, n_b, frac_b},
        {"area_a", "source cell areas", 1, kNcDouble, n_a, area_a},
        {"area_b", "destination cell areas", 2, kNcDouble, n_b, area_b},
    };
    // header with placeholder offsets, then the real ones (the header size
    // does not depend on the offset values)
    auto build = [&](const std::vector<uint64_t> &begin) {
        Header h;
        h.version = version;
        h.b = {'C', 'D', 'F', (uint8_t)version};
        h.count(0);  // numrecs: no record dimension
        h.u32(kNcDimension);
        h.count(3);
        h.name("n_s");
        h.count((uint64_t)n_s);
        h.name("n_a");
        h.count((uint64_t)n_a);
        h.name("n_b");
        h.count((uint64_t)n_b);
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
        h.count(natt);
        for (auto &kv : text) h.text_attr(kv.first, kv.second);
        if (at.src_ndim > 0) h.int_attr("source_grid_shape", std::vector<int64_t>(at.src_shape, at.src_shape + at.src_ndim));
        if (at.dst_ndim > 0)
            h.int_attr("destination_grid_shape", std::vector<int64_t>(at.dst_shape, at.dst_shape + at.dst_ndim));
        h.u32(kNcVariable);
        h.count(vars.size());
        for (size_t v = 0; v < vars.size(); ++v) {
            h.name(vars[v].name);
            h.count(1);
            h.count(vars[v].dim);
            h.u32(kNcAttribute);
            h.count(1);
            h.text_attr("long_name", vars[v].long_name);
            h.u32(vars[v].type);
            const uint64_t vs = (vars[v].bytes() + 3) / 4 * 4;
            if (version == 5) h.u64(vs);
            else h.u32(vs > 0xffffffffull ? 0xffffffffu : (uint32_t)vs);
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
            write_be(f, v.data, v.n, type_size(v.type));
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

// ---- reader -----------------------------------------------------------------
struct NcVar {
    std::string name;
    uint32_t type = 0;
    std::vector<uint64_t> dims;
    uint64_t begin = 0;
};

struct NcFile {
    int version = 0;
    std::vector<std::pair<std::string, uint64_t>> dims;
    std::vector<NcVar> vars;
};

// Big-endian reads from the header of an open file.
struct Cursor {
    FILE *f;
    int version;
    uint64_t u(int n) {
        unsigned char b[8];
        CR_REQUIRE(fread(b, 1, (size_t)n, f) == (size_t)n, CR_INVALID_ARGUMENT, "esmf: truncated netCDF header");
        uint64_t v = 0;
        for (int i = 0; i < n; ++i) v = (v << 8) | b[i];
        return v;
    }
    uint32_t u32() { return (uint32_t)u(4); }
    uint64_t count() { return u(version == 5 ? 8 : 4); }
    void skip(uint64_t n) {
        CR_REQUIRE(fseeko(f, (off_t)n, SEEK_CUR) == 0, CR_INVALID_ARGUMENT, "esmf: truncated netCDF header");
    }
    std::string name() {
        const uint64_t n = count();
        CR_REQUIRE(n <= 4096, CR_INVALID_ARGUMENT, "esmf: malformed netCDF name");
        std::string s(n, '\0');
        if (n) CR_REQUIRE(fread(&s[0], 1, n, f) == n, CR_INVALID_ARGUMENT, "esmf: truncated netCDF header");
        skip((4 - n % 4) % 4);
        return s;
    }
    void attributes() {  // an attribute list, skipped
        const uint32_t tag = u32();
        const uint64_t n = count();
        CR_REQUIRE(tag == kNcAttribute || (tag == 0 && n == 0), CR_INVALID_ARGUMENT, "esmf: malformed attribute list");
        for (uint64_t a = 0; a < n; ++a) {
            name();
            const uint32_t t = u32();
            const uint64_t m = count();
            CR_REQUIRE(type_size(t) > 0, CR_INVALID_ARGUMENT, "esmf: unknown netCDF type");
            const uint64_t bytes = m * (uint64_t)type_size(t);
            skip((bytes + 3) / 4 * 4);
        }
    }
};

NcFile read_header(FILE *f) {
    unsigned char magic[4];
    CR_REQUIRE(fread(magic, 1, 4, f) == 4 && magic[0] == 'C' && magic[1] == 'D' && magic[2] == 'F',
               CR_INVALID_ARGUMENT, "esmf: not a netCDF classic file");
    NcFile nc;
    nc.version = magic[3];
    CR_REQUIRE(nc.version == 1 || nc.version == 2 || nc.version == 5, CR_UNSUPPORTED,
               "esmf: netCDF classic version must be 1, 2 or 5 (netCDF-4/HDF5 files need the netCDF library)");
    Cursor c{f, nc.version};
    c.count();  // numrecs
    uint32_t tag = c.u32();
    uint64_t n = c.count();
    CR_REQUIRE(tag == kNcDimension || (tag == 0 && n == 0), CR_INVALID_ARGUMENT, "esmf: malformed dimension list");
    for (uint64_t d = 0; d < n; ++d) {
        std::string nm = c.name();
        nc.dims.push_back({nm, c.count()});
    }
    c.attributes();
    tag = c.u32();
    n = c.count();
    CR_REQUIRE(tag == kNcVariable || (tag == 0 && n == 0), CR_INVALID_ARGUMENT, "esmf: malformed variable list");
    for (uint64_t v = 0; v < n; ++v) {
        NcVar var;
        var.name = c.name();
        const uint64_t nd = c.count();
        CR_REQUIRE(nd <= 64, CR_INVALID_ARGUMENT, "esmf: malformed variable");
        for (uint64_t d = 0; d < nd; ++d) {
            const uint64_t id = c.count();
            CR_REQUIRE(id < nc.dims.size(), CR_INVALID_ARGUMENT, "esmf: bad dimension id");
            var.dims.push_back(id);
        }
        c.attributes();
        var.type = c.u32();
        CR_REQUIRE(type_size(var.type) > 0, CR_INVALID_ARGUMENT, "esmf: unknown netCDF type");
        c.count();  // vsize
        var.begin = c.u(nc.version == 1 ? 4 : 8);
        nc.vars.push_back(var);
    }
    return nc;
}

// A 1-D variable converted to T (double or int64_t); n values expected.
template <class T>
void read_var(FILE *f, const NcFile &nc, const char *name, int64_t n, T *out) {
    const NcVar *v = nullptr;
    for (const NcVar &x : nc.vars)
        if (x.name == name) v = &x;
    CR_REQUIRE(v, CR_INVALID_ARGUMENT, std::string("esmf: the file has no variable ") + name);
    CR_REQUIRE(v->dims.size() == 1, CR_INVALID_ARGUMENT, std::string("esmf: variable ") + name + " is not 1-D");
    const uint64_t len = nc.dims[v->dims[0]].second;
    CR_REQUIRE(len != 0 || n == 0, CR_UNSUPPORTED, std::string("esmf: record variable ") + name + " is not supported");
    CR_REQUIRE((int64_t)len == n, CR_DIMENSION_MISMATCH, std::string("esmf: variable ") + name + " has the wrong length");
    if (!out || n == 0) return;
    const int size = type_size(v->type);
    CR_REQUIRE(v->type != kNcChar, CR_INVALID_ARGUMENT, std::string("esmf: variable ") + name + " is text");
    CR_REQUIRE(fseeko(f, (off_t)v->begin, SEEK_SET) == 0, CR_INVALID_ARGUMENT, "esmf: bad variable offset");
    std::vector<unsigned char> blk((size_t)std::min<int64_t>(n, 1 << 20) * size);
    for (int64_t i0 = 0; i0 < n; i0 += 1 << 20) {
        const int64_t m = std::min<int64_t>(n - i0, 1 << 20);
        CR_REQUIRE(fread(blk.data(), 1, (size_t)(m * size), f) == (size_t)(m * size), CR_INVALID_ARGUMENT,
                   std::string("esmf: truncated data of ") + name);
        for (int64_t i = 0; i < m; ++i) {
            uint64_t bits = 0;
            for (int k = 0; k < size; ++k) bits = (bits << 8) | blk[i * size + k];
            T x;
            switch (v->type) {
                case kNcDouble: { double d; std::memcpy(&d, &bits, 8); x = (T)d; break; }
                case kNcFloat: { uint32_t b32 = (uint32_t)bits; float g; std::memcpy(&g, &b32, 4); x = (T)g; break; }
                case kNcByte: x = (T)(int8_t)bits; break;
                case kNcShort: x = (T)(int16_t)bits; break;
                case kNcInt: x = (T)(int32_t)bits; break;
                case kNcInt64: x = (T)(int64_t)bits; break;
                default: x = (T)bits; break;  // unsigned types
            }
            out[i0 + i] = x;
        }
    }
}

struct EsmfSizes {
    int64_t n_s, n_a, n_b;
};

EsmfSizes esmf_sizes(const NcFile &nc) {
    EsmfSizes z{-1, -1, -1};
    for (auto &d : nc.dims) {
        if (d.first == "n_s") z.n_s = (int64_t)d.second;
        if (d.first == "n_a") z.n_a = (int64_t)d.second;
        if (d.first == "n_b") z.n_b = (int64_t)d.second;
    }
    CR_REQUIRE(z.n_s >= 0 && z.n_a > 0 && z.n_b > 0, CR_INVALID_ARGUMENT,
               "esmf: the file lacks the dimensions n_s, n_a, n_b of an ESMF weights file");
    return z;
}

struct File {
    FILE *f;
    explicit File(const char *path) : f(fopen(path, "rb")) {
        CR_REQUIRE(f, CR_INVALID_ARGUMENT, std::string("esmf: cannot open ") + path);
    }
    ~File() { fclose(f); }
};

}  // namespace cr

using namespace cr;

extern "C" {

cr_status cr_esmf_read_sizes(const char *path, int64_t *n_s, int64_t *n_a, int64_t *n_b) {
    return guard([&] {
        CR_REQUIRE(path && n_s && n_a && n_b, CR_INVALID_ARGUMENT, "cr_esmf_read_sizes: null argument");
        File file(path);
        const EsmfSizes z = esmf_sizes(read_header(file.f));
        *n_s = z.n_s;
        *n_a = z.n_a;
        *n_b = z.n_b;
    });
}

cr_status cr_esmf_read(const char *path, double *S, int64_t *row, int64_t *col, double *frac_a, double *area_a,
                       double *frac_b, double *area_b) {
    return guard([&] {
        CR_REQUIRE(path, CR_INVALID_ARGUMENT, "cr_esmf_read: null path");
        File file(path);
        const NcFile nc = read_header(file.f);
        const EsmfSizes z = esmf_sizes(nc);
        if (S) read_var(file.f, nc, "S", z.n_s, S);
        if (row) read_var(file.f, nc, "row", z.n_s, row);
        if (col) read_var(file.f, nc, "col", z.n_s, col);
        if (frac_a) read_var(file.f, nc, "frac_a", z.n_a, frac_a);
        if (area_a) read_var(file.f, nc, "area_a", z.n_a, area_a);
        if (frac_b) read_var(file.f, nc, "frac_b", z.n_b, frac_b);
        if (area_b) read_var(file.f, nc, "area_b", z.n_b, area_b);
    });
}

cr_status cr_regridder_from_esmf(cr_regridder **R, const char *path, void *stream) {
    return guard([&] {
        CR_REQUIRE(R && path, CR_INVALID_ARGUMENT, "cr_regridder_from_esmf: null argument");
        std::vector<double> S, area_a, area_b;
        std::vector<int64_t> row, col;
        int64_t n_s, n_a, n_b;
        {
            File file(path);
            const NcFile nc = read_header(file.f);
            const EsmfSizes z = esmf_sizes(nc);
            n_s = z.n_s;
            n_a = z.n_a;
            n_b = z.n_b;
            S.resize(n_s);
            row.resize(n_s);
            col.resize(n_s);
            area_a.resize(n_a);
            area_b.resize(n_b);
            read_var(file.f, nc, "S", n_s, S.data());
            read_var(file.f, nc, "row", n_s, row.data());
            read_var(file.f, nc, "col", n_s, col.data());
            read_var(file.f, nc, "area_a", n_a, area_a.data());
            read_var(file.f, nc, "area_b", n_b, area_b.data());
        }
        CR_REQUIRE(n_a < ((int64_t)1 << 31) && n_b < ((int64_t)1 << 31), CR_UNSUPPORTED,
                   "cr_regridder_from_esmf: more than 2^31-1 cells");
        // CSR(A) with A_ij = S_k * area_b[i]: a stable counting sort by row keeps the
        // file's column-major order, so columns ascend inside each row
        std::vector<int64_t> rp(n_b + 1, 0);
        for (int64_t k = 0; k < n_s; ++k) {
            CR_REQUIRE(row[k] >= 1 && row[k] <= n_b && col[k] >= 1 && col[k] <= n_a, CR_DIMENSION_MISMATCH,
                       "cr_regridder_from_esmf: row or col index out of range at entry " + std::to_string(k));
            rp[row[k]]++;
        }
        for (int64_t i = 0; i < n_b; ++i) rp[i + 1] += rp[i];
        std::vector<int64_t> pos(rp.begin(), rp.end() - 1);
        std::vector<int32_t> ci(n_s);
        std::vector<double> cv(n_s);
        for (int64_t k = 0; k < n_s; ++k) {
            const int64_t i = row[k] - 1, p = pos[i]++;
            ci[p] = (int32_t)(col[k] - 1);
            cv[p] = S[k] * area_b[i];
        }
        const cr_status st = cr_regridder_from_csr(R, n_b, n_a, n_s, rp.data(), ci.data(), cv.data(), area_b.data(),
                                                   area_a.data(), stream);
        if (st != CR_OK) throw Error{st, cr_last_error()};
    });
}

cr_status cr_esmf_write(const char *path, int64_t n_s, const double *S, const int64_t *row, const int64_t *col,
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
        std::vector<int64_t> row(nnz), col(nnz);
        for (int64_t j = 0; j < n_a; ++j)
            for (int64_t k = rp[j]; k < rp[j + 1]; ++k) {
                const int32_t i = ci[k];
                S[k] = cv[k] / area_b[i];
                row[k] = (int64_t)i + 1;
                col[k] = j + 1;
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
