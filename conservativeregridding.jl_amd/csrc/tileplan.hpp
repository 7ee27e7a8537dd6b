// tileplan.hpp -- row tiles of a CSR matrix for the staged apply (host code,
// shared by the library and host-side checks).
//
// A tile is a contiguous range of rows with at most kTileRows rows,
// kTileCols distinct columns and kTileThreads * kTileEnt non-zeros.  Its
// non-zeros (row-major) are dealt out in equal runs of E = ceil(nnz/threads)
// to the workgroup's threads; a run of one row inside one thread is a
// segment.  Rows with more than kTileCols non-zeros ("huge") stay in their
// tile with no entries and are listed for another kernel.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace cr {

constexpr int kTileThreads = 256;
constexpr int kTileEnt = 16;
constexpr int kTileCols = 1024;
constexpr int kTileRows = 512;
constexpr int kTileSegs = kTileRows + kTileThreads;
constexpr int kColsPerThread = kTileCols / kTileThreads;
constexpr int kRowsPerThread = kTileRows / kTileThreads;
static_assert(kTileCols <= 65536 && kTileSegs < 65535, "local indices are 16 bits");

struct TileDesc {
    int64_t row0;     // first row
    int64_t ent_off;  // first entry slot (slots are entry-major: e * kTileThreads + thread)
    int64_t col_off;  // first distinct column in the column list
    int32_t nrows;    // rows in the tile
    int32_t ncols;    // distinct columns
    int32_t ent;      // entries per thread (slots per thread)
    int32_t nseg;     // row segments; segment id nseg is the padding sink
};

struct HostTilePlan {
    std::vector<TileDesc> desc;
    std::vector<uint32_t> key;   // per slot: local column | segment id << 16
    std::vector<float> v32;
    std::vector<double> v64;
    std::vector<int32_t> cols;   // distinct columns per tile, ascending
    std::vector<uint16_t> rseg;  // per tile nrows + 1: first segment of each row
    std::vector<int32_t> huge;
};

inline HostTilePlan make_host_tile_plan(int64_t nrows, int64_t ncols, const int64_t *rp,
                                        const int32_t *col, const double *val) {
    // pass 1: greedy tile boundaries (rows in order)
    struct Span {
        int64_t row0;
        int32_t nrows, nnz, ncols;
    };
    std::vector<Span> spans;
    std::vector<int32_t> huge;
    std::vector<int32_t> stamp(ncols, -1);
    const int64_t max_ent = (int64_t)kTileThreads * kTileEnt;
    int64_t r = 0;
    while (r < nrows) {
        const int32_t t = (int32_t)spans.size();
        Span cur{r, 0, 0, 0};
        while (r < nrows && cur.nrows < kTileRows) {
            const int64_t k0 = rp[r], k1 = rp[r + 1], n = k1 - k0;
            if (n > kTileCols) {  // a row alone over the column budget
                huge.push_back((int32_t)r);
                ++cur.nrows;
                ++r;
                continue;
            }
            if (cur.nnz + n > max_ent) break;
            int32_t nu = 0;
            for (int64_t k = k0; k < k1; ++k) nu += stamp[col[k]] != t;
            if (cur.ncols + nu > kTileCols) break;
            for (int64_t k = k0; k < k1; ++k) stamp[col[k]] = t;
            cur.ncols += nu;
            cur.nnz += (int32_t)n;
            ++cur.nrows;
            ++r;
        }
        spans.push_back(cur);
    }
    const int64_t ntiles = (int64_t)spans.size();
    std::vector<TileDesc> desc(ntiles);
    int64_t ent_total = 0, col_total = 0;
    for (int64_t t = 0; t < ntiles; ++t) {
        const Span &sp = spans[t];
        TileDesc &d = desc[t];
        d.row0 = sp.row0;
        d.nrows = sp.nrows;
        d.ent = (sp.nnz + kTileThreads - 1) / kTileThreads;
        d.ent_off = ent_total;
        d.col_off = col_total;
        ent_total += (int64_t)d.ent * kTileThreads;
        col_total += sp.ncols;  // capacity; the exact count is set in pass 2
    }
    std::vector<uint32_t> key(ent_total);
    std::vector<float> v32(ent_total);
    std::vector<double> v64(ent_total);
    std::vector<int32_t> ucols(col_total);
    std::vector<uint16_t> rseg(nrows + ntiles);

    // pass 2: per tile, distinct columns, merge-path split, segments
    auto fill = [&](int64_t t0, int64_t t1) {
        std::vector<int32_t> u;
        for (int64_t t = t0; t < t1; ++t) {
            TileDesc &d = desc[t];
            u.clear();
            for (int64_t rr = d.row0; rr < d.row0 + d.nrows; ++rr)
                if (rp[rr + 1] - rp[rr] <= kTileCols)
                    for (int64_t k = rp[rr]; k < rp[rr + 1]; ++k) u.push_back(col[k]);
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            d.ncols = (int32_t)u.size();
            std::copy(u.begin(), u.end(), ucols.begin() + d.col_off);
            const int64_t E = d.ent;
            int64_t q = 0;
            int32_t nseg = 0, cur = -1;
            int64_t prev_thread = -1;
            uint16_t *rs = rseg.data() + d.row0 + t;
            for (int32_t lr = 0; lr < d.nrows; ++lr) {
                const int64_t rr = d.row0 + lr;
                rs[lr] = (uint16_t)nseg;
                if (rp[rr + 1] - rp[rr] > kTileCols) continue;
                bool first = true;
                for (int64_t k = rp[rr]; k < rp[rr + 1]; ++k, ++q) {
                    const int64_t thr = q / E, e = q % E;
                    if (first || thr != prev_thread) {
                        cur = nseg++;
                        first = false;
                        prev_thread = thr;
                    }
                    const int32_t lc = (int32_t)(std::lower_bound(u.begin(), u.end(), col[k]) - u.begin());
                    const int64_t slot = d.ent_off + e * kTileThreads + thr;
                    key[slot] = (uint32_t)lc | ((uint32_t)cur << 16);
                    v32[slot] = (float)val[k];
                    v64[slot] = val[k];
                }
            }
            rs[d.nrows] = (uint16_t)nseg;
            d.nseg = nseg;
            for (; q < E * kTileThreads; ++q) {  // padding: value 0, sink segment
                const int64_t thr = q / E, e = q % E;
                const int64_t slot = d.ent_off + e * kTileThreads + thr;
                key[slot] = (uint32_t)nseg << 16;
                v32[slot] = 0.f;
                v64[slot] = 0.0;
            }
        }
    };
    {
        const int nth = (int)std::max<int64_t>(1, std::min<int64_t>(16, ntiles / 64));
        std::vector<std::thread> th;
        const int64_t per = (ntiles + nth - 1) / nth;
        for (int i = 0; i < nth; ++i) {
            const int64_t t0 = i * per, t1 = std::min(ntiles, t0 + per);
            if (t0 < t1) th.emplace_back(fill, t0, t1);
        }
        for (auto &x : th) x.join();
    }

    HostTilePlan H;
    H.desc = std::move(desc);
    H.key = std::move(key);
    H.v32 = std::move(v32);
    H.v64 = std::move(v64);
    H.cols = std::move(ucols);
    H.rseg = std::move(rseg);
    H.huge = std::move(huge);
    return H;
}

}  // namespace cr
