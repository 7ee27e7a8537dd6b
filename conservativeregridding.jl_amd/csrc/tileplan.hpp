// tileplan.hpp -- row tiles of a CSR matrix for the staged apply (host code,
// shared by the library and host-side checks).
//
// A tile is a contiguous range of rows with at most kRows rows and kCols
// distinct columns (TileGeom).  Two layouts of its non-zeros over the
// workgroup's threads:
//
// kRounds: rows are cut into pieces of at most kPieceLen non-zeros (a row of
// n non-zeros gives ceil(n / kPieceLen) pieces of near-equal length); a piece
// is a *segment*: the kernel sums it in one accumulator and writes the sum to
// LDS slot `segment id`.  Pieces are sorted by length (longest first) and
// dealt out in rounds of kThreads: round k gives piece k*kThreads + t to thread
// t.  In every round, the 64 pieces of one wave are padded to the longest of
// them, so all lanes of a wave end their segments on the same slot: the
// end-of-segment test is a wave-uniform bit mask (TileDesc::flush), not a
// per-lane branch.
//
// kEven: the non-zeros (row-major) are dealt out in equal runs; see
// make_even_tile_plan.
//
// Either way a thread's slots (at most kEnt) stay in registers for the whole
// launch.  Rows with more than kCols non-zeros ("huge") stay in their tile
// with no entries and are listed for another kernel.
#pragma once

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <thread>
#include <vector>

namespace cr {

// Tile geometry: threads per workgroup, rows and distinct columns per tile,
// segments (pieces) per tile; every thread holds at most kEnt entry slots.
template <int THREADS, int ROWS, int COLS, int SEGS>
struct TileGeom {
    static constexpr int kThreads = THREADS;
    static constexpr int kWaves = THREADS / 64;
    static constexpr int kEnt = 16;
    static constexpr int kCols = COLS;
    static constexpr int kRows = ROWS;
    static constexpr int kSegs = SEGS;
    static constexpr int kColsPerThread = COLS / THREADS;
    static constexpr int kRowsPerThread = ROWS / THREADS;
    static_assert(COLS <= 65536 && SEGS < 65535, "local indices are 16 bits");
    static_assert(kEnt <= 32, "flush masks are 32 bits");
    static_assert(SEGS >= ROWS + THREADS, "an even split makes up to ROWS + THREADS segments");
};
using GeomS = TileGeom<256, 512, 1024, 1024>;   // 4 waves, 512-row tiles
using GeomL = TileGeom<512, 1024, 2048, 1536>;  // 8 waves, 1024-row tiles
enum TileGeomId { kGeomS = 0, kGeomL = 1 };
constexpr int kMaxTileWaves = 8;
constexpr int kPieceLen = 8;
constexpr int kRowAlign = 32;  // rows per 128-byte line of a FP32 field

struct TileDesc {
    int64_t row0;     // first row
    int64_t ent_off;  // first entry slot (slots are entry-major: e * kThreads + thread)
    int64_t col_off;  // first distinct column in the column list
    int32_t nrows;    // rows in the tile
    int32_t ncols;    // distinct columns
    int32_t ent;      // entry slots per thread
    int32_t nseg;     // segments (pieces); segment id nseg is the padding sink
    uint32_t flush[kMaxTileWaves];  // per wave: bit e set = slot e ends a segment
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

// Number of pieces a row of n non-zeros is cut into, and the length of piece j.
inline int piece_count(int64_t n) { return (int)((n + kPieceLen - 1) / kPieceLen); }
inline int piece_len(int64_t n, int j) {
    const int p = piece_count(n);
    return (int)(n / p + (j < n % p ? 1 : 0));
}

// Slots per thread of a tile whose pieces have the given length histogram
// (hist[L] = pieces of length L, 1 <= L <= kPieceLen): the pieces sorted
// longest first, dealt in rounds of kThreads; per wave and round the slot
// count is the longest piece of that wave's 64.  Returns the largest per-wave
// total (the tile's slots per thread).
template <class G>
inline int tile_slots(const int *hist) {
    int npieces = 0;
    for (int L = 1; L <= kPieceLen; ++L) npieces += hist[L];
    int worst = 0;
    for (int w = 0; w < G::kWaves; ++w) {
        int total = 0;
        for (int start = w * 64; start < npieces; start += G::kThreads) {
            // length of the piece at sorted position `start` (descending order)
            int pos = start, L = kPieceLen;
            while (L > 0 && pos >= hist[L]) pos -= hist[L--];
            total += L;
        }
        worst = std::max(worst, total);
    }
    return worst;
}

template <class G>
inline HostTilePlan make_rounds_tile_plan(int64_t nrows, int64_t ncols, const int64_t *rp,
                                          const int32_t *col, const double *val) {
    // pass 1: greedy tile boundaries (rows in order)
    struct Span {
        int64_t row0;
        int32_t nrows, ncols;
    };
    std::vector<Span> spans;
    std::vector<int32_t> huge;
    std::vector<int32_t> stamp(ncols, -1);
    int64_t r = 0;
    while (r < nrows) {
        const int32_t t = (int32_t)spans.size();
        Span cur{r, 0, 0};
        int hist[kPieceLen + 1] = {0};
        int npieces = 0;
        while (r < nrows && cur.nrows < G::kRows) {
            const int64_t k0 = rp[r], k1 = rp[r + 1], n = k1 - k0;
            if (n > G::kCols) {  // a row alone over the column budget
                huge.push_back((int32_t)r);
                ++cur.nrows;
                ++r;
                continue;
            }
            const int p = piece_count(n);
            if (npieces + p > G::kSegs) break;
            int h2[kPieceLen + 1];
            std::copy(hist, hist + kPieceLen + 1, h2);
            for (int j = 0; j < p; ++j) ++h2[piece_len(n, j)];
            if (tile_slots<G>(h2) > G::kEnt) break;
            int32_t nu = 0;
            for (int64_t k = k0; k < k1; ++k) nu += stamp[col[k]] != t;
            if (cur.ncols + nu > G::kCols) break;
            for (int64_t k = k0; k < k1; ++k) stamp[col[k]] = t;
            std::copy(h2, h2 + kPieceLen + 1, hist);
            npieces += p;
            cur.ncols += nu;
            ++cur.nrows;
            ++r;
        }
        // end a full tile on a multiple of kRowAlign rows when it can give rows
        // back, so no 128-byte line of a FP32 output field is shared by two
        // tiles (the column count stays an upper bound)
        const int back = (int)(r % kRowAlign);
        if (r < nrows && back && cur.nrows > back) {
            r -= back;
            cur.nrows -= back;
            while (!huge.empty() && huge.back() >= r) huge.pop_back();  // listed again later
        }
        spans.push_back(cur);
    }
    const int64_t ntiles = (int64_t)spans.size();
    std::vector<TileDesc> desc(ntiles);
    // slots per thread of each tile (needs the piece lengths: recomputed here)
    int64_t ent_total = 0, col_total = 0;
    for (int64_t t = 0; t < ntiles; ++t) {
        const Span &sp = spans[t];
        int hist[kPieceLen + 1] = {0};
        for (int64_t rr = sp.row0; rr < sp.row0 + sp.nrows; ++rr) {
            const int64_t n = rp[rr + 1] - rp[rr];
            if (n > G::kCols) continue;
            for (int j = 0; j < piece_count(n); ++j) ++hist[piece_len(n, j)];
        }
        TileDesc &d = desc[t];
        d.row0 = sp.row0;
        d.nrows = sp.nrows;
        d.ent = tile_slots<G>(hist);
        d.ent_off = ent_total;
        d.col_off = col_total;
        ent_total += (int64_t)d.ent * G::kThreads;
        col_total += sp.ncols;  // capacity; the exact count is set in pass 2
    }
    std::vector<uint32_t> key(ent_total);
    std::vector<float> v32(ent_total);
    std::vector<double> v64(ent_total);
    std::vector<int32_t> ucols(col_total);
    std::vector<uint16_t> rseg(nrows + ntiles);

    // pass 2: per tile, distinct columns, pieces, rounds, slots
    auto fill = [&](int64_t t0, int64_t t1) {
        std::vector<int32_t> u;
        struct Piece {
            int64_t k0;  // first non-zero
            int32_t len, sid;
        };
        std::vector<Piece> pieces;
        std::vector<int32_t> order;
        for (int64_t t = t0; t < t1; ++t) {
            TileDesc &d = desc[t];
            u.clear();
            pieces.clear();
            uint16_t *rs = rseg.data() + d.row0 + t;
            for (int32_t lr = 0; lr < d.nrows; ++lr) {
                const int64_t rr = d.row0 + lr;
                const int64_t n = rp[rr + 1] - rp[rr];
                rs[lr] = (uint16_t)pieces.size();
                if (n > G::kCols) continue;
                for (int64_t k = rp[rr]; k < rp[rr + 1]; ++k) u.push_back(col[k]);
                int64_t k = rp[rr];
                for (int j = 0; j < piece_count(n); ++j) {
                    const int len = piece_len(n, j);
                    pieces.push_back(Piece{k, len, (int32_t)pieces.size()});
                    k += len;
                }
            }
            const int32_t nseg = (int32_t)pieces.size();
            rs[d.nrows] = (uint16_t)nseg;
            d.nseg = nseg;
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            d.ncols = (int32_t)u.size();
            std::copy(u.begin(), u.end(), ucols.begin() + d.col_off);
            // longest first; ties keep row order
            order.resize(nseg);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(),
                             [&](int32_t a, int32_t b) { return pieces[a].len > pieces[b].len; });
            // every slot starts as padding of the sink segment
            for (int64_t q = 0; q < (int64_t)d.ent * G::kThreads; ++q) {
                key[d.ent_off + q] = (uint32_t)nseg << 16;
                v32[d.ent_off + q] = 0.f;
                v64[d.ent_off + q] = 0.0;
            }
            for (int w = 0; w < G::kWaves; ++w) {
                uint32_t fm = 0u;
                int off = 0;  // slots used by this wave's earlier rounds
                for (int start = w * 64; start < nseg; start += G::kThreads) {
                    const int L = pieces[order[start]].len;  // longest of this wave's 64
                    for (int l = 0; l < 64 && start + l < nseg; ++l) {
                        const Piece &pc = pieces[order[start + l]];
                        const int thr = w * 64 + l;
                        for (int e = 0; e < L; ++e) {
                            const int64_t slot = d.ent_off + (int64_t)(off + e) * G::kThreads + thr;
                            uint32_t lc = 0;
                            double vv = 0.0;
                            if (e < pc.len) {
                                const int64_t k = pc.k0 + e;
                                lc = (uint32_t)(std::lower_bound(u.begin(), u.end(), col[k]) - u.begin());
                                vv = val[k];
                            }
                            key[slot] = lc | ((uint32_t)pc.sid << 16);
                            v32[slot] = (float)vv;
                            v64[slot] = vv;
                        }
                    }
                    fm |= 1u << (off + L - 1);
                    off += L;
                }
                d.flush[w] = fm;
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

// Even split: a tile's non-zeros (row-major) are dealt out in equal runs of
// E = ceil(nnz / threads) to the threads, cutting rows anywhere; a run of one
// row inside one thread is a segment, and each lane ends its segments where
// its own run changes row (a per-lane test in the kernel: TileDesc::flush is
// unused).  No padding beyond the last run: the cheaper layout when few fields
// share the per-tile set-up.
template <class G>
inline HostTilePlan make_even_tile_plan(int64_t nrows, int64_t ncols, const int64_t *rp,
                                        const int32_t *col, const double *val) {
    // pass 1: greedy tile boundaries (rows in order)
    struct Span {
        int64_t row0;
        int32_t nrows, nnz, ncols;
    };
    std::vector<Span> spans;
    std::vector<int32_t> huge;
    std::vector<int32_t> stamp(ncols, -1);
    const int64_t max_ent = (int64_t)G::kThreads * G::kEnt;
    int64_t r = 0;
    while (r < nrows) {
        const int32_t t = (int32_t)spans.size();
        Span cur{r, 0, 0, 0};
        while (r < nrows && cur.nrows < G::kRows) {
            const int64_t k0 = rp[r], k1 = rp[r + 1], n = k1 - k0;
            if (n > G::kCols) {  // a row alone over the column budget
                huge.push_back((int32_t)r);
                ++cur.nrows;
                ++r;
                continue;
            }
            if (cur.nnz + n > max_ent) break;
            int32_t nu = 0;
            for (int64_t k = k0; k < k1; ++k) nu += stamp[col[k]] != t;
            if (cur.ncols + nu > G::kCols) break;
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
        d.ent = (sp.nnz + G::kThreads - 1) / G::kThreads;
        d.ent_off = ent_total;
        d.col_off = col_total;
        for (int w = 0; w < kMaxTileWaves; ++w) d.flush[w] = 0u;
        ent_total += (int64_t)d.ent * G::kThreads;
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
                if (rp[rr + 1] - rp[rr] <= G::kCols)
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
                if (rp[rr + 1] - rp[rr] > G::kCols) continue;
                bool first = true;
                for (int64_t k = rp[rr]; k < rp[rr + 1]; ++k, ++q) {
                    const int64_t thr = q / E, e = q % E;
                    if (first || thr != prev_thread) {
                        cur = nseg++;
                        first = false;
                        prev_thread = thr;
                    }
                    const int32_t lc = (int32_t)(std::lower_bound(u.begin(), u.end(), col[k]) - u.begin());
                    const int64_t slot = d.ent_off + e * G::kThreads + thr;
                    key[slot] = (uint32_t)lc | ((uint32_t)cur << 16);
                    v32[slot] = (float)val[k];
                    v64[slot] = val[k];
                }
            }
            rs[d.nrows] = (uint16_t)nseg;
            d.nseg = nseg;
            for (; q < E * G::kThreads; ++q) {  // padding: value 0, sink segment
                const int64_t thr = q / E, e = q % E;
                const int64_t slot = d.ent_off + e * G::kThreads + thr;
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

// Plan layouts: kRounds (wave-uniform segment ends, for many fields) and kEven
// (no padding, for one or two fields).
enum TileStyle { kRounds = 0, kEven = 1 };

template <class G>
inline HostTilePlan make_host_tile_plan(int style, int64_t nrows, int64_t ncols, const int64_t *rp,
                                        const int32_t *col, const double *val) {
    return style == kEven ? make_even_tile_plan<G>(nrows, ncols, rp, col, val)
                          : make_rounds_tile_plan<G>(nrows, ncols, rp, col, val);
}

}  // namespace cr
