// cr_tileplan.cpp -- host restatement of the library's tile plans (the layout
// of a CSR matrix that the staged batched apply runs on), the checker of the
// device planner.
//
// TEST INFRASTRUCTURE ONLY: only tests/ load it.  The reference has no
// counterpart: its batched regrid! is a loop of sparse matrix-vector products;
// the plan is this library's own layout, so the check here is that the device
// planner lays out every slot exactly as this plain sequential statement does.
//
// Plan (the same definitions the device code states):
//   * tiles: contiguous rows, cut greedily in row order.  A tile takes rows
//     while it has at most kRows rows and its budgets hold; a row with more
//     than kCols non-zeros ("huge") joins with no entries and is listed apart.
//     The tile then gives rows back down to a multiple of kRowAlign when that
//     leaves it non-empty (the rows go to the next tile).  The distinct-column
//     count of the greedy pass counts a column once per row that brings it
//     first (a column twice in one row counts twice); it is the tile's column
//     capacity.
//   * kRounds: rows cut into pieces of at most 8 non-zeros (ceil(n/8) pieces of
//     near-equal length, the longer ones first), sorted longest first (ties in
//     row order), dealt in rounds of kThreads; per round a wave's 64 pieces are
//     padded to its first (longest) one.  Budgets: kSegs pieces, kEnt slots per
//     thread, kCols columns.
//   * kEven: the tile's non-zeros (row-major) dealt in runs of
//     E = ceil(nnz / kThreads); a segment starts at a row's first non-zero and
//     at each run's first.  Budgets: kThreads * kEnt non-zeros, kCols columns.
//   * slots are entry-major (slot e of thread t at ent_off + e * kThreads + t);
//     a key is local column | segment << 16; padding has value 0 and segment
//     nseg (the sink); inside a piece's padded round the padding keeps the
//     piece's segment.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

struct Geom {
    int threads, rows, cols, segs, waves, ent;
};
constexpr Geom kGeoms[2] = {{256, 512, 1024, 1024, 4, 16}, {512, 1024, 2048, 1536, 8, 16}};
constexpr int kPieceLen = 8;
constexpr int kRowAlign = 32;

struct Desc {  // same field order and sizes as the library's tile descriptor
    int64_t row0, ent_off, col_off;
    int32_t nrows, ncols, ent, nseg;
    uint32_t flush[8];
};
static_assert(sizeof(Desc) == 72, "descriptor size");

struct Plan {
    std::vector<Desc> desc;
    std::vector<uint32_t> key;
    std::vector<float> v32;
    std::vector<double> v64;
    std::vector<int32_t> cols;
    std::vector<uint16_t> rseg;
    std::vector<int32_t> huge;
};

int piece_count(int64_t n) { return (int)((n + kPieceLen - 1) / kPieceLen); }
int piece_len(int64_t n, int j) {
    const int p = piece_count(n);
    return (int)(n / p + (j < n % p ? 1 : 0));
}

// Slots per thread of a tile with piece-length histogram hist[1..8]: walk the
// sorted (descending) positions each wave takes its round leaders from.
int slots_of(const Geom &G, const int *hist) {
    int npieces = 0;
    for (int L = 1; L <= kPieceLen; ++L) npieces += hist[L];
    int worst = 0;
    for (int w = 0; w < G.waves; ++w) {
        int total = 0;
        for (int start = w * 64; start < npieces; start += G.threads) {
            int pos = start, L = kPieceLen;
            while (L > 0 && pos >= hist[L]) pos -= hist[L--];
            total += L;
        }
        worst = std::max(worst, total);
    }
    return worst;
}

Plan make_plan(int style, const Geom &G, int64_t nrows, int64_t ncols, const int64_t *rp, const int32_t *col,
               const double *val) {
    struct Span {
        int64_t row0;
        int32_t nrows, cap;
    };
    std::vector<Span> spans;
    std::vector<int32_t> stamp(ncols, -1);
    const int64_t max_nnz = (int64_t)G.threads * G.ent;
    int64_t r = 0;
    while (r < nrows) {
        const int32_t t = (int32_t)spans.size();
        Span cur{r, 0, 0};
        int hist[kPieceLen + 1] = {0};
        int npieces = 0;
        int64_t nnz = 0;
        while (r < nrows && cur.nrows < G.rows) {
            const int64_t k0 = rp[r], k1 = rp[r + 1], n = k1 - k0;
            if (n > G.cols) {
                ++cur.nrows;
                ++r;
                continue;
            }
            int h2[kPieceLen + 1];
            std::copy(hist, hist + kPieceLen + 1, h2);
            const int p = piece_count(n);
            for (int j = 0; j < p; ++j) ++h2[piece_len(n, j)];
            if (style == 0) {
                if (npieces + p > G.segs) break;
                if (slots_of(G, h2) > G.ent) break;
            } else {
                if (nnz + n > max_nnz) break;
            }
            int32_t nu = 0;
            for (int64_t k = k0; k < k1; ++k) nu += stamp[col[k]] != t;
            if (cur.cap + nu > G.cols) break;
            for (int64_t k = k0; k < k1; ++k) stamp[col[k]] = t;
            std::copy(h2, h2 + kPieceLen + 1, hist);
            npieces += p;
            nnz += n;
            cur.cap += nu;
            ++cur.nrows;
            ++r;
        }
        const int back = (int)(r % kRowAlign);
        if (r < nrows && back && cur.nrows > back) {
            r -= back;
            cur.nrows -= back;
        }
        spans.push_back(cur);
    }
    Plan P;
    const int64_t ntiles = (int64_t)spans.size();
    P.desc.assign(ntiles, Desc{});
    int64_t ent_total = 0, col_total = 0;
    for (int64_t t = 0; t < ntiles; ++t) {
        const Span &sp = spans[t];
        int hist[kPieceLen + 1] = {0};
        int64_t nnz = 0;
        for (int64_t rr = sp.row0; rr < sp.row0 + sp.nrows; ++rr) {
            const int64_t n = rp[rr + 1] - rp[rr];
            if (n > G.cols) continue;
            nnz += n;
            for (int j = 0; j < piece_count(n); ++j) ++hist[piece_len(n, j)];
        }
        Desc &d = P.desc[t];
        d.row0 = sp.row0;
        d.nrows = sp.nrows;
        d.ent = style == 0 ? slots_of(G, hist) : (int32_t)((nnz + G.threads - 1) / G.threads);
        d.ent_off = ent_total;
        d.col_off = col_total;
        ent_total += (int64_t)d.ent * G.threads;
        col_total += sp.cap;
    }
    P.key.assign(ent_total, 0u);
    P.v32.assign(ent_total, 0.f);
    P.v64.assign(ent_total, 0.0);
    P.cols.assign(col_total, 0);
    P.rseg.assign(nrows + ntiles, 0);
    for (int64_t rr = 0; rr < nrows; ++rr)
        if (rp[rr + 1] - rp[rr] > G.cols) P.huge.push_back((int32_t)rr);

    std::vector<int32_t> u;
    struct Piece {
        int64_t k0;
        int32_t len, sid;
    };
    std::vector<Piece> pieces;
    std::vector<int32_t> order;
    for (int64_t t = 0; t < ntiles; ++t) {
        Desc &d = P.desc[t];
        u.clear();
        for (int64_t rr = d.row0; rr < d.row0 + d.nrows; ++rr)
            if (rp[rr + 1] - rp[rr] <= G.cols)
                for (int64_t k = rp[rr]; k < rp[rr + 1]; ++k) u.push_back(col[k]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        d.ncols = (int32_t)u.size();
        std::copy(u.begin(), u.end(), P.cols.begin() + d.col_off);
        auto local = [&](int32_t c) { return (uint32_t)(std::lower_bound(u.begin(), u.end(), c) - u.begin()); };
        uint16_t *rs = P.rseg.data() + d.row0 + t;
        const int64_t slots = (int64_t)d.ent * G.threads;
        if (style == 0) {
            pieces.clear();
            for (int32_t lr = 0; lr < d.nrows; ++lr) {
                const int64_t rr = d.row0 + lr, n = rp[rr + 1] - rp[rr];
                rs[lr] = (uint16_t)pieces.size();
                if (n > G.cols) continue;
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
            order.resize(nseg);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(),
                             [&](int32_t a, int32_t b) { return pieces[a].len > pieces[b].len; });
            for (int64_t q = 0; q < slots; ++q) P.key[d.ent_off + q] = (uint32_t)nseg << 16;
            for (int w = 0; w < G.waves; ++w) {
                uint32_t fm = 0u;
                int off = 0;
                for (int start = w * 64; start < nseg; start += G.threads) {
                    const int L = pieces[order[start]].len;
                    for (int l = 0; l < 64 && start + l < nseg; ++l) {
                        const Piece &pc = pieces[order[start + l]];
                        for (int e = 0; e < L; ++e) {
                            const int64_t slot = d.ent_off + (int64_t)(off + e) * G.threads + w * 64 + l;
                            const bool in = e < pc.len;
                            P.key[slot] = (in ? local(col[pc.k0 + e]) : 0u) | ((uint32_t)pc.sid << 16);
                            P.v32[slot] = in ? (float)val[pc.k0 + e] : 0.f;
                            P.v64[slot] = in ? val[pc.k0 + e] : 0.0;
                        }
                    }
                    fm |= 1u << (off + L - 1);
                    off += L;
                }
                d.flush[w] = fm;
            }
        } else {
            const int64_t E = d.ent;
            int64_t q = 0;
            int32_t nseg = 0;
            for (int32_t lr = 0; lr < d.nrows; ++lr) {
                const int64_t rr = d.row0 + lr;
                rs[lr] = (uint16_t)nseg;
                if (rp[rr + 1] - rp[rr] > G.cols) continue;
                for (int64_t k = rp[rr]; k < rp[rr + 1]; ++k, ++q) {
                    const int64_t thr = q / E, e = q % E;
                    if (k == rp[rr] || e == 0) ++nseg;
                    const int64_t slot = d.ent_off + e * G.threads + thr;
                    P.key[slot] = local(col[k]) | ((uint32_t)(nseg - 1) << 16);
                    P.v32[slot] = (float)val[k];
                    P.v64[slot] = val[k];
                }
            }
            rs[d.nrows] = (uint16_t)nseg;
            d.nseg = nseg;
            for (; q < slots; ++q) {
                const int64_t slot = d.ent_off + (q % E) * G.threads + q / E;
                P.key[slot] = (uint32_t)nseg << 16;
            }
        }
    }
    return P;
}

}  // namespace

extern "C" {

// The plan of CSR (rp, col, val) in layout `style` (0: rounds, 1: even) and
// geometry `geom` (0: 256 threads / 512 rows / 1024 columns, 1: 512 / 1024 /
// 2048).  sizes: tiles, slots, column capacity, row-map entries, huge rows.
void *oracle_tile_plan_make(int32_t style, int32_t geom, int64_t nrows, int64_t ncols, const int64_t *rp,
                            const int32_t *col, const double *val, int64_t *sizes) {
    Plan *P = new Plan(make_plan(style, kGeoms[geom ? 1 : 0], nrows, ncols, rp, col, val));
    sizes[0] = (int64_t)P->desc.size();
    sizes[1] = (int64_t)P->key.size();
    sizes[2] = (int64_t)P->cols.size();
    sizes[3] = (int64_t)P->rseg.size();
    sizes[4] = (int64_t)P->huge.size();
    return P;
}

void oracle_tile_plan_copy(void *h, void *desc, uint32_t *key, float *v32, double *v64, int32_t *cols,
                           uint16_t *rseg, int32_t *huge) {
    const Plan *P = (const Plan *)h;
    std::memcpy(desc, P->desc.data(), P->desc.size() * sizeof(Desc));
    std::memcpy(key, P->key.data(), P->key.size() * sizeof(uint32_t));
    std::memcpy(v32, P->v32.data(), P->v32.size() * sizeof(float));
    std::memcpy(v64, P->v64.data(), P->v64.size() * sizeof(double));
    std::memcpy(cols, P->cols.data(), P->cols.size() * sizeof(int32_t));
    std::memcpy(rseg, P->rseg.data(), P->rseg.size() * sizeof(uint16_t));
    std::memcpy(huge, P->huge.data(), P->huge.size() * sizeof(int32_t));
}

void oracle_tile_plan_free(void *h) { delete (Plan *)h; }

}  // extern "C"
