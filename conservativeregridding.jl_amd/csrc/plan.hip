// plan.hip -- tile plans (tileplan.hpp) of a device CSR matrix, made on the
// device.  The staged apply needs one per (layout, geometry); it is made on the
// first apply that uses it:
//
//   1. k_plan_prev: for every non-zero k of a row that is not huge, prev[k] =
//      the last earlier non-huge row holding the same column (-1: none), from
//      a stable radix sort of the non-zeros by column.  A tile that starts at
//      row s counts, for each row, the non-zeros with prev[k] < s as new
//      distinct columns, which is the greedy column budget's count.
//   2. k_plan_table: for every start row s on the kRowAlign grid, the tile a
//      greedy pass from s makes (one workgroup per start: per-row counts, a
//      block scan, the first row that breaks a budget): end, column capacity,
//      slots per thread.
//   3. k_plan_walk: one workgroup follows the chain 0 -> end(0) -> ... through
//      that table (held in LDS in chunks).  A start off the grid (after a tile
//      shorter than kRowAlign rows) is evaluated in place.  Running sums give
//      every tile its first slot and first column.
//   4. one readback of the totals, the plan's buffers allocated;
//   5. k_plan_fill: one workgroup per tile: the distinct columns (block radix
//      sort of the tile's column indices), the pieces and their rounds
//      (kRounds) or the even runs (kEven), every slot, the row -> segment map;
//      it also sums the tiles' distinct columns and the 32-column lines they
//      start (the apply's choice between plain and block-ordered plans).
//
// The huge rows (more than kCols non-zeros) are listed by a flag scan.  The
// device-wide sort and scan come from build.hip (sort_pairs, exclusive_sum).
#include <rocprim/block/block_radix_sort.hpp>

#include "regridder.hpp"

namespace cr {

namespace {

constexpr int kPlanThreads = 256;  // workgroup of the table and walk kernels
constexpr int kWalkChunk = 8192;   // table entries the walk holds in LDS

struct TileRec {  // a tile found by the walk
    int64_t row0, ent_off, col_off;
    int32_t nrows, ent, cap, pad;
};

// Exclusive scan over the workgroup of K counters per thread (NT threads):
// v becomes the sum over earlier threads, tot the workgroup total.
// lds: K * NT / 64 words.  Starts and ends with every thread past a barrier.
template <int K, int NT>
__device__ __forceinline__ void block_scan(uint32_t (&v)[K], uint32_t (&tot)[K], uint32_t *lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc[K];
#pragma unroll
    for (int c = 0; c < K; ++c) inc[c] = v[c];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
        for (int c = 0; c < K; ++c) {
            const uint32_t t = __shfl_up(inc[c], d, 64);
            if (lane >= d) inc[c] += t;
        }
    }
    if (lane == 63) {
#pragma unroll
        for (int c = 0; c < K; ++c) lds[w * K + c] = inc[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < K; ++c) {
        uint32_t b = 0, s = 0;
        for (int i = 0; i < NT / 64; ++i) {
            const uint32_t x = lds[i * K + c];
            b += i < w ? x : 0u;
            s += x;
        }
        v[c] = inc[c] - v[c] + b;
        tot[c] = s;
    }
    __syncthreads();
}

// Piece-length histogram packed two lengths per word (16 bits each): length L
// sits in word (L-1)/2 at bit 16*((L-1)%2).  Exact while the pieces counted
// stay below 65536.
__device__ __forceinline__ void hist_add(uint32_t *h, int L, uint32_t cnt) {
#pragma unroll
    for (int j = 0; j < 4; ++j) h[j] += ((L - 1) >> 1) == j ? cnt << (16 * ((L - 1) & 1)) : 0u;
}
__device__ __forceinline__ void hist_cum(const uint32_t *h, int (&cum)[kPieceLen + 2]) {
    cum[kPieceLen + 1] = 0;
    cum[0] = 0;
#pragma unroll
    for (int L = kPieceLen; L >= 1; --L)
        cum[L] = cum[L + 1] + (int)((h[(L - 1) >> 1] >> (16 * ((L - 1) & 1))) & 0xffffu);
}

struct WindowTile {
    int64_t end;
    int cap, ent;
};

// The tile a greedy pass starting at row s makes (every thread of a
// kPlanThreads workgroup calls it).  Rows are added in order while the tile
// stays within kRows rows and its budgets; huge rows join with no entries.
// The tile then ends on a multiple of kRowAlign if that leaves it non-empty.
// cap: distinct columns counted when the greedy pass stopped (duplicates
// within a row counted each), ent: slots per thread of the final rows.
template <class G, int STYLE>
__device__ WindowTile window_tile(int64_t s, int64_t nrows, const int64_t *__restrict__ rp,
                                  const int32_t *__restrict__ prev, uint32_t *scan_lds, int *sh) {
    constexpr int NT = kPlanThreads, RPT = G::kRows / NT;
    constexpr int K = STYLE == kRounds ? 6 : 2;  // [0] pieces or non-zeros, [1] new columns, [2..5] histogram
    static_assert(G::kRows % NT == 0, "rows per thread");
    const int tid = threadIdx.x;
    if (tid == 0) sh[0] = G::kRows;
    uint32_t inc[RPT][K];
    bool live[RPT];
    uint32_t tot[K];
#pragma unroll
    for (int c = 0; c < K; ++c) tot[c] = 0u;
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int64_t r = s + tid * RPT + i;
        uint32_t v[K];
#pragma unroll
        for (int c = 0; c < K; ++c) v[c] = 0u;
        live[i] = false;
        if (r < nrows) {
            const int64_t k0 = rp[r], k1 = rp[r + 1], n = k1 - k0;
            if (n <= G::kCols) {
                live[i] = true;
                uint32_t nu = 0;
                for (int64_t k = k0; k < k1; ++k) nu += prev[k] < s ? 1u : 0u;
                v[1] = nu;
                if constexpr (STYLE == kRounds) {
                    if (n > 0) {
                        const int p = piece_count(n), q = (int)(n / p), x = (int)(n % p);
                        v[0] = (uint32_t)p;
                        hist_add(v + 2, q, (uint32_t)(p - x));
                        if (x) hist_add(v + 2, q + 1, (uint32_t)x);
                    }
                } else {
                    v[0] = (uint32_t)n;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < K; ++c) {
            tot[c] += v[c];
            inc[i][c] = tot[c];
        }
    }
    uint32_t wtot[K];
    block_scan<K, NT>(tot, wtot, scan_lds);  // tot: the rows of earlier threads
#pragma unroll
    for (int i = 0; i < RPT; ++i)
#pragma unroll
        for (int c = 0; c < K; ++c) inc[i][c] += tot[c];
    // the first row that breaks a budget (the budgets only grow along the rows)
    int fail = G::kRows;
#pragma unroll
    for (int i = RPT - 1; i >= 0; --i) {
        if (!live[i]) continue;
        bool f;
        if constexpr (STYLE == kRounds) {
            f = inc[i][0] > (uint32_t)G::kSegs || inc[i][1] > (uint32_t)G::kCols;
            // at most kEnt / kPieceLen rounds cannot break the slot budget
            if (!f && inc[i][0] > (uint32_t)(G::kEnt / kPieceLen * G::kThreads)) {
                int cum[kPieceLen + 2];
                hist_cum(inc[i] + 2, cum);
                f = tile_slots<G>(cum) > G::kEnt;
            }
        } else {
            f = inc[i][0] > (uint32_t)(G::kThreads * G::kEnt) || inc[i][1] > (uint32_t)G::kCols;
        }
        if (!f) break;  // earlier rows of this thread pass too
        fail = tid * RPT + i;
    }
    if (fail < G::kRows) atomicMin(&sh[0], fail);
    __syncthreads();
    const int64_t e = min(nrows, s + sh[0]);
    const int64_t down = e & ~(int64_t)(kRowAlign - 1);
    const int64_t end = (e < nrows && down > s) ? down : e;
    // capacity: new columns through row e-1; slots: histogram through row end-1
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int64_t r = s + tid * RPT + i;
        if (r == e - 1) sh[1] = (int)inc[i][1];
        if (r == end - 1) {
            if constexpr (STYLE == kRounds) {
                int cum[kPieceLen + 2];
                hist_cum(inc[i] + 2, cum);
                sh[2] = tile_slots<G>(cum);
            } else {
                sh[2] = (int)((inc[i][0] + G::kThreads - 1) / G::kThreads);
            }
        }
    }
    __syncthreads();
    WindowTile out{end, sh[1], sh[2]};
    __syncthreads();
    return out;
}

// Row of every non-zero, and the non-zero indices 0..nnz-1 (sort payload).
__global__ void k_plan_rows(int64_t nrows, const int64_t *__restrict__ rp, int32_t *__restrict__ row,
                            int32_t *__restrict__ iota) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
        row[k] = (int32_t)r;
        iota[k] = (int32_t)k;
    }
}

// prev[k] for the non-zeros of non-huge rows, from the non-zeros sorted by
// column (a stable sort of row-major order: each column's non-zeros in
// ascending row order, copies inside a row side by side): the nearest earlier
// entry of the same column whose row is earlier and not huge.  Looking back
// further than kRows rows is not needed: every tile that holds row r starts
// after r - kRows, so such a row counts as none.
__global__ void k_plan_prev(int64_t nnz, const int32_t *__restrict__ scol, const int32_t *__restrict__ sk,
                            const int32_t *__restrict__ row, const int64_t *__restrict__ rp, int64_t kcols,
                            int64_t krows, int32_t *__restrict__ prev) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int32_t c = scol[p], k = sk[p];
    const int64_t r = row[k];
    if (rp[r + 1] - rp[r] > kcols) return;
    int32_t found = -1;
    for (int64_t q = p - 1; q >= 0 && scol[q] == c; --q) {
        const int64_t r2 = row[sk[q]];
        if (r2 <= r - krows) break;
        if (r2 < r && rp[r2 + 1] - rp[r2] <= kcols) {
            found = (int32_t)r2;
            break;
        }
    }
    prev[k] = found;
}

template <class G, int STYLE>
__global__ void __launch_bounds__(kPlanThreads) k_plan_table(int64_t nrows, const int64_t *__restrict__ rp,
                                                             const int32_t *__restrict__ prev,
                                                             int2 *__restrict__ table) {
    __shared__ uint32_t scan_lds[6 * kPlanThreads / 64];
    __shared__ int sh[4];
    const int64_t s = (int64_t)blockIdx.x * kRowAlign;
    const WindowTile w = window_tile<G, STYLE>(s, nrows, rp, prev, scan_lds, sh);
    if (threadIdx.x == 0) table[blockIdx.x] = make_int2((int)w.end, w.cap | (w.ent << 16));
}

template <class G, int STYLE>
__global__ void __launch_bounds__(kPlanThreads) k_plan_walk(int64_t nrows, const int64_t *__restrict__ rp,
                                                            const int32_t *__restrict__ prev,
                                                            const int2 *__restrict__ table, TileRec *__restrict__ recs,
                                                            int64_t *__restrict__ totals) {
    __shared__ uint32_t scan_lds[6 * kPlanThreads / 64];
    __shared__ int sh[4];
    __shared__ int2 chunk[kWalkChunk];
    __shared__ int64_t bc[4];
    const int tid = threadIdx.x;
    const int64_t nA = (nrows + kRowAlign - 1) / kRowAlign;
    int64_t s = 0, t = 0, ent_total = 0, col_total = 0, cbase = -1;
    auto emit = [&](int64_t end, int cap, int ent) {
        TileRec rec;
        rec.row0 = s;
        rec.ent_off = ent_total;
        rec.col_off = col_total;
        rec.nrows = (int32_t)(end - s);
        rec.ent = ent;
        rec.cap = cap;
        rec.pad = 0;
        recs[t] = rec;
        ++t;
        ent_total += (int64_t)ent * G::kThreads;
        col_total += cap;
        s = end;
    };
    while (s < nrows) {
        if (s % kRowAlign == 0) {
            const int64_t a = s / kRowAlign;
            if (cbase < 0 || a < cbase || a >= cbase + kWalkChunk) {
                cbase = a;
                for (int i = tid; i < kWalkChunk && cbase + i < nA; i += kPlanThreads) chunk[i] = table[cbase + i];
                __syncthreads();
            }
            if (tid == 0) {
                while (s < nrows && s % kRowAlign == 0 && s / kRowAlign < cbase + kWalkChunk) {
                    const int2 e = chunk[s / kRowAlign - cbase];
                    emit(e.x, e.y & 0xffff, e.y >> 16);
                }
            }
        } else {
            const WindowTile w = window_tile<G, STYLE>(s, nrows, rp, prev, scan_lds, sh);
            if (tid == 0) emit(w.end, w.cap, w.ent);
        }
        if (tid == 0) {
            bc[0] = s;
            bc[1] = t;
            bc[2] = ent_total;
            bc[3] = col_total;
        }
        __syncthreads();
        s = bc[0];
        t = bc[1];
        ent_total = bc[2];
        col_total = bc[3];
        __syncthreads();
    }
    if (tid == 0) {
        totals[0] = t;
        totals[1] = ent_total;
        totals[2] = col_total;
    }
}

// Row holding tile-local non-zero q: the last row whose first non-zero is at
// or before q (rows with no non-zeros share the start of the next row and
// come before it).
__device__ __forceinline__ int row_of(const int32_t *qs, int nrows, int q) {
    int lo = 0, hi = nrows;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (qs[mid] <= q) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

template <class G, int STYLE>
__global__ void __launch_bounds__(G::kThreads) k_plan_fill(const TileRec *__restrict__ recs, int col_bits,
                                                           const int64_t *__restrict__ rp,
                                                           const int32_t *__restrict__ col,
                                                           const double *__restrict__ val, TileDesc *__restrict__ desc,
                                                           uint32_t *__restrict__ keys, float *__restrict__ v32,
                                                           double *__restrict__ v64, int32_t *__restrict__ cols,
                                                           uint16_t *__restrict__ rseg, int *__restrict__ err,
                                                           unsigned long long *__restrict__ stats) {
    constexpr int NT = G::kThreads, RPT = G::kRows / NT, IPT = G::kEnt, ZMAX = NT * IPT;
    using Sort = rocprim::block_radix_sort<uint32_t, NT, IPT, uint16_t>;
    // the sort's storage is free once every thread is past the barrier after it
    __shared__ union {
        typename Sort::storage_type sort;
        uint16_t lc[ZMAX];  // local column of each non-zero
    } u;
    uint16_t *lc = u.lc;
    __shared__ int32_t qs[G::kRows + 1];  // first tile-local non-zero of each row
    __shared__ int64_t rk[G::kRows];      // first non-zero of each row in A
    __shared__ uint32_t scan_lds[8 * NT / 64];
    __shared__ uint32_t lastkey[NT];
    const TileRec R = recs[blockIdx.x];
    const int tid = threadIdx.x;
    uint16_t *rs = rseg + R.row0 + blockIdx.x;

    // 1. rows: non-zeros (0 for huge rows) and pieces, scanned
    uint32_t rn[RPT], rpc[RPT];
    uint32_t b[2] = {0u, 0u}, tot[2];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int lr = tid * RPT + i;
        rn[i] = 0u;
        rpc[i] = 0u;
        if (lr < R.nrows) {
            const int64_t r = R.row0 + lr, k0 = rp[r], n = rp[r + 1] - k0;
            rk[lr] = k0;
            if (n <= G::kCols) {
                rn[i] = (uint32_t)n;
                rpc[i] = STYLE == kRounds ? (uint32_t)piece_count(n) : 0u;
            }
        }
        b[0] += rn[i];
        b[1] += rpc[i];
    }
    block_scan<2, NT>(b, tot, scan_lds);
    const int Z = (int)tot[0];
    // budgets the walk guarantees (a matrix whose transpose is not exact could break them)
    if (Z > ZMAX || (STYLE == kRounds && tot[1] > (uint32_t)G::kSegs) ||
        (STYLE == kEven && Z > R.ent * NT)) {
        if (tid == 0) err[0] = 1;
        return;
    }
    {
        uint32_t q = b[0], pc = b[1];
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int lr = tid * RPT + i;
            if (lr < R.nrows) {
                qs[lr] = (int32_t)q;
                if constexpr (STYLE == kRounds) rs[lr] = (uint16_t)pc;
            }
            q += rn[i];
            pc += rpc[i];
        }
        if (tid == 0) {
            qs[R.nrows] = Z;
            if constexpr (STYLE == kRounds) rs[R.nrows] = (uint16_t)tot[1];
        }
    }
    __syncthreads();

    // 2. distinct columns: sort (column, non-zero) pairs, rank the distinct ones
    uint32_t key[IPT];
    uint16_t qv[IPT];
    const uint32_t pad_key = (1u << col_bits) - 1u;
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int q = tid * IPT + i;
        qv[i] = (uint16_t)q;
        key[i] = pad_key;
        if (q < Z) {
            const int lr = row_of(qs, R.nrows, q);
            key[i] = (uint32_t)col[rk[lr] + (q - qs[lr])];
        }
    }
    Sort().sort(key, qv, u.sort, 0, col_bits);
    lastkey[tid] = key[IPT - 1];
    __syncthreads();
    uint32_t fl[1] = {0u}, ftot[1];
    bool first[IPT];
    {
        uint32_t pk = tid > 0 ? lastkey[tid - 1] : 0xffffffffu;
        uint32_t nlines = 0;  // 32-column lines (128 bytes of an FP32 field) the tile's columns start
#pragma unroll
        for (int i = 0; i < IPT; ++i) {
            const int pos = tid * IPT + i;
            first[i] = pos < Z && (pos == 0 || key[i] != pk);
            fl[0] += first[i] ? 1u : 0u;
            nlines += (pos < Z && (pos == 0 || (key[i] >> 5) != (pk >> 5))) ? 1u : 0u;
            pk = key[i];
        }
        if (nlines) atomicAdd(stats + 1, (unsigned long long)nlines);
    }
    block_scan<1, NT>(fl, ftot, scan_lds);
    {
        int rank = (int)fl[0] - 1;
#pragma unroll
        for (int i = 0; i < IPT; ++i) {
            const int pos = tid * IPT + i;
            if (pos < Z) {
                if (first[i]) {
                    ++rank;
                    if (rank < R.cap) cols[R.col_off + rank] = (int32_t)key[i];
                }
                lc[qv[i]] = (uint16_t)rank;
            }
        }
    }
    __syncthreads();
    const int ncols = (int)ftot[0];
    if (tid == 0) atomicAdd(stats, (unsigned long long)ncols);
    if (ncols > R.cap) {
        if (tid == 0) err[0] = 1;
        return;
    }

    if constexpr (STYLE == kRounds) {
        // 3r. pieces in row order, then sorted longest first (ties in row order)
        constexpr int PPT = (G::kSegs + NT - 1) / NT;
        __shared__ uint8_t plen[G::kSegs];
        __shared__ uint16_t pq0[G::kSegs];
        __shared__ uint16_t prow[G::kSegs];  // local row of each piece
        __shared__ uint16_t order[G::kSegs];
        const int nseg = (int)tot[1];
        {
            uint32_t q = b[0], pc = b[1];
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int lr = tid * RPT + i;
                if (rpc[i]) {
                    int off = 0;
                    for (int j = 0; j < (int)rpc[i]; ++j) {
                        const int len = piece_len(rn[i], j);
                        plen[pc + j] = (uint8_t)len;
                        pq0[pc + j] = (uint16_t)(q + off);
                        prow[pc + j] = (uint16_t)lr;
                        off += len;
                    }
                }
                q += rn[i];
                pc += rpc[i];
            }
        }
        __syncthreads();
        uint32_t cnt[kPieceLen], ctot[kPieceLen];
#pragma unroll
        for (int m = 0; m < kPieceLen; ++m) cnt[m] = 0u;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int sid = tid * PPT + j;
            if (sid < nseg) {
                const int L = plen[sid];
#pragma unroll
                for (int m = 0; m < kPieceLen; ++m) cnt[m] += L == m + 1 ? 1u : 0u;
            }
        }
        block_scan<kPieceLen, NT>(cnt, ctot, scan_lds);
        uint32_t start[kPieceLen];
        {
            uint32_t acc = 0u;
#pragma unroll
            for (int m = kPieceLen - 1; m >= 0; --m) {
                start[m] = acc + cnt[m];
                acc += ctot[m];
            }
        }
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int sid = tid * PPT + j;
            if (sid < nseg) {
                const int L = plen[sid];
                uint32_t pos = 0u;
#pragma unroll
                for (int m = 0; m < kPieceLen; ++m) {
                    if (L == m + 1) pos = start[m]++;
                }
                order[pos] = (uint16_t)sid;
            }
        }
        __syncthreads();
        // 4r. slots: round k gives sorted piece k*NT + tid to this thread; the
        // wave's slots for the round are its first (longest) piece's length
        const int w = tid >> 6;
        const uint32_t sink = (uint32_t)nseg << 16;
        {
            int need = 0;
            for (int st = w * 64; st < nseg; st += NT) need += plen[order[st]];
            if (need > R.ent) {  // the wave's rounds must fit the tile's slots
                err[0] = 1;
                return;
            }
        }
        uint32_t fm = 0u;
        int off = 0;
        for (int k = 0;; ++k) {
            const int st = k * NT + w * 64;
            if (st >= nseg) break;
            const int L = plen[order[st]];
            const int pos = k * NT + tid;
            if (pos < nseg) {
                const int sid = order[pos], len = plen[sid], q0 = pq0[sid];
                const int64_t kk = rk[prow[sid]] + (q0 - qs[prow[sid]]);
                for (int e = 0; e < L; ++e) {
                    const int64_t slot = R.ent_off + (int64_t)(off + e) * NT + tid;
                    uint32_t kv = (uint32_t)sid << 16;
                    double vv = 0.0;
                    if (e < len) {
                        kv |= lc[q0 + e];
                        vv = val[kk + e];
                    }
                    keys[slot] = kv;
                    v32[slot] = (float)vv;
                    v64[slot] = vv;
                }
            } else {
                for (int e = 0; e < L; ++e) {
                    const int64_t slot = R.ent_off + (int64_t)(off + e) * NT + tid;
                    keys[slot] = sink;
                    v32[slot] = 0.f;
                    v64[slot] = 0.0;
                }
            }
            fm |= 1u << (off + L - 1);
            off += L;
        }
        for (int e = off; e < R.ent; ++e) {
            const int64_t slot = R.ent_off + (int64_t)e * NT + tid;
            keys[slot] = sink;
            v32[slot] = 0.f;
            v64[slot] = 0.0;
        }
        TileDesc *d = desc + blockIdx.x;
        if ((tid & 63) == 0) d->flush[w] = fm;
        if (tid == 0) {
            for (int x = G::kWaves; x < kMaxTileWaves; ++x) d->flush[x] = 0u;
            d->nseg = nseg;
        }
    } else {
        // 3e. even runs: thread t holds non-zeros [t*E, t*E + E); a segment
        // starts at a row's first non-zero and at each run's first
        __shared__ uint8_t sflag[ZMAX];
        __shared__ uint32_t sbase[NT];
        const int E = R.ent;
        uint32_t c[1] = {0u}, ctot[1];
        int lrow[IPT];
#pragma unroll
        for (int e = 0; e < IPT; ++e) {
            const int q = tid * E + e;
            lrow[e] = -1;
            if (e < E && q < Z) {
                const int lr = row_of(qs, R.nrows, q);
                const bool st = e == 0 || q == qs[lr];
                lrow[e] = lr;
                sflag[q] = st ? 1 : 0;
                c[0] += st ? 1u : 0u;
            }
        }
        block_scan<1, NT>(c, ctot, scan_lds);
        const int nseg = (int)ctot[0];
        sbase[tid] = c[0];
        __syncthreads();
        const uint32_t sink = (uint32_t)nseg << 16;
        uint32_t sg = c[0];
#pragma unroll
        for (int e = 0; e < IPT; ++e) {
            if (e < E) {
                const int q = tid * E + e;
                const int64_t slot = R.ent_off + (int64_t)e * NT + tid;
                if (q < Z) {
                    const int lr = lrow[e];
                    sg += sflag[q];
                    keys[slot] = (uint32_t)lc[q] | ((sg - 1u) << 16);
                    const double vv = val[rk[lr] + (q - qs[lr])];
                    v32[slot] = (float)vv;
                    v64[slot] = vv;
                } else {
                    keys[slot] = sink;
                    v32[slot] = 0.f;
                    v64[slot] = 0.0;
                }
            }
        }
        // first segment of each row: the segments started before its first non-zero
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int lr = tid * RPT + i;
            if (lr < R.nrows) {
                const int q = qs[lr];
                int S = nseg;
                if (q < Z) {
                    const int t = q / E;
                    S = (int)sbase[t];
                    for (int x = t * E; x < q; ++x) S += sflag[x];
                }
                rs[lr] = (uint16_t)S;
            }
        }
        if (tid == 0) rs[R.nrows] = (uint16_t)nseg;
        TileDesc *d = desc + blockIdx.x;
        if (tid < kMaxTileWaves) d->flush[tid] = 0u;
        if (tid == 0) d->nseg = nseg;
    }
    if (tid == 0) {
        TileDesc *d = desc + blockIdx.x;
        d->row0 = R.row0;
        d->ent_off = R.ent_off;
        d->col_off = R.col_off;
        d->nrows = R.nrows;
        d->ncols = ncols;
        d->ent = R.ent;
    }
}

// Rows over kcols non-zeros: flag[r] (0/1), then their ascending list.
__global__ void k_huge_flags(int64_t nrows, const int64_t *__restrict__ rp, int64_t kcols,
                             int64_t *__restrict__ flag) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows) flag[r] = rp[r + 1] - rp[r] > kcols ? 1 : 0;
}
__global__ void k_huge_list(int64_t nrows, const int64_t *__restrict__ rp, int64_t kcols,
                            const int64_t *__restrict__ pos, int32_t *__restrict__ huge) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows && rp[r + 1] - rp[r] > kcols) huge[pos[r]] = (int32_t)r;
}

}  // namespace

template <class G, int STYLE>
void make_plan(const Csr &A, TilePlan &P, hipStream_t s) {
    const int64_t nrows = A.nrows, nnz = A.nnz;
    const int64_t nA = (nrows + kRowAlign - 1) / kRowAlign;
    CR_REQUIRE(nrows < ((int64_t)1 << 31) - kRowAlign && A.ncols < ((int64_t)1 << 31), CR_UNSUPPORTED,
               "tile plans take fewer than 2^31 rows and columns");
    CR_REQUIRE(nnz < ((int64_t)1 << 31), CR_UNSUPPORTED, "tile plans take fewer than 2^31 non-zeros");
    if (nrows == 0) return;
    Temp T(s);
    int col_bits = 1;
    while (((int64_t)1 << col_bits) <= A.ncols) ++col_bits;  // (1 << col_bits) - 1 > every column
    DevBuf<int32_t> prev(nnz > 0 ? nnz : 1, s);
    if (nnz > 0) {
        DevBuf<int32_t> row(nnz, s), iota(nnz, s), scol(nnz, s), sk(nnz, s);
        hipLaunchKernelGGL(k_plan_rows, dim3(grid_for(nrows, 256)), dim3(256), 0, s, nrows, A.rowptr.p, row.p,
                           iota.p);
        sort_pairs<int32_t, int32_t>(T, A.col.p, scol.p, iota.p, sk.p, nnz, col_bits, s);
        hipLaunchKernelGGL(k_plan_prev, dim3(grid_for(nnz, 256)), dim3(256), 0, s, nnz, scol.p, sk.p, row.p,
                           A.rowptr.p, (int64_t)G::kCols, (int64_t)G::kRows, prev.p);
    }
    DevBuf<int2> table(nA > 0 ? nA : 1, s);
    DevBuf<TileRec> recs(nrows > 0 ? nrows : 1, s);
    DevBuf<int64_t> totals(3, s);
    DevBuf<int64_t> hflag(nrows + 1, s), hpos(nrows + 1, s);
    hipLaunchKernelGGL((k_plan_table<G, STYLE>), dim3((unsigned)nA), dim3(kPlanThreads), 0, s, nrows, A.rowptr.p,
                       prev.p, table.p);
    hipLaunchKernelGGL((k_plan_walk<G, STYLE>), dim3(1), dim3(kPlanThreads), 0, s, nrows, A.rowptr.p, prev.p,
                       table.p, recs.p, totals.p);
    CR_HIP(hipMemsetAsync(hflag.p + nrows, 0, sizeof(int64_t), s));
    hipLaunchKernelGGL(k_huge_flags, dim3(grid_for(nrows, 256)), dim3(256), 0, s, nrows, A.rowptr.p,
                       (int64_t)G::kCols, hflag.p);
    exclusive_sum<int64_t *, int64_t *>(T, hflag.p, hpos.p, nrows + 1, s);
    CR_HIP(hipGetLastError());
    int64_t h[3], nh = 0;
    CR_HIP(hipMemcpyAsync(h, totals.p, sizeof(h), hipMemcpyDeviceToHost, s));
    CR_HIP(hipMemcpyAsync(&nh, hpos.p + nrows, sizeof(nh), hipMemcpyDeviceToHost, s));
    CR_HIP(hipStreamSynchronize(s));
    P.ntiles = h[0];
    P.nslots = h[1];
    P.nhuge = nh;
    P.tiles.alloc_kept(P.ntiles, s);
    P.key.alloc_kept(P.nslots, s);
    P.val32.alloc_kept(P.nslots, s);
    P.val64.alloc_kept(P.nslots, s);
    P.cols.alloc_kept(h[2], s);
    P.rseg.alloc_kept(nrows + P.ntiles, s);
    P.huge.alloc_kept(P.nhuge, s);
    if (P.nhuge)
        hipLaunchKernelGGL(k_huge_list, dim3(grid_for(nrows, 256)), dim3(256), 0, s, nrows, A.rowptr.p,
                           (int64_t)G::kCols, hpos.p, P.huge.p);
    // err, then the tiles' distinct columns and the 32-column lines they start
    DevBuf<unsigned long long> fstat(3, s);
    CR_HIP(hipMemsetAsync(fstat.p, 0, fstat.bytes(), s));
    if (P.ntiles)
        hipLaunchKernelGGL((k_plan_fill<G, STYLE>), dim3((unsigned)P.ntiles), dim3(G::kThreads), 0, s, recs.p,
                           col_bits, A.rowptr.p, A.col.p, A.val.p, P.tiles.p, P.key.p, P.val32.p, P.val64.p,
                           P.cols.p, P.rseg.p, (int *)fstat.p, fstat.p + 1);
    CR_HIP(hipGetLastError());
    unsigned long long f[3] = {0, 0, 0};
    CR_HIP(hipMemcpyAsync(f, fstat.p, sizeof(f), hipMemcpyDeviceToHost, s));
    CR_HIP(hipStreamSynchronize(s));
    CR_REQUIRE(!f[0], CR_INTERNAL, "tile plan: a tile broke its budgets");
    P.tile_cols = (int64_t)f[1];
    P.tile_lines = (int64_t)f[2];
}

template void make_plan<GeomS, kRounds>(const Csr &, TilePlan &, hipStream_t);
template void make_plan<GeomS, kEven>(const Csr &, TilePlan &, hipStream_t);
template void make_plan<GeomL, kRounds>(const Csr &, TilePlan &, hipStream_t);
template void make_plan<GeomL, kEven>(const Csr &, TilePlan &, hipStream_t);

template <class G, int STYLE>
void make_block_plan(const Csr &A, TilePlan &P, hipStream_t s);  // planblocks.hip

const TilePlan &tile_plan(const Csr &A, int style, int geom, hipStream_t s, bool blocks) {
    std::lock_guard<std::mutex> lock(A.plan_mu);
    auto &slot = A.tiles[style][geom][blocks ? 1 : 0];
    if (slot) return *slot;
    CR_REQUIRE_NOT_CAPTURING(s, "the tile plan of this direction");
    auto P = std::make_shared<TilePlan>();
    P->geom = geom;
    if (geom == kGeomL) {
        if (style == kEven) blocks ? make_block_plan<GeomL, kEven>(A, *P, s) : make_plan<GeomL, kEven>(A, *P, s);
        else blocks ? make_block_plan<GeomL, kRounds>(A, *P, s) : make_plan<GeomL, kRounds>(A, *P, s);
    } else {
        if (style == kEven) blocks ? make_block_plan<GeomS, kEven>(A, *P, s) : make_plan<GeomS, kEven>(A, *P, s);
        else blocks ? make_block_plan<GeomS, kRounds>(A, *P, s) : make_plan<GeomS, kRounds>(A, *P, s);
    }
    slot = P;
    return *slot;
}

bool plan_ready(const Csr &A, int style, int geom, bool blocks) {
    std::lock_guard<std::mutex> lock(A.plan_mu);
    return (bool)A.tiles[style][geom][blocks ? 1 : 0];
}

void drop_tile_plan(const Csr &A) {
    std::lock_guard<std::mutex> lock(A.plan_mu);
    for (auto &byStyle : A.tiles)
        for (auto &byGeom : byStyle)
            for (auto &p : byGeom) p.reset();
}

}  // namespace cr

using namespace cr;

extern "C" {

cr_status cr_regrid_tile_plan(const cr_regridder *R, int32_t transpose, int32_t style, int32_t geom,
                              int64_t *sizes, void *stream) {
    return guard([&] {
        CR_REQUIRE(R && sizes, CR_INVALID_ARGUMENT, "cr_regrid_tile_plan: null argument");
        CR_REQUIRE(style >= 0 && style <= 3 && (geom == kGeomS || geom == kGeomL), CR_INVALID_ARGUMENT,
                   "cr_regrid_tile_plan: style is 0..3 and geom 0 or 1");
        const Csr &A = transpose ? R->bwd : R->fwd;
        hipStream_t s = as_stream(stream);
        for (int i = 0; i < 6; ++i) sizes[i] = 0;
        if (A.nrows == 0) return;
        const TilePlan &P = tile_plan(A, style & 1, geom, s, (style & 2) != 0);
        CR_HIP(hipStreamSynchronize(s));
        sizes[0] = P.ntiles;
        sizes[1] = P.nslots;
        sizes[2] = (int64_t)P.cols.n;
        sizes[3] = (int64_t)P.rseg.n;
        sizes[4] = P.nhuge;
        sizes[5] = (int64_t)P.rowmap.n;
    });
}

cr_status cr_regrid_copy_tile_plan(const cr_regridder *R, int32_t transpose, int32_t style, int32_t geom,
                                   void *desc, uint32_t *key, float *v32, double *v64, int32_t *cols,
                                   uint16_t *rseg, int32_t *huge, int32_t *rowmap) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "cr_regrid_copy_tile_plan: null regridder");
        CR_REQUIRE(style >= 0 && style <= 3 && (geom == kGeomS || geom == kGeomL), CR_INVALID_ARGUMENT,
                   "cr_regrid_copy_tile_plan: style is 0..3 and geom 0 or 1");
        const Csr &A = transpose ? R->bwd : R->fwd;
        std::shared_ptr<TilePlan> P;
        {
            std::lock_guard<std::mutex> lock(A.plan_mu);
            P = A.tiles[style & 1][geom][(style & 2) ? 1 : 0];
        }
        CR_REQUIRE(P, CR_INVALID_ARGUMENT, "cr_regrid_copy_tile_plan: the plan is not made (cr_regrid_tile_plan)");
        CR_HIP(hipDeviceSynchronize());
        auto down = [](void *dst, const void *src, size_t bytes) {
            if (bytes) {
                CR_REQUIRE(dst, CR_INVALID_ARGUMENT, "cr_regrid_copy_tile_plan: null buffer");
                CR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
            }
        };
        down(desc, P->tiles.p, P->tiles.bytes());
        down(key, P->key.p, P->key.bytes());
        down(v32, P->val32.p, P->val32.bytes());
        down(v64, P->val64.p, P->val64.bytes());
        down(cols, P->cols.p, P->cols.bytes());
        down(rseg, P->rseg.p, P->rseg.bytes());
        down(huge, P->huge.p, P->huge.bytes());
        down(rowmap, P->rowmap.p, P->rowmap.bytes());
    });
}

}  // extern "C"
