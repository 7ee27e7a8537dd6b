// tileplan.hpp -- row tiles of a CSR matrix for the staged apply: geometry and
// layout (plan.hip makes the plans on the device, tiles.hip runs them).
//
// A tile is a contiguous range of rows with at most kRows rows and kCols
// distinct columns (TileGeom).  Tiles are cut greedily in row order: a tile
// takes rows while its budgets hold, then ends on a multiple of kRowAlign rows
// when it can give rows back (so no 128-byte line of a FP32 output field is
// shared by two tiles).  Two layouts of its non-zeros over the workgroup's
// threads:
//
// kRounds: rows are cut into pieces of at most kPieceLen non-zeros (a row of
// n non-zeros gives ceil(n / kPieceLen) pieces of near-equal length); a piece
// is a *segment*: the kernel sums it in one accumulator and writes the sum to
// LDS slot `segment id`.  Pieces are sorted by length (longest first, ties in
// row order) and dealt out in rounds of kThreads: round k gives piece
// k*kThreads + t to thread t.  In every round, the 64 pieces of one wave are
// padded to the longest of them, so all lanes of a wave end their segments on
// the same slot: the end-of-segment test is a wave-uniform bit mask
// (TileDesc::flush), not a per-lane branch.  Budgets: kSegs pieces, kEnt slots
// per thread, kCols columns.
//
// kEven: the non-zeros (row-major) are dealt out in equal runs of
// E = ceil(nnz / kThreads) to the threads, cutting rows anywhere; a run of one
// row inside one thread is a segment, and each lane ends its segments where
// its own run changes row (a per-lane test in the kernel: TileDesc::flush is
// zero).  No padding beyond the last run: the cheaper layout when few fields
// share the per-tile set-up.  Budgets: kThreads * kEnt non-zeros, kCols columns.
//
// Either way a thread's slots (at most kEnt) stay in registers for the whole
// launch.  Rows with more than kCols non-zeros ("huge") stay in their tile
// with no entries and are listed for another kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

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
// Plan layouts: kRounds (wave-uniform segment ends, for many fields) and kEven
// (no padding, for one or two fields).
enum TileStyle { kRounds = 0, kEven = 1 };
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
static_assert(sizeof(TileDesc) == 72, "TileDesc is 72 bytes");

// Number of pieces a row of n non-zeros is cut into, and the length of piece j.
__host__ __device__ __forceinline__ int piece_count(int64_t n) { return (int)((n + kPieceLen - 1) / kPieceLen); }
__host__ __device__ __forceinline__ int piece_len(int64_t n, int j) {
    const int p = piece_count(n);
    return (int)(n / p + (j < n % p ? 1 : 0));
}

// Slots per thread of a tile whose pieces, sorted longest first, are dealt in
// rounds of kThreads; per wave and round the slot count is the longest piece
// of that wave's 64.  cum[L] = pieces of length >= L (cum[1]: all pieces).
// Returns the largest per-wave total.
template <class G>
__host__ __device__ __forceinline__ int tile_slots(const int (&cum)[kPieceLen + 2]) {
    int worst = 0;
    for (int w = 0; w < G::kWaves; ++w) {
        int total = 0;
        for (int start = w * 64; start < cum[1]; start += G::kThreads) {
            // length of the piece at sorted position `start`
#pragma unroll
            for (int L = 1; L <= kPieceLen; ++L) total += cum[L] > start ? 1 : 0;
        }
        worst = std::max(worst, total);
    }
    return worst;
}

}  // namespace cr
