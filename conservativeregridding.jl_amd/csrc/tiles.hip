// tiles.hip -- staged batched apply: y[:, f] = A x[:, f] for many fields f.
//
// Reference: the batched regrid! is a serial loop of 1-D mul! calls over the
// non-spatial indices (NDSliceLoop, src/regridder/regrid.jl:303-318).  Here one
// launch does every field, organised so that each byte of the fields crosses
// HBM about once:
//
//   * rows are cut into tiles (contiguous rows, at most kTileRows rows,
//     kTileCols distinct columns and kTileThreads * kTileEnt non-zeros);
//   * a tile's rows are cut into pieces of at most 8 non-zeros, sorted by
//     length and dealt out to the workgroup's threads in rounds, each wave's
//     64 pieces of a round padded to the longest (tileplan.hpp): long and
//     short rows cost the same per thread, and every lane of a wave ends its
//     segments on the same slot (a wave-uniform mask, no divergent branch);
//     the plans are made on the device (plan.hip);
//   * each thread keeps its slots (packed local column + segment id, and
//     the value) in registers for all fields of the launch;
//   * per field, the tile's distinct columns of x are loaded once (ascending
//     column order: runs of neighbouring cells, coalesced) into LDS, every
//     thread sums its row segments from LDS, and a second pass adds the
//     segments of each row in order and stores the rows coalesced.
//
// Launch: one workgroup per (tile, field chunk).  Workgroups go round-robin to
// the 8 XCDs; workgroup b works on tile (b % 8) * tiles_per_xcd + (b / 8) /
// nchunks, so every XCD owns a contiguous range of tiles (shared source cells
// stay in its L2) and the chunks of one tile run side by side (the tile's
// entries are read from HBM once).
//
// Numerics (MODE of the kernel, chosen per variant in apply.hip): FP32 fields
// either take products and segment sums in FP64 from the FP64 A (MODE 1, the
// default) or from A rounded to FP32 (MODE 2), or FP32 products of A rounded to
// FP32 with Kahan-compensated FP32 segment sums (MODE 0); in every mode the
// segments of a row are added in FP64, scaled by the FP64 reciprocal area and
// rounded to FP32 once.  FP64 fields use fused multiply-adds and a true
// division by the area.  Results are bitwise reproducible for a given variant
// (a fixed summation order, no atomics); another variant can differ in the
// last bits.
#include <algorithm>
#include <type_traits>

#include "regridder.hpp"
#include "tileplan.hpp"

namespace cr {


template <class T>
struct SegAcc;

template <>
struct SegAcc<double> {
    double s;
    __device__ void init() { s = 0.0; }
    __device__ void add(double v, double x) { s = fma(v, x, s); }
    __device__ double value() const { return s; }
};

template <>
struct SegAcc<float> {
    float s, c;  // Kahan: the compensated sum is s - c
    __device__ void init() {
        s = 0.f;
        c = 0.f;
    }
    __device__ void add(float v, float x) {
        float yk = v * x - c;
        float t = s + yk;
        c = (t - s) - yk;
        s = t;
    }
    __device__ float value() const { return (float)((double)s - (double)c); }
};

// FP32 fields summed in FP64: the FP64 A value times the (exactly widened)
// FP32 field value, fused multiply-add in FP64.  On gfx950 an FP64 FMA issues
// at the rate of an unpacked FP32 FMA, so this costs 2 VALU operations per
// product (convert + FMA) against 4 for a Kahan step, with smaller error.
struct SegAcc64 {
    double s;
    __device__ void init() { s = 0.0; }
    __device__ void add(double v, float x) { s = fma(v, (double)x, s); }
    __device__ float value() const { return (float)s; }
};

// scale = 1/area (FP32 fields) or area (FP64 fields), nullptr: no normalisation
__device__ __forceinline__ float row_finish(double acc, double scale, bool has, float *) {
    return (float)(has ? acc * scale : acc);
}
__device__ __forceinline__ double row_finish(double acc, double scale, bool has, double *) {
    return has ? acc / scale : acc;
}

// F field values of one cell (or one segment) side by side, so one LDS
// instruction moves all F of them.
template <class T, int F>
struct alignas(sizeof(T) * F >= 16 ? 16 : sizeof(T) * F) Pack {
    T v[F];
};

// MODE 0: T accumulation (FP64 fields: FMA; FP32 fields: Kahan);
// MODE 1: FP32 fields accumulated in FP64 (values from the FP64 A);
// MODE 2: as MODE 1 with the values rounded to FP32.
// OFF32: every cell offset (cell * stride) fits 32 bits, so loads and stores
// use a per-field base address plus a 32-bit offset.
// val: the register type of A's values, load: the type they are read as.
template <int MODE, class T>
struct TileAcc {
    using type = SegAcc<T>;
    using val = T;
    using load = T;
};
template <>
struct TileAcc<1, float> {
    using type = SegAcc64;
    using val = double;
    using load = double;
};
// MODE 2: FP32 fields accumulated in FP64 from A rounded to FP32 (8 bytes per
// slot instead of 12; A's rounding, 2^-24 relative, is below the FP32 result's).
// The values stay FP32 in registers and are widened at each product (widening
// them once per launch measured slower: twice the registers per slot).
template <>
struct TileAcc<2, float> {
    using type = SegAcc64;
    using val = float;
    using load = float;
};

// UNI: segment ends are the plan's wave-uniform masks (kRounds plans); else
// each lane finds them from its own slots (kEven plans).
template <class G, class T, int F, int MODE, bool OFF32, bool UNI>
__global__ void __launch_bounds__(G::kThreads) k_spmm_tiles(
    const TileDesc *__restrict__ tiles, int64_t ntiles, int64_t tiles_per_xcd, int nchunks,
    int64_t chunk, int chunk_major, const uint32_t *__restrict__ keys,
    const typename TileAcc<MODE, T>::load *__restrict__ vals,
    const int32_t *__restrict__ cols, const uint16_t *__restrict__ rseg, const T *__restrict__ x,
    int64_t sc, int64_t sf, T *__restrict__ y, int64_t dc, int64_t df, int64_t nf,
    const double *__restrict__ scale, const int32_t *__restrict__ rowmap) {
    using Acc = typename TileAcc<MODE, T>::type;
    using V = typename TileAcc<MODE, T>::val;
    using Off = typename std::conditional<OFF32, uint32_t, int64_t>::type;
    __shared__ Pack<T, F> xs[G::kCols];
    __shared__ Pack<T, F> seg[G::kSegs + 1];
    const int64_t b = blockIdx.x;
    const int64_t local = b >> 3;
    // tile-major: the chunks of a tile run side by side; chunk-major: the whole
    // GPU sweeps one field chunk at a time
    const int64_t tl = chunk_major ? local % tiles_per_xcd : local / nchunks;
    const int64_t ch = chunk_major ? local / tiles_per_xcd : local % nchunks;
    const int64_t tile = (b & 7) * tiles_per_xcd + tl;
    if (tile >= ntiles) return;
    const int64_t fb = ch * chunk;
    const int64_t fe = min(nf, fb + chunk);
    if (fb >= fe) return;
    const TileDesc d = tiles[tile];
    const int tid = threadIdx.x;
    const int E = d.ent;

    uint32_t key[G::kEnt];
    V v[G::kEnt];
#pragma unroll
    for (int e = 0; e < G::kEnt; ++e) {
        key[e] = 0u;
        v[e] = V(0);
        if (e < E) {
            const int64_t q = d.ent_off + (int64_t)e * G::kThreads + tid;
            key[e] = keys[q];
            v[e] = (V)vals[q];
        }
    }
    // bit e: slot e ends a segment (for every lane of the wave when UNI)
    uint32_t flush = 0u;
    if constexpr (UNI) {
        // read from the descriptor in memory: indexing the local copy by wave
        // would put the copy in scratch
        flush = __builtin_amdgcn_readfirstlane(tiles[tile].flush[tid >> 6]);
    } else {
#pragma unroll
        for (int e = 0; e < G::kEnt; ++e) {
            const uint32_t nxt = (e + 1 < G::kEnt) ? (key[e + 1] >> 16) : 0xffffffffu;
            if (e < E && (e + 1 == E || nxt != (key[e] >> 16))) flush |= 1u << e;
        }
    }
    Off xo[G::kColsPerThread];
    bool xin[G::kColsPerThread];
#pragma unroll
    for (int j = 0; j < G::kColsPerThread; ++j) {
        const int u = tid + j * G::kThreads;
        xin[j] = u < d.ncols;
        xo[j] = xin[j] ? (Off)((int64_t)cols[d.col_off + u] * sc) : (Off)0;
    }
    const uint16_t *rs = rseg + d.row0 + tile;
    int sb[G::kRowsPerThread], se[G::kRowsPerThread];
    bool rin[G::kRowsPerThread];
    Off yo[G::kRowsPerThread];
    double rscale[G::kRowsPerThread];
#pragma unroll
    for (int i = 0; i < G::kRowsPerThread; ++i) {
        const int r = tid + i * G::kThreads;
        // the output row: the plan row itself, or its matrix row in a block-ordered plan
        int64_t orow = d.row0 + r;
        if (rowmap && r < d.nrows) orow = rowmap[orow];
        rin[i] = r < d.nrows && orow >= 0;
        sb[i] = 0;
        se[i] = 0;
        rscale[i] = 1.0;
        yo[i] = rin[i] ? (Off)(orow * dc) : (Off)0;
        if (rin[i]) {
            sb[i] = rs[r];
            se[i] = rs[r + 1];
            if (scale) rscale[i] = scale[orow];
        }
    }

    T pre[F][G::kColsPerThread];
#pragma unroll
    for (int f = 0; f < F; ++f) {
        const T *xb = x + (fb + f) * sf;  // uniform per field
        const bool fok = fb + f < fe;
#pragma unroll
        for (int j = 0; j < G::kColsPerThread; ++j) pre[f][j] = (xin[j] && fok) ? xb[xo[j]] : T(0);
    }

    for (int64_t f0 = fb; f0 < fe; f0 += F) {
#pragma unroll
        for (int j = 0; j < G::kColsPerThread; ++j) {
            Pack<T, F> p;
#pragma unroll
            for (int f = 0; f < F; ++f) p.v[f] = pre[f][j];
            xs[tid + j * G::kThreads] = p;
        }
        const int64_t fn = f0 + F;
        if (fn < fe) {
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const T *xb = x + (fn + f) * sf;
                const bool fok = fn + f < fe;
#pragma unroll
                for (int j = 0; j < G::kColsPerThread; ++j) pre[f][j] = (xin[j] && fok) ? xb[xo[j]] : T(0);
            }
        }
        __syncthreads();
        // segment sums for F fields at once (padding slots: value 0, sink segment)
        {
            Acc a[F];
#pragma unroll
            for (int f = 0; f < F; ++f) a[f].init();
#pragma unroll
            for (int e = 0; e < G::kEnt; ++e) {
                if (e < E) {
                    const Pack<T, F> xv = xs[key[e] & 0xffffu];
#pragma unroll
                    for (int f = 0; f < F; ++f) a[f].add(v[e], xv.v[f]);
                    if (flush & (1u << e)) {
                        Pack<T, F> o;
#pragma unroll
                        for (int f = 0; f < F; ++f) {
                            o.v[f] = a[f].value();
                            a[f].init();
                        }
                        seg[key[e] >> 16] = o;
                    }
                }
            }
        }
        __syncthreads();
        // rows: add their segments in order, store coalesced
#pragma unroll
        for (int i = 0; i < G::kRowsPerThread; ++i) {
            if (rin[i]) {
                double acc[F];
#pragma unroll
                for (int f = 0; f < F; ++f) acc[f] = 0.0;
                for (int q = sb[i]; q < se[i]; ++q) {
                    const Pack<T, F> sp = seg[q];
#pragma unroll
                    for (int f = 0; f < F; ++f) acc[f] += (double)sp.v[f];
                }
#pragma unroll
                for (int f = 0; f < F; ++f) {
                    if (f0 + f < fe) {
                        T *yb = y + (f0 + f) * df;  // uniform per field
                        // written once: keep it from displacing source lines in L2
                        __builtin_nontemporal_store(row_finish(acc[f], rscale[i], scale != nullptr, (T *)nullptr),
                                                    yb + yo[i]);
                    }
                }
            }
        }
        // the next iteration's LDS writes of xs follow the barrier above (all
        // segment sums are done); its segment writes follow its first barrier
        // (all row passes of this iteration are done)
    }
}

// ---- launch --------------------------------------------------------------------------
template <class G, class T, int F, int MODE, bool UNI>
static void launch_f(const TilePlan &P, int nchunks, int chunk_major, const T *x, int64_t sc, int64_t sf,
                     T *y, int64_t dc, int64_t df, int64_t nf, const double *scale, int64_t nrows,
                     int64_t ncols, hipStream_t s) {
    const int64_t tiles_per_xcd = (P.ntiles + 7) / 8;
    int64_t chunk = (nf + nchunks - 1) / nchunks;
    chunk = (chunk + F - 1) / F * F;
    nchunks = (int)((nf + chunk - 1) / chunk);
    const int64_t blocks = tiles_per_xcd * 8 * nchunks;
    CR_REQUIRE(blocks < (int64_t)1 << 31, CR_UNSUPPORTED, "too many tiles x field chunks for one launch");
    using V = typename TileAcc<MODE, T>::load;
    const V *vals;
    if constexpr (sizeof(V) == 4) vals = (const V *)P.val32.p;
    else vals = (const V *)P.val64.p;
    const bool off32 = (ncols - 1) * sc < ((int64_t)1 << 32) && (nrows - 1) * dc < ((int64_t)1 << 32);
    if (off32)
        hipLaunchKernelGGL((k_spmm_tiles<G, T, F, MODE, true, UNI>), dim3((unsigned)blocks), dim3(G::kThreads), 0, s,
                           P.tiles.p, P.ntiles, tiles_per_xcd, nchunks, chunk, chunk_major, P.key.p, vals,
                           P.cols.p, P.rseg.p,
                           x, sc, sf, y, dc, df, nf, scale, P.rowmap.p);
    else
        hipLaunchKernelGGL((k_spmm_tiles<G, T, F, MODE, false, UNI>), dim3((unsigned)blocks), dim3(G::kThreads), 0, s,
                           P.tiles.p, P.ntiles, tiles_per_xcd, nchunks, chunk, chunk_major, P.key.p, vals,
                           P.cols.p, P.rseg.p,
                           x, sc, sf, y, dc, df, nf, scale, P.rowmap.p);
}

// Field chunks of a tile launch (each chunk is a workgroup per tile): enough
// (tile, chunk) workgroups to keep every CU busy several times over when the
// plan has few tiles, but each chunk past the first reads the tile's slots
// again, so no more chunks than keep that re-read within a tenth of the
// launch's field bytes (a plan with many tiles over a few dozen fields, whose
// slots are a large share of the traffic, gets one chunk).
int default_chunks(const TilePlan &P, int64_t nf, int F, int slot_bytes, int64_t field_bytes) {
    const int64_t want = P.geom == kGeomL ? 8192 : 2048;
    int64_t c = (want + P.ntiles - 1) / std::max<int64_t>(1, P.ntiles);
    c = std::max<int64_t>(1, std::min<int64_t>(c, (nf + F - 1) / F));
    const double reread = (double)P.nslots * slot_bytes;
    while (c > 1 && (double)(c - 1) * reread > 0.1 * (double)field_bytes) --c;
    return (int)c;
}

// F: fields per LDS stage; acc: FP32 fields summed with Kahan (0), in FP64 from
// the FP64 A (1), in FP64 from A rounded to FP32 (2, 1024-row tiles only).
// FP32 fields use kRounds plans (wave-uniform segment ends), FP64 fields kEven
// plans (no padding; the same segments for any field count, so a batched FP64
// apply equals the loop of single-field applies bit for bit).
template <class T>
void launch_tiles(const TilePlan &P, int F, int acc, int nchunks, int chunk_major, const T *x, int64_t sc,
                  int64_t sf, T *y, int64_t dc, int64_t df, int64_t nf, const double *scale, int64_t nrows,
                  int64_t ncols, hipStream_t s) {
    if (P.ntiles == 0) return;
    const bool acc64 = acc != 0;
    if (nchunks <= 0) {
        const int slot_bytes = 4 + (sizeof(T) == 4 && acc != 1 ? 4 : 8);  // key + value as the kernel reads it
        nchunks = default_chunks(P, nf, F, slot_bytes, nf * (nrows + ncols) * (int64_t)sizeof(T));
    }
    nchunks = (int)std::max<int64_t>(1, std::min<int64_t>(nchunks, (nf + F - 1) / F));
#define CR_TILE_ARGS P, nchunks, chunk_major, x, sc, sf, y, dc, df, nf, scale, nrows, ncols, s
    if constexpr (sizeof(T) == 4) {
        if (P.geom == kGeomL) {
            if (!acc64) launch_f<GeomL, T, 4, 0, true>(CR_TILE_ARGS);
            else if (acc == 2) launch_f<GeomL, T, 4, 2, true>(CR_TILE_ARGS);
            else if (F == 2) launch_f<GeomL, T, 2, 1, true>(CR_TILE_ARGS);
            else launch_f<GeomL, T, 4, 1, true>(CR_TILE_ARGS);
        } else if (acc64) {
            CR_REQUIRE(acc == 1, CR_UNSUPPORTED, "FP32 values of A are kept for 1024-row tiles only");
            switch (F) {
                case 2: launch_f<GeomS, T, 2, 1, true>(CR_TILE_ARGS); break;
                case 8: launch_f<GeomS, T, 8, 1, true>(CR_TILE_ARGS); break;
                default: launch_f<GeomS, T, 4, 1, true>(CR_TILE_ARGS); break;
            }
        } else {
            switch (F) {
                case 2: launch_f<GeomS, T, 2, 0, true>(CR_TILE_ARGS); break;
                case 8: launch_f<GeomS, T, 8, 0, true>(CR_TILE_ARGS); break;
                default: launch_f<GeomS, T, 4, 0, true>(CR_TILE_ARGS); break;
            }
        }
    } else {
        (void)acc64;
        CR_REQUIRE(P.geom == kGeomS, CR_UNSUPPORTED, "FP64 fields use 512-row tiles");
        switch (F) {
            case 1: launch_f<GeomS, T, 1, 0, false>(CR_TILE_ARGS); break;
            case 2: launch_f<GeomS, T, 2, 0, false>(CR_TILE_ARGS); break;
            default: launch_f<GeomS, T, 4, 0, false>(CR_TILE_ARGS); break;
        }
    }
#undef CR_TILE_ARGS
}

template void launch_tiles<float>(const TilePlan &, int, int, int, int, const float *, int64_t, int64_t,
                                  float *, int64_t, int64_t, int64_t, const double *, int64_t, int64_t,
                                  hipStream_t);
template void launch_tiles<double>(const TilePlan &, int, int, int, int, const double *, int64_t, int64_t,
                                   double *, int64_t, int64_t, int64_t, const double *, int64_t, int64_t,
                                   hipStream_t);

}  // namespace cr
