// build.hip -- Regridder(dst, src) on the device.
//
// Reference pipeline (src/regridder/regridder.jl:125-163,
// src/regridder/intersection_areas.jl:67-157) and what replaces each step:
//   treeify + dual-tree DFS over spherical caps (MultithreadedDualDepthFirstSearch.jl)
//     -> k_cell_bounds: FP64 area + bounding cap of every cell
//        space-filling-curve sort (cube face + Hilbert) of source caps +
//        implicit 8-ary tree of caps (k_build_level)
//        k_search<false/true>: one wave per curve-order cluster of 64 destination
//        cells walks the tree with the cluster's cap, then every lane tests its
//        cell against the cluster's leaf list; count pass + scan + fill pass
//        emit candidates grouped by destination row in a deterministic order
//   compute_intersection_areas with ConvexConvexSutherlandHodgman + GO.area
//     -> k_clip_compact<4>: a uniform outside-one-edge test settles about a
//        quarter of the candidates; the rest are queued in LDS and clipped 64
//        at a time, one thread per candidate: streaming FP64 clip + area
//   keep iff area > 0 (here > threshold_rel * min(a_src, a_dst)) and
//   sparse(dst, src, areas)
//     -> k_keep_count / k_keep_scatter: wave64 ballot compaction of kept pairs;
//        the rows of CSR(A) stay grouped as the candidates were (row pointers
//        from the keep mask), one stable radix sort by column gives CSR(A^T)
//        (host copies of CSR(A) sort each row's columns, like sparse())
//   areas(manifold, tree)  -> the k_cell_bounds areas, times radius^2
//   normalize!             -> k_scale by 1 / max(A)
#include <limits>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <vector>

#include "planar.hpp"
#include "regridder.hpp"
#include "sphgeom.hpp"
#include "tree.hpp"

namespace cr {

// ---- kernels over tree.hpp -------------------------------------------------
// Area (unit sphere) and bounding cap of each cell.
__global__ void k_cell_bounds(const double *__restrict__ cells, int64_t n, int nv,
                              double *__restrict__ area, float4 *__restrict__ cap) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double *p = cells + 3 * nv * c;
    area[c] = ring_area(p, nv);
    cap[c] = cell_cap(p, nv);
}

// The same for cells of NV vertices, 64 per block: the block's cells are one
// contiguous run of 64*NV*3 doubles, read with coalesced 16-byte loads into LDS
// (a thread-per-cell read strides 24*NV bytes between lanes), then every lane
// works on its own cell from LDS.
template <int NV>
__global__ void __launch_bounds__(64) k_cell_bounds_staged(const double *__restrict__ cells, int64_t n,
                                                           double *__restrict__ area, float4 *__restrict__ cap) {
    constexpr int kPer = 3 * NV;
    __shared__ double2 buf[64 * kPer / 2];
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int64_t cnt = min((int64_t)64, n - c0);
    const double2 *src = reinterpret_cast<const double2 *>(cells + kPer * c0);
    const int nload = (int)(cnt * kPer / 2);  // kPer*cnt is even for even NV
    for (int t = threadIdx.x; t < nload; t += 64) buf[t] = src[t];
    __syncthreads();
    if (threadIdx.x >= cnt) return;
    const double *p = reinterpret_cast<const double *>(buf) + kPer * threadIdx.x;
    area[c0 + threadIdx.x] = ring_area(p, NV);
    cap[c0 + threadIdx.x] = cell_cap(p, NV);
}

void launch_cell_bounds(const cr_grid *g, double *area, float4 *cap, hipStream_t s) {
    const int64_t n = g->ncells;
    const unsigned blocks = grid_for(n, 64);
    switch (g->nv) {
    case 4: hipLaunchKernelGGL(k_cell_bounds_staged<4>, dim3(blocks), dim3(64), 0, s, g->cells.p, n, area, cap); break;
    case 8: hipLaunchKernelGGL(k_cell_bounds_staged<8>, dim3(blocks), dim3(64), 0, s, g->cells.p, n, area, cap); break;
    case 16: hipLaunchKernelGGL(k_cell_bounds_staged<16>, dim3(blocks), dim3(64), 0, s, g->cells.p, n, area, cap); break;
    default:
        hipLaunchKernelGGL(k_cell_bounds, dim3(grid_for(n, 256)), dim3(256), 0, s, g->cells.p, n, g->nv, area, cap);
    }
}

// Curve keys of n caps (with `tag` or-ed in: the top bit tells grids apart in
// a combined sort) and the identity permutation.
__global__ void k_sfc_keys(const float4 *__restrict__ cap, int64_t n, uint32_t tag, uint32_t *__restrict__ key,
                           int32_t *__restrict__ idx) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    key[c] = sfc_key(cap[c]) | tag;
    idx[c] = (int32_t)c;
}

__global__ void k_gather_caps(const float4 *__restrict__ cap, const int32_t *__restrict__ perm,
                              int64_t n, float4 *__restrict__ leaf) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) leaf[i] = cap[perm[i]];
}

__global__ void k_build_level(const float4 *__restrict__ child, int64_t nchild,
                              float4 *__restrict__ parent, int64_t nparent) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nparent) return;
    int64_t c0 = p * kTreeFanout, c1 = min(nchild, c0 + kTreeFanout);
    parent[p] = parent_cap(child, c0, c1);
}

// Candidate search, one wave per 64 destination cells taken in space-filling-
// curve order of their caps (a compact cluster).  The wave bounds its 64 caps
// by one cap, walks the source tree level by level with that cap (lanes test
// the children of the current frontier side by side; ballot + prefix count
// compact the survivors into LDS), and ends with the list of source leaves
// that touch the cluster.  Each lane then tests its own cap against that list
// (LDS broadcast reads, no divergence).  A cluster whose frontier or leaf list
// overflows the LDS lists is retried as 4 clusters of 16 lanes; the cells of
// a 16-lane cluster that still overflows are walked one at a time by the
// whole wave, which counts or writes the hits straight from its ballots (no
// leaf list); only a cell whose frontier overflows even then walks the tree
// alone (tree_query).  Every leaf that touches a cell's cap touches any cluster cap
// that contains it and all of its ancestors, so every route emits exactly the
// leaves whose caps touch the cell's cap, in a fixed order.  FILL=false
// counts; FILL=true writes at the scanned offsets (never past the cell's
// segment).
constexpr int kSearchLeaves = 1024;
constexpr int kSearchFront = 256;

template <int W>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int m = W / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
template <int W>
__device__ __forceinline__ float group_max(float v) {
#pragma unroll
    for (int m = W / 2; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    return v;
}

// Cap around the caps of the W-lane group of this lane (lanes whose `in` is
// false are left out); w < 0 marks an empty group, w >= 2 a group too wide to
// bound (both handled as overflow).  Uniform over each group.
template <int W>
__device__ float4 group_cap(const float4 &c, bool in) {
    const float sx = group_sum<W>(in ? c.x : 0.f), sy = group_sum<W>(in ? c.y : 0.f),
                sz = group_sum<W>(in ? c.z : 0.f);
    const float L = sqrtf(sx * sx + sy * sy + sz * sz);
    const float wide = group_max<W>(in && c.w >= 2.0f ? 1.f : 0.f);
    const float cnt = group_sum<W>(in ? 1.f : 0.f);
    if (cnt == 0.f) return make_float4(0.f, 0.f, 1.f, -1.f);
    if (wide > 0.f || L < 1e-3f) return make_float4(0.f, 0.f, 1.f, kCapAll);
    const float mx = sx / L, my = sy / L, mz = sz / L;
    float dd = 0.f;
    if (in) {
        const float dx = c.x - mx, dy = c.y - my, dz = c.z - mz;
        dd = sqrtf(dx * dx + dy * dy + dz * dz) + c.w;
    }
    return make_float4(mx, my, mz, group_max<W>(dd) * (1.0f + 1e-5f) + 1e-5f);
}

__device__ __forceinline__ float4 shfl4(const float4 &v, int src) {
    return make_float4(__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64), __shfl(v.w, src, 64));
}

// The whole wave walks the tree with the (wave-uniform) cap wc and lists the
// leaves that touch it in lcap/lid.  Returns the count, or -1 when a frontier
// or the leaf list overflows (or wc cannot be bounded).
__device__ int coop_collect(const float4 &wc, const float4 *__restrict__ nodes, const Levels &lv,
                            const int32_t *__restrict__ perm, float4 *lcap, int32_t *lid,
                            int32_t (*front)[kSearchFront]) {
    const int lane = threadIdx.x;
    if (wc.w < 0.0f || wc.w >= 1.4f) return -1;
    if (lv.top == 0) {  // a single leaf
        __syncthreads();
        if (lane == 0) {
            lcap[0] = nodes[0];
            lid[0] = perm[0];
        }
        __syncthreads();
        return 1;
    }
    __syncthreads();  // the previous list is no longer read
    if (lane == 0) front[0][0] = 0;
    __syncthreads();
    int cur = 0, nf = 1;
    for (int lev = lv.top; lev >= 1; --lev) {
        const int64_t csize = lv.size[lev - 1], coff = lv.off[lev - 1];
        const int tot = nf * kTreeFanout;
        int nn = 0;
        // four chunks of 64 children per step: their loads are issued together
        for (int base = 0; base < tot; base += 4 * 64) {
            int64_t child[4];
            float4 cc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = base + 64 * u + lane;
                child[u] = -1;
                cc[u] = make_float4(0.f, 0.f, 0.f, -8.0f);
                if (t < tot) {
                    const int64_t ch = (int64_t)front[cur][t >> 3] * kTreeFanout + (t & 7);
                    if (ch < csize) {
                        child[u] = ch;
                        cc[u] = nodes[coff + ch];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool hit = child[u] >= 0 && caps_overlap(wc, cc[u]);
                const uint64_t m = __ballot(hit);
                const int pos = nn + (int)__popcll(m & ((1ull << lane) - 1ull));
                if (hit) {
                    if (lev > 1) {
                        if (pos < kSearchFront) front[cur ^ 1][pos] = (int32_t)child[u];
                    } else if (pos < kSearchLeaves) {
                        lcap[pos] = cc[u];
                        lid[pos] = perm[child[u]];
                    }
                }
                nn += (int)__popcll(m);
            }
        }
        __syncthreads();
        if (nn > (lev > 1 ? kSearchFront : kSearchLeaves)) return -1;
        cur ^= 1;
        nf = nn;
    }
    return nf;
}

// The whole wave walks the tree once with the four 16-lane cluster caps
// gc[0..3] (wave-uniform): the frontier is the union of what touches any of
// them, and the leaf level lists, for each cluster g, the leaves that touch
// gc[g] in lcap/lid[g * kSubLeaves ...] (n[g] of them, in tree order).  Each
// lane then tests only its own cluster's list, a quarter of the lanes' region
// and far fewer leaves than one 64-cell list.  Returns false when the union
// frontier or a list overflows or a cap cannot be bounded; an empty cluster
// (w < 0) gets an empty list.
constexpr int kSubLeaves = kSearchLeaves / 4;

__device__ bool coop_collect4(const float4 (&gc)[4], const float4 *__restrict__ nodes, const Levels &lv,
                              const int32_t *__restrict__ perm, float4 *lcap, int32_t *lid,
                              int32_t (*front)[kSearchFront], int (&n)[4]) {
    const int lane = threadIdx.x;
    bool live[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        if (gc[g].w >= 1.4f) return false;
        live[g] = gc[g].w >= 0.0f;
        n[g] = 0;
    }
    if (lv.top == 0) {  // a single leaf
        __syncthreads();
#pragma unroll
        for (int g = 0; g < 4; ++g)
            if (live[g] && caps_overlap(gc[g], nodes[0])) {
                if (lane == 0) {
                    lcap[g * kSubLeaves] = nodes[0];
                    lid[g * kSubLeaves] = perm[0];
                }
                n[g] = 1;
            }
        __syncthreads();
        return true;
    }
    __syncthreads();  // the previous lists are no longer read
    if (lane == 0) front[0][0] = 0;
    __syncthreads();
    int cur = 0, nf = 1;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int lev = lv.top; lev >= 1; --lev) {
        const int64_t csize = lv.size[lev - 1], coff = lv.off[lev - 1];
        const int tot = nf * kTreeFanout;
        int nn = 0;
        for (int base = 0; base < tot; base += 4 * 64) {
            int64_t child[4];
            float4 cc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = base + 64 * u + lane;
                child[u] = -1;
                cc[u] = make_float4(0.f, 0.f, 0.f, -8.0f);
                if (t < tot) {
                    const int64_t ch = (int64_t)front[cur][t >> 3] * kTreeFanout + (t & 7);
                    if (ch < csize) {
                        child[u] = ch;
                        cc[u] = nodes[coff + ch];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                bool h[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) h[g] = child[u] >= 0 && live[g] && caps_overlap(gc[g], cc[u]);
                if (lev > 1) {
                    const bool hit = h[0] || h[1] || h[2] || h[3];
                    const uint64_t m = __ballot(hit);
                    const int pos = nn + (int)__popcll(m & below);
                    if (hit && pos < kSearchFront) front[cur ^ 1][pos] = (int32_t)child[u];
                    nn += (int)__popcll(m);
                } else {
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const uint64_t m = __ballot(h[g]);
                        const int pos = n[g] + (int)__popcll(m & below);
                        if (h[g] && pos < kSubLeaves) {
                            lcap[g * kSubLeaves + pos] = cc[u];
                            lid[g * kSubLeaves + pos] = perm[child[u]];
                        }
                        n[g] += (int)__popcll(m);
                    }
                }
            }
        }
        __syncthreads();
        if (lev > 1) {
            if (nn > kSearchFront) return false;
            cur ^= 1;
            nf = nn;
        }
    }
#pragma unroll
    for (int g = 0; g < 4; ++g)
        if (n[g] > kSubLeaves) return false;
    return true;
}

// One lane's tests against the collected list (four independent tests per
// step: a wave per SIMD has no other wave to hide a dependent chain).
template <bool FILL>
__device__ void lane_emit(const float4 &c, int64_t d, int nleaf, const float4 *lcap, const int32_t *lid,
                          int64_t *__restrict__ count_or_offset, int32_t *__restrict__ cand_src,
                          int32_t *__restrict__ cand_dst) {
    if (FILL) {
        int64_t pos = count_or_offset[d];
        const int64_t end = count_or_offset[d + 1];
        int j = 0;
        for (; j + 4 <= nleaf; j += 4) {
            bool h[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) h[u] = caps_overlap(c, lcap[j + u]);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (h[u] && pos < end) {
                    cand_src[pos] = lid[j + u];
                    cand_dst[pos] = (int32_t)d;
                    ++pos;
                }
        }
        for (; j < nleaf; ++j)
            if (caps_overlap(c, lcap[j]) && pos < end) {
                cand_src[pos] = lid[j];
                cand_dst[pos] = (int32_t)d;
                ++pos;
            }
    } else {
        int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        int j = 0;
        for (; j + 4 <= nleaf; j += 4) {
            c0 += caps_overlap(c, lcap[j]);
            c1 += caps_overlap(c, lcap[j + 1]);
            c2 += caps_overlap(c, lcap[j + 2]);
            c3 += caps_overlap(c, lcap[j + 3]);
        }
        for (; j < nleaf; ++j) c0 += caps_overlap(c, lcap[j]);
        count_or_offset[d] = (int64_t)(c0 + c1 + c2 + c3);
    }
}

template <bool FILL>
__device__ void lane_walk(const float4 &c, int64_t d, const float4 *__restrict__ nodes, const Levels &lv,
                          const int32_t *__restrict__ perm, int64_t *__restrict__ count_or_offset,
                          int32_t *__restrict__ cand_src, int32_t *__restrict__ cand_dst) {
    if (FILL) {
        int64_t pos = count_or_offset[d];
        const int64_t end = count_or_offset[d + 1];
        tree_query(c, nodes, lv, [&](int64_t leaf) {
            if (pos < end) {
                cand_src[pos] = perm[leaf];
                cand_dst[pos] = (int32_t)d;
                ++pos;
            }
        });
    } else {
        int64_t cnt = 0;
        tree_query(c, nodes, lv, [&](int64_t) { ++cnt; });
        count_or_offset[d] = cnt;
    }
}

// One cell (cap qc, row d, wave-uniform) walked by the whole wave; the leaf
// level is tested 64 children at a time and the hits are counted, or written
// in order at the cell's offset, straight from the ballot (no leaf list).
// The frontier lives in `scratch` (kSearchBigFront entries).  Returns false
// when a frontier overflows (nothing was counted or written then).
constexpr int kSearchBigFront = kSearchLeaves * 4;  // int32 entries in the lcap storage

template <bool FILL>
__device__ bool coop_emit_one(const float4 &qc, int64_t d, const float4 *__restrict__ nodes, const Levels &lv,
                              const int32_t *__restrict__ perm, int32_t *scratch,
                              int64_t *__restrict__ count_or_offset, int32_t *__restrict__ cand_src,
                              int32_t *__restrict__ cand_dst) {
    const int lane = threadIdx.x;
    int32_t *fa = scratch, *fb = scratch + kSearchBigFront / 2;
    const int cap = kSearchBigFront / 2;
    __syncthreads();
    if (lane == 0) fa[0] = 0;
    __syncthreads();
    int nf = 1;
    int64_t pos = 0, end = 0;
    if (FILL) {
        pos = count_or_offset[d];
        end = count_or_offset[d + 1];
    }
    int64_t cnt = 0;
    if (lv.top == 0) {
        const bool hit = lane == 0 && caps_overlap(qc, nodes[0]);
        if (hit) {
            if (FILL && pos < end) {
                cand_src[pos] = perm[0];
                cand_dst[pos] = (int32_t)d;
            }
        }
        cnt = (int64_t)__popcll(__ballot(hit));
    }
    for (int lev = lv.top; lev >= 1; --lev) {
        const int64_t csize = lv.size[lev - 1], coff = lv.off[lev - 1];
        const int tot = nf * kTreeFanout;
        int nn = 0;
        for (int base = 0; base < tot; base += 4 * 64) {
            int64_t child[4];
            float4 cc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = base + 64 * u + lane;
                child[u] = -1;
                cc[u] = make_float4(0.f, 0.f, 0.f, -8.0f);
                if (t < tot) {
                    const int64_t ch = (int64_t)fa[t >> 3] * kTreeFanout + (t & 7);
                    if (ch < csize) {
                        child[u] = ch;
                        cc[u] = nodes[coff + ch];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool hit = child[u] >= 0 && caps_overlap(qc, cc[u]);
                const uint64_t m = __ballot(hit);
                const int rank = (int)__popcll(m & ((1ull << lane) - 1ull));
                if (lev > 1) {
                    if (hit && nn + rank < cap) fb[nn + rank] = (int32_t)child[u];
                    nn += (int)__popcll(m);
                } else {
                    if (FILL && hit && pos + cnt + rank < end) {
                        cand_src[pos + cnt + rank] = perm[child[u]];
                        cand_dst[pos + cnt + rank] = (int32_t)d;
                    }
                    cnt += (int64_t)__popcll(m);
                }
            }
        }
        __syncthreads();
        if (lev > 1) {
            if (nn > cap) return false;
            int32_t *t = fa;
            fa = fb;
            fb = t;
            nf = nn;
        }
    }
    if (!FILL && lane == 0) count_or_offset[d] = cnt;
    return true;
}

template <bool FILL>
__global__ void __launch_bounds__(64) k_search(const float4 *__restrict__ qcap,
                                               const int32_t *__restrict__ qorder, int64_t nq,
                                               const float4 *__restrict__ nodes, Levels lv,
                                               const int32_t *__restrict__ perm,
                                               int64_t *__restrict__ count_or_offset,
                                               int32_t *__restrict__ cand_src,
                                               int32_t *__restrict__ cand_dst,
                                               int32_t *__restrict__ defer, unsigned *__restrict__ ndefer) {
    __shared__ float4 lcap[kSearchLeaves];
    __shared__ int32_t lid[kSearchLeaves];
    __shared__ int32_t front[2][kSearchFront];
    const int lane = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const bool active = i < nq;
    const int64_t d = active ? qorder[i] : 0;
    const float4 c = active ? qcap[d] : make_float4(0.f, 0.f, 0.f, 0.f);

    // four 16-cell clusters in one walk
    const float4 gcap = group_cap<16>(c, active);
    {
        float4 gcs[4];
        int n4[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) gcs[g] = shfl4(gcap, 16 * g);
        if (coop_collect4(gcs, nodes, lv, perm, lcap, lid, front, n4)) {
            const int g = lane >> 4;
            const int n = g == 0 ? n4[0] : (g == 1 ? n4[1] : (g == 2 ? n4[2] : n4[3]));
            if (active)
                lane_emit<FILL>(c, d, n, lcap + g * kSubLeaves, lid + g * kSubLeaves, count_or_offset, cand_src,
                                cand_dst);
            return;
        }
    }
    // the four clusters one after the other with full-size lists; the lanes of
    // a cluster that still overflows walk the tree alone, all together at the end
    int n;
    bool walk = false;
    for (int g = 0; g < 4; ++g) {
        const float4 gc = shfl4(gcap, 16 * g);
        n = coop_collect(gc, nodes, lv, perm, lcap, lid, front);
        if ((lane >> 4) != g) continue;
        if (n >= 0) {
            if (active) lane_emit<FILL>(c, d, n, lcap, lid, count_or_offset, cand_src, cand_dst);
        } else {
            walk = true;
        }
    }
    // cells left over (about 0.15% of them, with hundreds to thousands of
    // candidates each): listed by the count pass for k_search_deferred, which
    // gives each its own wave (walked here one at a time they would hold this
    // wave for a long tail)
    if (!FILL) {
        const bool mine = walk && active;
        const uint64_t m = __ballot(mine);
        if (m) {
            const int leader = (int)__ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(ndefer, (unsigned)__popcll(m));
            base = (unsigned)__shfl((int)base, leader, 64);
            if (mine) defer[base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = (int32_t)d;
        }
    }
}

// The listed cells, one wave each (grid-stride over the list): the whole wave
// walks the tree with the cell's cap and counts or writes the hits straight
// from its ballots; a cell whose frontier overflows even then is walked by
// one lane alone (tree_query).
template <bool FILL>
__global__ void __launch_bounds__(64) k_search_deferred(const float4 *__restrict__ qcap,
                                                        const int32_t *__restrict__ defer,
                                                        const unsigned *__restrict__ ndefer,
                                                        const float4 *__restrict__ nodes, Levels lv,
                                                        const int32_t *__restrict__ perm,
                                                        int64_t *__restrict__ count_or_offset,
                                                        int32_t *__restrict__ cand_src,
                                                        int32_t *__restrict__ cand_dst) {
    __shared__ int32_t scratch[kSearchBigFront];
    const unsigned n = *ndefer;
    for (unsigned i = blockIdx.x; i < n; i += gridDim.x) {
        const int64_t d = defer[i];
        const float4 qc = qcap[d];
        if (!coop_emit_one<FILL>(qc, d, nodes, lv, perm, scratch, count_or_offset, cand_src, cand_dst) &&
            threadIdx.x == 0)
            lane_walk<FILL>(qc, d, nodes, lv, perm, count_or_offset, cand_src, cand_dst);
    }
}

// Candidates whose subject lies outside one clip edge (about a quarter on the
// benchmark grids) or entirely inside the clip polygon (area = the subject's
// area) are settled by a cheap uniform test; the others are queued in LDS by
// the number of clip edges their subject crosses and clipped 64 at a time by a
// pipeline with one stage per crossed edge, so every lane of the wave runs the
// same pipeline on a real pair.  One wave per kClipChunks * 64 candidates;
// results go to each candidate's own slot.  (Used with a positive drop
// threshold: a pair the test settles could at most leave a sub-rounding sliver
// in the pipeline, which the threshold drops too.)
constexpr int kClipChunks = 8;

// The kept value of a pair: its area if above threshold_rel times the smaller
// cell area, else 0.  A zero-area cell (a padding ghost such as the duplicated
// half of a tripolar fold row, ext/ConservativeRegriddingOceananigansExt.jl:65-115,
// never a tree candidate there) intersects nothing: as a clip polygon its
// zero-length edges would pass the whole subject.
// The compaction counts the two other outcomes: kZeroCellPair marks a pair
// dropped because one of its cells has zero area, and a non-finite area stays
// as it is (the build then fails and names the pair).
constexpr double kZeroCellPair = -1.0;
__device__ __forceinline__ double kept_area(double a, double amin, double thr) {
    if (!(amin > 0.0)) return kZeroCellPair;
    if (!isfinite(a)) return a;
    return a > thr * amin ? a : 0.0;
}

template <int NS, int NC>
__device__ __forceinline__ void clip_candidate_masked(int64_t c, unsigned mask, const int32_t *__restrict__ cand_dst,
                                                      const int32_t *__restrict__ cand_src,
                                                      const double *__restrict__ dst_cells, int dst_nv,
                                                      const double *__restrict__ src_cells, int src_nv,
                                                      int clip_is_dst, const double *__restrict__ a_dst,
                                                      const double *__restrict__ a_src, double thr,
                                                      double *__restrict__ val) {
    const int64_t d = cand_dst[c], s = cand_src[c];
    const double *dp = dst_cells + 3 * (int64_t)dst_nv * d;
    const double *sp = src_cells + 3 * (int64_t)src_nv * s;
    const double *clip = clip_is_dst ? dp : sp, *subj = clip_is_dst ? sp : dp;
    const int nc = clip_is_dst ? dst_nv : src_nv, ns = clip_is_dst ? src_nv : dst_nv;
    const double a = clip_area_stream_masked<NS, NC>(clip, nc, mask, subj, ns);
    val[c] = kept_area(a, fmin(a_src[s], a_dst[d]), thr);
}

// Clip the queued candidates 64 at a time with an NS-stage pipeline (entries:
// local index | crossed-edge mask << 16; a mask of fewer bits gets pass-through
// stages).  Full batches only, unless `drain`.  Block = one wave, so qn is
// uniform.
template <int NS, int NC>
__device__ __forceinline__ void clip_queued(int32_t *q, int &qn, bool drain, int64_t base,
                                            const int32_t *__restrict__ cand_dst, const int32_t *__restrict__ cand_src,
                                            const double *__restrict__ dst_cells, int dst_nv,
                                            const double *__restrict__ src_cells, int src_nv, int clip_is_dst,
                                            const double *__restrict__ a_dst, const double *__restrict__ a_src,
                                            double thr, double *__restrict__ val) {
    const int lane = threadIdx.x;
    while (qn >= 64 || (drain && qn > 0)) {
        if (lane < qn) {
            const int32_t e = q[lane];
            clip_candidate_masked<NS, NC>(base + (e & 0xffff), pad_mask((unsigned)e >> 16, NS), cand_dst, cand_src,
                                          dst_cells, dst_nv, src_cells, src_nv, clip_is_dst, a_dst, a_src, thr, val);
        }
        const int left = qn > 64 ? qn - 64 : 0;
        const int moved = lane < left ? q[64 + lane] : 0;
        __syncthreads();
        if (lane < left) q[lane] = moved;
        qn = left;
        __syncthreads();
    }
}

// Move a queue's entries (fewer than 64) to the end of a longer-stage queue.
__device__ __forceinline__ void promote(const int32_t *from, int &nfrom, int32_t *to, int &nto) {
    if ((int)threadIdx.x < nfrom) to[nto + threadIdx.x] = from[threadIdx.x];
    nto += nfrom;
    nfrom = 0;
    __syncthreads();
}

template <int NC>
__global__ void __launch_bounds__(64) k_clip_compact(const int32_t *__restrict__ cand_dst,
                                                     const int32_t *__restrict__ cand_src, int64_t ncand,
                                                     const double *__restrict__ dst_cells, int dst_nv,
                                                     const double *__restrict__ src_cells, int src_nv,
                                                     int clip_is_dst, const double *__restrict__ a_dst,
                                                     const double *__restrict__ a_src, double thr,
                                                     double *__restrict__ val) {
    static_assert(NC == 4, "stage queues are for 4-vertex clip polygons");
    // queues of candidates that cross 1-2, 3 and 4 clip edges (a pipeline stage
    // per crossed edge: about 2.3 stages instead of 4 on the benchmark grids)
    __shared__ int32_t queue[3][128];
    const int lane = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * (64 * kClipChunks);
    int qn0 = 0, qn1 = 0, qn2 = 0;
    for (int ch = 0; ch <= kClipChunks; ++ch) {
        if (ch < kClipChunks) {
            const int64_t c = base + ch * 64 + lane;
            bool keep = false;
            unsigned crossed = 0;
            if (c < ncand) {
                const int64_t d = cand_dst[c], s = cand_src[c];
                const double *dp = dst_cells + 3 * (int64_t)dst_nv * d;
                const double *sp = src_cells + 3 * (int64_t)src_nv * s;
                const double *clip = clip_is_dst ? dp : sp, *subj = clip_is_dst ? sp : dp;
                const int nc = clip_is_dst ? dst_nv : src_nv, ns = clip_is_dst ? src_nv : dst_nv;
                const int where = classify_subject<NC>(clip, nc, subj, ns, &crossed);
                keep = where == kSubjectCrosses;
                if (where == kSubjectOutside) val[c] = 0.0;
                if (where == kSubjectInside) {  // the subject's own (unit-sphere) area
                    const double a = clip_is_dst ? a_src[s] : a_dst[d];
                    val[c] = kept_area(a, fmin(a_src[s], a_dst[d]), thr);
                }
            }
            const int stages = __builtin_popcount(crossed);
            const int qi = stages <= 2 ? 0 : stages - 2;
            const int32_t entry = (ch * 64 + lane) | (int32_t)(crossed << 16);
            const uint64_t below = (1ull << lane) - 1ull;
            const uint64_t m0 = __ballot(keep && qi == 0), m1 = __ballot(keep && qi == 1),
                           m2 = __ballot(keep && qi == 2);
            if (keep && qi == 0) queue[0][qn0 + (int)__popcll(m0 & below)] = entry;
            if (keep && qi == 1) queue[1][qn1 + (int)__popcll(m1 & below)] = entry;
            if (keep && qi == 2) queue[2][qn2 + (int)__popcll(m2 & below)] = entry;
            qn0 += (int)__popcll(m0);
            qn1 += (int)__popcll(m1);
            qn2 += (int)__popcll(m2);
            __syncthreads();
        }
        // full batches; after the last chunk the short queues' remainders move
        // up (a longer pipeline clips them too) and the 4-stage queue drains.
        // One call site per pipeline length.
        const bool last = ch == kClipChunks;
        if (last) promote(queue[0], qn0, queue[1], qn1);
        clip_queued<2, NC>(queue[0], qn0, false, base, cand_dst, cand_src, dst_cells, dst_nv, src_cells, src_nv,
                           clip_is_dst, a_dst, a_src, thr, val);
        clip_queued<3, NC>(queue[1], qn1, false, base, cand_dst, cand_src, dst_cells, dst_nv, src_cells, src_nv,
                           clip_is_dst, a_dst, a_src, thr, val);
        if (last) promote(queue[1], qn1, queue[2], qn2);
        clip_queued<4, NC>(queue[2], qn2, last, base, cand_dst, cand_src, dst_cells, dst_nv, src_cells, src_nv,
                           clip_is_dst, a_dst, a_src, thr, val);
    }
}

template <int NC>
__global__ void __launch_bounds__(256) k_clip(const int32_t *__restrict__ cand_dst,
                                              const int32_t *__restrict__ cand_src, int64_t ncand,
                                              const double *__restrict__ dst_cells, int dst_nv,
                                              const double *__restrict__ src_cells, int src_nv,
                                              int clip_is_dst, const double *__restrict__ a_dst,
                                              const double *__restrict__ a_src, double thr,
                                              double *__restrict__ val) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand) return;
    int64_t d = cand_dst[c], s = cand_src[c];
    const double *dp = dst_cells + 3 * (int64_t)dst_nv * d;
    const double *sp = src_cells + 3 * (int64_t)src_nv * s;
    const double *clip = clip_is_dst ? dp : sp, *subj = clip_is_dst ? sp : dp;
    const int nc = clip_is_dst ? dst_nv : src_nv, ns = clip_is_dst ? src_nv : dst_nv;
    double a = clip_area_stream<NC>(clip, nc, subj, ns);
    val[c] = kept_area(a, fmin(a_src[s], a_dst[d]), thr);
}

// Generic clip for convex clip polygons with more than 4 vertices: classic
// buffered Sutherland-Hodgman in private memory.  Each clip edge adds at most
// one vertex to a convex subject's ring, so ns + nc <= 2 * kMaxCellVertices
// vertices always fit; the overflow branch (a NaN area, which fails the build)
// is a guard that cannot trigger.
constexpr int kGenericMaxV = 2 * kMaxCellVertices;

__device__ double clip_area_buffered(const double *clip, int nc, const double *subj, int ns) {
    d3 bufA[kGenericMaxV], bufB[kGenericMaxV];
    int m = min(ns, kGenericMaxV);
    for (int i = 0; i < m; ++i) bufA[i] = load3(subj + 3 * i);
    d3 *cur = bufA, *nxt = bufB;
    for (int k = 0; k < nc && m > 0; ++k) {
        d3 a = load3(clip + 3 * k), b = load3(clip + 3 * ((k + 1) % nc));
        d3 n = cross(a, b - a);
        if (dot(n, n) < kDegenerateEdge2) continue;
        int w = 0;
        d3 p = cur[m - 1];
        double sp = dot(n, p - a);
        for (int i = 0; i < m; ++i) {
            if (w > kGenericMaxV - 2) return __builtin_nan("");
            d3 q = cur[i];
            double sq = dot(n, q - a);
            if (sq >= 0.0) {
                if (sp < 0.0) nxt[w++] = crossing(p, sp, q, sq);
                nxt[w++] = q;
            } else if (sp >= 0.0) {
                nxt[w++] = crossing(p, sp, q, sq);
            }
            p = q;
            sp = sq;
        }
        m = w;
        d3 *t = cur;
        cur = nxt;
        nxt = t;
    }
    AreaAcc acc;
    acc.init();
    for (int i = 0; i < m; ++i) acc.push(cur[i]);
    return acc.area();
}

__global__ void k_clip_generic(const int32_t *__restrict__ cand_dst,
                               const int32_t *__restrict__ cand_src, int64_t ncand,
                               const double *__restrict__ dst_cells, int dst_nv,
                               const double *__restrict__ src_cells, int src_nv, int clip_is_dst,
                               const double *__restrict__ a_dst, const double *__restrict__ a_src,
                               double thr, double *__restrict__ val) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand) return;
    int64_t d = cand_dst[c], s = cand_src[c];
    const double *dp = dst_cells + 3 * (int64_t)dst_nv * d;
    const double *sp = src_cells + 3 * (int64_t)src_nv * s;
    double a = clip_is_dst ? clip_area_buffered(dp, dst_nv, sp, src_nv)
                           : clip_area_buffered(sp, src_nv, dp, dst_nv);
    val[c] = kept_area(a, fmin(a_src[s], a_dst[d]), thr);
}

// Planar clip + area per candidate (the Planar manifold): both rings gathered
// from the caller vertices, Sutherland-Hodgman in the plane (planar.hpp).
// MAXV holds a cell (nv <= MAXV) and bounds the clipped ring.
template <int MAXV>
__global__ void k_clip_planar(const int32_t *__restrict__ cand_dst, const int32_t *__restrict__ cand_src,
                              int64_t ncand, const double *__restrict__ dst_xy, const int32_t *__restrict__ dst_vi,
                              int dst_nv, const double *__restrict__ src_xy, const int32_t *__restrict__ src_vi,
                              int src_nv, int clip_is_dst, const double *__restrict__ a_dst,
                              const double *__restrict__ a_src, double thr, double *__restrict__ val) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand) return;
    const int64_t d = cand_dst[c], s = cand_src[c];
    p2 dr[MAXV], sr[MAXV];
    for (int k = 0; k < dst_nv; ++k) dr[k] = planar_vertex(dst_xy, dst_vi, d * dst_nv + k);
    for (int k = 0; k < src_nv; ++k) sr[k] = planar_vertex(src_xy, src_vi, s * src_nv + k);
    const double a = clip_is_dst ? planar_clip_area<MAXV>(dr, dst_nv, sr, src_nv)
                                 : planar_clip_area<MAXV>(sr, src_nv, dr, dst_nv);
    val[c] = kept_area(a, fmin(a_src[s], a_dst[d]), thr);
}

// ---- wave64 ballot compaction of kept candidates --------------------------
constexpr int kCompactBlock = 256;

// Failure accounting of a build (one device block, zeroed per build):
enum {
    kAcctNonfinitePairs = 0,  // candidate pairs with a non-finite area
    kAcctZeroCellPairs,       // candidate pairs dropped for a zero-area cell
    kAcctFirstBadPair,        // lowest such candidate index (initialised to all ones)
    kAcctZeroCells,           // zero-area cells, both grids
    kAcctNonfiniteCells,      // non-finite cell areas, both grids
    kAcctFirstBadCell,        // lowest such cell: destination index, or n_dst + source index
    kAcctWords
};

// Zero-area and non-finite cell areas of one grid (base: 0 for the
// destination, n_dst for the source in kAcctFirstBadCell).
__global__ void k_count_cells(const double *__restrict__ a, int64_t n, int64_t base,
                              unsigned long long *__restrict__ acct) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double v = c < n ? a[c] : 1.0;
    const uint64_t mz = __ballot(v == 0.0), mb = __ballot(!isfinite(v));
    if ((threadIdx.x & 63) == 0) {
        if (mz) atomicAdd(&acct[kAcctZeroCells], (unsigned long long)__popcll(mz));
        if (mb) {
            atomicAdd(&acct[kAcctNonfiniteCells], (unsigned long long)__popcll(mb));
            atomicMin(&acct[kAcctFirstBadCell], (unsigned long long)(base + c + __ffsll((long long)mb) - 1));
        }
    }
}

__global__ void __launch_bounds__(kCompactBlock) k_keep_count(const double *__restrict__ val,
                                                              int64_t n,
                                                              int64_t *__restrict__ block_cnt,
                                                              uint64_t *__restrict__ keep_mask,
                                                              unsigned long long *__restrict__ acct) {
    __shared__ int wave_cnt[kCompactBlock / 64];
    int64_t c = (int64_t)blockIdx.x * kCompactBlock + threadIdx.x;
    const double v = c < n ? val[c] : 0.0;
    bool keep = v > 0.0;  // not kZeroCellPair, 0 or NaN
    uint64_t m = __ballot(keep);
    const uint64_t mz = __ballot(v == kZeroCellPair), mb = __ballot(!isfinite(v));
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        wave_cnt[wave] = __popcll(m);
        keep_mask[(int64_t)blockIdx.x * (kCompactBlock / 64) + wave] = m;
        if (mz) atomicAdd(&acct[kAcctZeroCellPairs], (unsigned long long)__popcll(mz));
        if (mb) {
            atomicAdd(&acct[kAcctNonfinitePairs], (unsigned long long)__popcll(mb));
            atomicMin(&acct[kAcctFirstBadPair], (unsigned long long)(c + __ffsll((long long)mb) - 1));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kCompactBlock / 64; ++w) t += wave_cnt[w];
        block_cnt[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kCompactBlock) k_keep_scatter(
    const double *__restrict__ val, const int32_t *__restrict__ cand_dst,
    const int32_t *__restrict__ cand_src, int64_t n, const int64_t *__restrict__ block_off,
    int32_t *__restrict__ out_row, int32_t *__restrict__ out_col, double *__restrict__ out_val) {
    __shared__ int wave_base[kCompactBlock / 64];
    int64_t c = (int64_t)blockIdx.x * kCompactBlock + threadIdx.x;
    double v = c < n ? val[c] : 0.0;
    bool keep = v > 0.0;
    uint64_t m = __ballot(keep);
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_base[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kCompactBlock / 64; ++w) {
            int x = wave_base[w];
            wave_base[w] = t;
            t += x;
        }
    }
    __syncthreads();
    if (keep) {
        uint64_t below = lane ? (m & ((~0ull) >> (64 - lane))) : 0ull;
        int64_t o = block_off[blockIdx.x] + wave_base[wave] + __popcll(below);
        int32_t r = cand_dst[c];
        out_row[o] = r;
        out_col[o] = cand_src[c];
        out_val[o] = v;
    }
}

// Row pointers of the compacted pairs: candidates are grouped by destination
// row (row d owns candidates [cand_off[d], cand_off[d+1])), so row d starts at
// the number of kept candidates before cand_off[d] = block offset + popcounts
// of the keep mask inside the block.
__global__ void k_rowptr_from_mask(const int64_t *__restrict__ cand_off, int64_t nrows, int64_t ncand,
                                   const int64_t *__restrict__ block_off,
                                   const uint64_t *__restrict__ keep_mask, int64_t *__restrict__ rowptr) {
    int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d > nrows) return;
    const int64_t k = cand_off[d];
    if (k >= ncand) {
        rowptr[d] = block_off[(ncand + kCompactBlock - 1) / kCompactBlock];
        return;
    }
    const int64_t b = k / kCompactBlock;
    int64_t p = block_off[b];
    const int64_t w0 = b * (kCompactBlock / 64), kw = k >> 6;
    for (int64_t w = w0; w < kw; ++w) p += __popcll(keep_mask[w]);
    const int bit = (int)(k & 63);
    if (bit) p += __popcll(keep_mask[kw] & ((1ull << bit) - 1ull));
    rowptr[d] = p;
}

// Row pointers of a matrix whose entries are sorted by row key: rowptr[r] =
// first position with key >= r (binary search; empty rows need no care).
__global__ void k_rowptr_from_sorted(const int32_t *__restrict__ key, int64_t n, int64_t nrows,
                                     int64_t *__restrict__ rowptr) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrows) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (key[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    rowptr[r] = lo;
}

__global__ void k_iota(int32_t *__restrict__ x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (int32_t)i;
}

template <class T>
__global__ void k_gather(const T *__restrict__ src, const int32_t *__restrict__ idx, int64_t n,
                         T *__restrict__ dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_scale(double *__restrict__ x, int64_t n, double mul, double div) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (x[i] * mul) / div;
}

struct Timer {
    hipStream_t s;
    std::vector<hipEvent_t> ev;
    explicit Timer(hipStream_t st) : s(st) {}
    void mark() {
        hipEvent_t e;
        CR_HIP(hipEventCreate(&e));
        CR_HIP(hipEventRecord(e, s));
        ev.push_back(e);
    }
    double ms(int a, int b) const {
        float t = 0.f;
        if (hipEventElapsedTime(&t, ev[a], ev[b]) != hipSuccess) return 0.0;
        return t;
    }
    ~Timer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

template <class In, class Out>
void exclusive_sum(Temp &T, In in, Out out, int64_t n, hipStream_t s) {
    using V = typename std::iterator_traits<In>::value_type;
    size_t bytes = 0;
    CR_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, V(0), (size_t)n, rocprim::plus<V>(), s));
    void *p = T.get(bytes);
    CR_HIP(rocprim::exclusive_scan(p, bytes, in, out, V(0), (size_t)n, rocprim::plus<V>(), s));
}

template <class K, class V>
void sort_pairs(Temp &T, const K *kin, K *kout, const V *vin, V *vout, int64_t n, int end_bit,
                hipStream_t s) {
    size_t bytes = 0;
    CR_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)end_bit, s));
    void *p = T.get(bytes);
    CR_HIP(rocprim::radix_sort_pairs(p, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)end_bit, s));
}

// Instantiated here for the planner (plan.hip, planblocks.hip) too: their code
// objects then carry no device-wide sort or scan of their own, which keeps the
// first plan's one-time code-object load short.
template void exclusive_sum<int64_t *, int64_t *>(Temp &, int64_t *, int64_t *, int64_t, hipStream_t);
template void sort_pairs<int32_t, int32_t>(Temp &, const int32_t *, int32_t *, const int32_t *, int32_t *, int64_t,
                                           int, hipStream_t);
template void sort_pairs<uint32_t, int32_t>(Temp &, const uint32_t *, uint32_t *, const int32_t *, int32_t *,
                                            int64_t, int, hipStream_t);

int bits_for(int64_t n) {
    int b = 1;
    while ((int64_t(1) << b) < n) ++b;
    return b;
}

// CSR(A^T) of a device CSR(A) whose entry k lies in row coo_row[k]: a stable
// radix sort by column keeps the rows of A ascending inside each row of A^T.
void csr_transpose(const Csr &F, const int32_t *coo_row, Csr &Bk, Temp &T, hipStream_t s) {
    const int64_t nnz = F.nnz, ns = F.ncols;
    const int B = 256;
    Bk.nrows = ns; Bk.ncols = F.nrows; Bk.nnz = nnz;
    Bk.rowptr.alloc_kept(ns + 1, s);
    Bk.col.alloc_kept(nnz, s);
    Bk.val.alloc_kept(nnz, s);
    if (nnz == 0) {
        CR_HIP(hipMemsetAsync(Bk.rowptr.p, 0, Bk.rowptr.bytes(), s));
        return;
    }
    DevBuf<int32_t> iota(nnz, s), perm_t(nnz, s), tkey(nnz, s);
    hipLaunchKernelGGL(k_iota, dim3(grid_for(nnz, B)), dim3(B), 0, s, iota.p, nnz);
    sort_pairs(T, F.col.p, tkey.p, iota.p, perm_t.p, nnz, bits_for(ns), s);
    hipLaunchKernelGGL(k_gather<int32_t>, dim3(grid_for(nnz, B)), dim3(B), 0, s, coo_row, perm_t.p, nnz,
                       Bk.col.p);
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(nnz, B)), dim3(B), 0, s, F.val.p, perm_t.p, nnz,
                       Bk.val.p);
    hipLaunchKernelGGL(k_rowptr_from_sorted, dim3(grid_for(ns + 1, B)), dim3(B), 0, s, tkey.p, nnz, ns,
                       Bk.rowptr.p);
    CR_HIP(hipGetLastError());
}

static thread_local int64_t g_failed_dst = -1, g_failed_src = -1;
void set_failed_pair(int64_t d, int64_t s) {
    g_failed_dst = d;
    g_failed_src = s;
}

void build_regridder(cr_regridder *R, const cr_grid *dst, const cr_grid *src,
                     const cr_build_opts &o, hipStream_t s) {
    const int64_t nd = dst->ncells, ns = src->ncells;
    const int B = 256;
    Timer tm(s);
    Temp T(s);
    tm.mark();  // 0
    // the Planar manifold needs both grids on it (regridder.jl:105-123 errors on a mismatch)
    CR_REQUIRE(dst->planar == src->planar, CR_INVALID_ARGUMENT,
               std::string("Destination and source manifolds must be the same.  Got ") +
                   (dst->planar ? "Planar() and Spherical()" : "Spherical() and Planar()"));
    const bool planar = dst->planar != 0;

    // clip side: a convex side clips; prefer the destination (GO clips p1=src by p2=dst)
    int clip_is_dst;
    if (o.clip_side == CR_CLIP_WITH_DST) clip_is_dst = 1;
    else if (o.clip_side == CR_CLIP_WITH_SRC) clip_is_dst = 0;
    else {
        CR_REQUIRE(dst->convex || src->convex, CR_UNSUPPORTED,
                   "neither grid is declared convex: the Sutherland-Hodgman clip needs a convex side");
        clip_is_dst = dst->convex ? 1 : 0;
    }
    const int clip_nv = clip_is_dst ? dst->nv : src->nv;

    // failure accounting (first-index words start at all ones for atomicMin)
    DevBuf<unsigned long long> acct(kAcctWords, s);
    CR_HIP(hipMemsetAsync(acct.p, 0, acct.bytes(), s));
    CR_HIP(hipMemsetAsync(acct.p + kAcctFirstBadPair, 0xff, sizeof(unsigned long long), s));
    CR_HIP(hipMemsetAsync(acct.p + kAcctFirstBadCell, 0xff, sizeof(unsigned long long), s));
    unsigned long long h_acct[kAcctWords] = {};

    // 1. areas + caps
    DevBuf<double> a_dst, a_src;
    a_dst.alloc_kept(nd, s);
    a_src.alloc_kept(ns, s);
    DevBuf<float4> cap_dst(nd, s), cap_src(ns, s);
    launch_cell_bounds(dst, a_dst.p, cap_dst.p, s);
    launch_cell_bounds(src, a_src.p, cap_src.p, s);
    if (planar) {  // caps of the unit-vector images, areas in the plane
        launch_planar_areas(dst, a_dst.p, s);
        launch_planar_areas(src, a_src.p, s);
    }
    hipLaunchKernelGGL(k_count_cells, dim3(grid_for(nd, B)), dim3(B), 0, s, a_dst.p, nd, int64_t(0), acct.p);
    hipLaunchKernelGGL(k_count_cells, dim3(grid_for(ns, B)), dim3(B), 0, s, a_src.p, ns, nd, acct.p);
    CR_HIP(hipGetLastError());
    tm.mark();  // 1

    // 2. tree over source caps; one sort of both grids' curve keys (source
    // keys first: the destination keys carry the top bit) gives the tree order
    // of the source cells and the query order of the destination cells
    CR_REQUIRE(ns + nd < (int64_t(1) << 31), CR_UNSUPPORTED, "more than 2^31-1 cells in the two grids");
    Levels lv = make_levels(ns);
    int64_t nnodes = lv.off[lv.top] + 1;
    DevBuf<float4> nodes(nnodes, s);
    DevBuf<int32_t> order(ns + nd, s);
    const int32_t *perm = order.p, *qorder = order.p + ns;
    {
        DevBuf<uint32_t> key(ns + nd, s), key_sorted(ns + nd, s);
        DevBuf<int32_t> idx(ns + nd, s);
        hipLaunchKernelGGL(k_sfc_keys, dim3(grid_for(ns, B)), dim3(B), 0, s, cap_src.p, ns, 0u, key.p, idx.p);
        hipLaunchKernelGGL(k_sfc_keys, dim3(grid_for(nd, B)), dim3(B), 0, s, cap_dst.p, nd, 1u << 31, key.p + ns,
                           idx.p + ns);
        sort_pairs(T, key.p, key_sorted.p, idx.p, order.p, ns + nd, 32, s);
        hipLaunchKernelGGL(k_gather_caps, dim3(grid_for(ns, B)), dim3(B), 0, s, cap_src.p, perm, ns,
                           nodes.p);
        for (int L = 1; L <= lv.top; ++L)
            hipLaunchKernelGGL(k_build_level, dim3(grid_for(lv.size[L], B)), dim3(B), 0, s,
                               nodes.p + lv.off[L - 1], lv.size[L - 1], nodes.p + lv.off[L], lv.size[L]);
        CR_HIP(hipGetLastError());
    }
    tm.mark();  // 2

    // 3. candidate search: destination cells in curve order, count, scan, fill
    DevBuf<int64_t> cand_off(nd + 1, s);
    const unsigned qblocks = grid_for(nd, 64);
    DevBuf<int32_t> defer(nd, s);  // cells whose clusters overflow, listed by the count pass
    DevBuf<unsigned> ndefer(1, s);
    CR_HIP(hipMemsetAsync(ndefer.p, 0, sizeof(unsigned), s));
    constexpr unsigned kDeferBlocks = 2048;
    {
        DevBuf<int64_t> cand_cnt(nd + 1, s);
        CR_HIP(hipMemsetAsync(cand_cnt.p + nd, 0, sizeof(int64_t), s));
        hipLaunchKernelGGL(k_search<false>, dim3(qblocks), dim3(64), 0, s, cap_dst.p, qorder, nd, nodes.p,
                           lv, perm, cand_cnt.p, (int32_t *)nullptr, (int32_t *)nullptr, defer.p, ndefer.p);
        hipLaunchKernelGGL(k_search_deferred<false>, dim3(kDeferBlocks), dim3(64), 0, s, cap_dst.p, defer.p,
                           ndefer.p, nodes.p, lv, perm, cand_cnt.p, (int32_t *)nullptr, (int32_t *)nullptr);
        CR_HIP(hipGetLastError());
        exclusive_sum(T, cand_cnt.p, cand_off.p, nd + 1, s);
    }
    int64_t ncand = 0;
    CR_HIP(hipMemcpyAsync(&ncand, cand_off.p + nd, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CR_HIP(hipMemcpyAsync(h_acct, acct.p, sizeof(h_acct), hipMemcpyDeviceToHost, s));
    CR_HIP(hipStreamSynchronize(s));
    if (h_acct[kAcctNonfiniteCells]) {
        const int64_t c = (int64_t)h_acct[kAcctFirstBadCell];
        set_failed_pair(c < nd ? c : -1, c < nd ? -1 : c - nd);
        throw Error{CR_DEGENERATE_GEOMETRY,
                    std::to_string(h_acct[kAcctNonfiniteCells]) + " cell(s) have a non-finite area, the first " +
                        (c < nd ? "destination cell " + std::to_string(c) : "source cell " + std::to_string(c - nd))};
    }
    CR_REQUIRE(ncand < (int64_t(1) << 31), CR_UNSUPPORTED, "more than 2^31 candidate pairs");
    DevBuf<int32_t> cand_src(ncand, s), cand_dst(ncand, s);
    if (ncand) {  // zero-fill: any slot the fill pass skipped is the valid pair (0, 0)
        CR_HIP(hipMemsetAsync(cand_src.p, 0, cand_src.bytes(), s));
        CR_HIP(hipMemsetAsync(cand_dst.p, 0, cand_dst.bytes(), s));
    }
    if (ncand) {
        hipLaunchKernelGGL(k_search<true>, dim3(qblocks), dim3(64), 0, s, cap_dst.p, qorder, nd, nodes.p,
                           lv, perm, cand_off.p, cand_src.p, cand_dst.p, defer.p, ndefer.p);
        hipLaunchKernelGGL(k_search_deferred<true>, dim3(kDeferBlocks), dim3(64), 0, s, cap_dst.p, defer.p,
                           ndefer.p, nodes.p, lv, perm, cand_off.p, cand_src.p, cand_dst.p);
    }
    CR_HIP(hipGetLastError());
    tm.mark();  // 3

    // 4. clip + area per candidate
    DevBuf<double> cval(ncand, s);
    if (ncand && planar) {
        if (std::max(dst->nv, src->nv) <= 8)
            hipLaunchKernelGGL(k_clip_planar<32>, dim3(grid_for(ncand, 64)), dim3(64), 0, s, cand_dst.p, cand_src.p,
                               ncand, dst->xy.p, dst->vidx.p, dst->nv, src->xy.p, src->vidx.p, src->nv, clip_is_dst,
                               a_dst.p, a_src.p, o.threshold_rel, cval.p);
        else
            hipLaunchKernelGGL(k_clip_planar<2 * kMaxCellVertices>, dim3(grid_for(ncand, 64)), dim3(64), 0, s,
                               cand_dst.p, cand_src.p, ncand, dst->xy.p, dst->vidx.p, dst->nv, src->xy.p,
                               src->vidx.p, src->nv, clip_is_dst, a_dst.p, a_src.p, o.threshold_rel, cval.p);
        CR_HIP(hipGetLastError());
    } else if (ncand) {
        if (clip_nv <= 4 && o.threshold_rel > 0.0)
            hipLaunchKernelGGL(k_clip_compact<4>, dim3(grid_for(ncand, 64 * kClipChunks)), dim3(64), 0, s,
                               cand_dst.p, cand_src.p, ncand, dst->cells.p, dst->nv, src->cells.p, src->nv,
                               clip_is_dst, a_dst.p, a_src.p, o.threshold_rel, cval.p);
        else if (clip_nv <= 4)
            hipLaunchKernelGGL(k_clip<4>, dim3(grid_for(ncand, B)), dim3(B), 0, s, cand_dst.p,
                               cand_src.p, ncand, dst->cells.p, dst->nv, src->cells.p, src->nv,
                               clip_is_dst, a_dst.p, a_src.p, o.threshold_rel, cval.p);
        else
            hipLaunchKernelGGL(k_clip_generic, dim3(grid_for(ncand, 64)), dim3(64), 0, s, cand_dst.p,
                               cand_src.p, ncand, dst->cells.p, dst->nv, src->cells.p, src->nv,
                               clip_is_dst, a_dst.p, a_src.p, o.threshold_rel, cval.p);
        CR_HIP(hipGetLastError());
    }
    tm.mark();  // 4

    // 5. compaction (rows stay grouped in candidate order) and the two CSRs
    int64_t nblk = (ncand + kCompactBlock - 1) / kCompactBlock;
    DevBuf<int64_t> blk(nblk + 1, s);
    DevBuf<uint64_t> keep_mask(nblk * (kCompactBlock / 64) + 1, s);
    {
        DevBuf<int64_t> blk_cnt(nblk + 1, s);
        CR_HIP(hipMemsetAsync(blk_cnt.p, 0, blk_cnt.bytes(), s));
        if (ncand)
            hipLaunchKernelGGL(k_keep_count, dim3((unsigned)nblk), dim3(kCompactBlock), 0, s, cval.p,
                               ncand, blk_cnt.p, keep_mask.p, acct.p);
        exclusive_sum(T, blk_cnt.p, blk.p, nblk + 1, s);
    }
    int64_t nnz = 0;
    CR_HIP(hipMemcpyAsync(&nnz, blk.p + nblk, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CR_HIP(hipMemcpyAsync(h_acct, acct.p, sizeof(h_acct), hipMemcpyDeviceToHost, s));
    CR_HIP(hipStreamSynchronize(s));
    if (h_acct[kAcctNonfinitePairs]) {
        // the reference's DefaultIntersectionFailureError (regridder.jl:64-73): name the pair
        const int64_t c = (int64_t)h_acct[kAcctFirstBadPair];
        int32_t pd = -1, ps = -1;
        CR_HIP(hipMemcpyAsync(&pd, cand_dst.p + c, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        CR_HIP(hipMemcpyAsync(&ps, cand_src.p + c, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        CR_HIP(hipStreamSynchronize(s));
        set_failed_pair(pd, ps);
        throw Error{CR_DEGENERATE_GEOMETRY, std::to_string(h_acct[kAcctNonfinitePairs]) +
                                                " intersection(s) have a non-finite area, the first of source cell " +
                                                std::to_string(ps) + " and destination cell " + std::to_string(pd)};
    }

    Csr &F = R->fwd;
    F.nrows = nd; F.ncols = ns; F.nnz = nnz;
    F.rowptr.alloc_kept(nd + 1, s);
    F.col.alloc_kept(nnz, s);
    F.val.alloc_kept(nnz, s);
    {
        // CSR(A): kept pairs in candidate order (columns in search order within a row)
        DevBuf<int32_t> coo_row(nnz, s);
        if (ncand)
            hipLaunchKernelGGL(k_keep_scatter, dim3((unsigned)nblk), dim3(kCompactBlock), 0, s, cval.p,
                               cand_dst.p, cand_src.p, ncand, blk.p, coo_row.p, F.col.p, F.val.p);
        hipLaunchKernelGGL(k_rowptr_from_mask, dim3(grid_for(nd + 1, B)), dim3(B), 0, s, cand_off.p, nd, ncand,
                           blk.p, keep_mask.p, F.rowptr.p);
        CR_HIP(hipGetLastError());
        csr_transpose(F, coo_row.p, R->bwd, T, s);
    }

    // 6. radius^2 scaling (none on the plane), areas
    const double r2 = planar ? 1.0 : o.radius * o.radius;
    R->a_dst = std::move(a_dst);
    R->a_src = std::move(a_src);
    if (r2 != 1.0) {
        hipLaunchKernelGGL(k_scale, dim3(grid_for(nnz, B)), dim3(B), 0, s, F.val.p, nnz, r2, 1.0);
        hipLaunchKernelGGL(k_scale, dim3(grid_for(nnz, B)), dim3(B), 0, s, R->bwd.val.p, nnz, r2, 1.0);
        hipLaunchKernelGGL(k_scale, dim3(grid_for(nd, B)), dim3(B), 0, s, R->a_dst.p, nd, r2, 1.0);
        hipLaunchKernelGGL(k_scale, dim3(grid_for(ns, B)), dim3(B), 0, s, R->a_src.p, ns, r2, 1.0);
        CR_HIP(hipGetLastError());
    }
    R->area_scale = r2;
    finalize_regridder(R, s);
    tm.mark();  // 5
    CR_HIP(hipStreamSynchronize(s));

    auto &st = R->stats;
    st.n_dst = nd;
    st.n_src = ns;
    st.nnz = nnz;
    st.n_candidates = ncand;
    st.ms_bounds = tm.ms(0, 1);
    st.ms_tree = tm.ms(1, 2);
    st.ms_search = tm.ms(2, 3);
    st.ms_clip = tm.ms(3, 4);
    st.ms_assemble = tm.ms(4, 5);
    st.ms_total = tm.ms(0, 5);
    st.clip_with_dst = clip_is_dst;
    st.flipped_cells = dst->flipped + src->flipped;
    st.n_zero_area_cells = (int64_t)h_acct[kAcctZeroCells];
    st.n_zero_cell_pairs = (int64_t)h_acct[kAcctZeroCellPairs];
    st.n_nonfinite_pairs = 0;
    st.n_nonfinite_cells = 0;
}

// normalize!(R): every stored value and both area vectors divided by max(A).
void normalize_regridder(cr_regridder *R, hipStream_t s) {
    Csr &F = R->fwd;
    if (F.nnz == 0) return;
    Temp T(s);
    DevBuf<double> mx(1, s);
    size_t bytes = 0;
    const double lowest = -std::numeric_limits<double>::infinity();
    CR_HIP(rocprim::reduce(nullptr, bytes, F.val.p, mx.p, lowest, (size_t)F.nnz, rocprim::maximum<double>(), s));
    void *p = T.get(bytes);
    CR_HIP(rocprim::reduce(p, bytes, F.val.p, mx.p, lowest, (size_t)F.nnz, rocprim::maximum<double>(), s));
    double m = 0;
    CR_HIP(hipMemcpyAsync(&m, mx.p, sizeof(double), hipMemcpyDeviceToHost, s));
    CR_HIP(hipStreamSynchronize(s));
    CR_REQUIRE(m > 0.0, CR_DEGENERATE_GEOMETRY, "normalize!: maximum(A) is not positive");
    const int B = 256;
    hipLaunchKernelGGL(k_scale, dim3(grid_for(F.nnz, B)), dim3(B), 0, s, F.val.p, F.nnz, 1.0, m);
    hipLaunchKernelGGL(k_scale, dim3(grid_for(R->bwd.nnz, B)), dim3(B), 0, s, R->bwd.val.p, R->bwd.nnz,
                       1.0, m);
    hipLaunchKernelGGL(k_scale, dim3(grid_for(F.nrows, B)), dim3(B), 0, s, R->a_dst.p, F.nrows, 1.0, m);
    hipLaunchKernelGGL(k_scale, dim3(grid_for(F.ncols, B)), dim3(B), 0, s, R->a_src.p, F.ncols, 1.0, m);
    CR_HIP(hipGetLastError());
    R->area_scale /= m;
    drop_tile_plan(F);
    drop_tile_plan(R->bwd);
    finalize_regridder(R, s);
    CR_HIP(hipStreamSynchronize(s));
}

}  // namespace cr

using namespace cr;

extern "C" {

void cr_build_opts_default(cr_build_opts *o) {
    if (!o) return;
    o->threshold_rel = 1e-14;
    o->radius = 6371008.8;
    o->normalize = 0;
    o->clip_side = CR_CLIP_AUTO;
    o->reserved0 = 0;
    o->reserved1 = 0;
}

cr_status cr_regridder_build(cr_regridder **R, const cr_grid *dst, const cr_grid *src,
                             const cr_build_opts *opts, void *stream) {
    return guard([&] {
        CR_REQUIRE(R && dst && src, CR_INVALID_ARGUMENT, "cr_regridder_build: null argument");
        cr_build_opts o;
        if (opts) o = *opts; else cr_build_opts_default(&o);
        CR_REQUIRE(o.threshold_rel >= 0.0, CR_INVALID_ARGUMENT, "threshold_rel must be >= 0");
        CR_REQUIRE(o.radius > 0.0, CR_INVALID_ARGUMENT, "radius must be > 0");
        CR_REQUIRE(o.clip_side >= 0 && o.clip_side <= 2, CR_INVALID_ARGUMENT, "bad clip_side");
        set_failed_pair(-1, -1);
        check_device();
        const auto t0 = std::chrono::steady_clock::now();
        hipStream_t s = as_stream(stream);
        auto *r = new cr_regridder();
        try {
            build_regridder(r, dst, src, o, s);
            for (int side = 0; side < 2; ++side) {
                const cr_grid *g = side ? src : dst;
                DevBuf<int32_t> &f = side ? r->fold_src : r->fold_dst;
                (side ? r->nfold_src : r->nfold_dst) = g->nfold;
                if (g->nfold) {
                    f.alloc(2 * g->nfold);
                    CR_HIP(hipMemcpyAsync(f.p, g->fold.p, f.bytes(), hipMemcpyDeviceToDevice, s));
                }
            }
            if (o.normalize) normalize_regridder(r, s);
            r->stats.ms_host =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        } catch (...) {
            delete r;
            throw;
        }
        *R = r;
    });
}

cr_status cr_last_failed_pair(int64_t *dst_cell, int64_t *src_cell) {
    return guard([&] {
        CR_REQUIRE(dst_cell && src_cell, CR_INVALID_ARGUMENT, "cr_last_failed_pair: null argument");
        *dst_cell = g_failed_dst;
        *src_cell = g_failed_src;
    });
}

cr_status cr_regridder_normalize(cr_regridder *R, void *stream) {
    return guard([&] {
        CR_REQUIRE(R, CR_INVALID_ARGUMENT, "null regridder");
        normalize_regridder(R, as_stream(stream));
    });
}

}  // extern "C"
