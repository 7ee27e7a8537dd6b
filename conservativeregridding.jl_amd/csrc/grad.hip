// grad.hip -- reverse mode of regrid!(y, R, x) w.r.t. the field and the grid
// vertices (no reference counterpart: the reference has no AD; this is what a
// ChainRulesCore rrule around regrid! calls).
//
// With y = (A x) ./ a_to (normalize) and upstream gradient gy:
//   dL/dx          = A^T (gy ./ a_to)                      (the transpose apply)
//   dL/dA_ij       = sum_f gy_if x_jf / a_to,i             (per intersection pair)
//   dL/da_to,i     = -sum_f gy_if y_if / a_to,i            (per destination cell)
// and every A_ij, a_to,i is an area of unit-sphere polygons, so the vertex
// gradient is the sum over pairs of dL/dA_ij * d(area of cell_i ∩ cell_j)/dV
// plus the sum over cells of dL/da_to,i * d(area of cell_i)/dV, scattered into
// per-vertex gradients with FP64 atomics (order-nondeterministic in the last
// bits).  Gradients are tangential: the areas depend on the vertices through
// v/|v|.  Areas here are on the unit sphere; `area_scale` converts.
#include <mutex>
#include <string>

#include "regridder.hpp"
#include "sphgeom.hpp"

namespace cr {

// Row index of every non-zero of CSR(A) (rows = dst).
__global__ void k_pair_rows(const int64_t *__restrict__ rowptr, int64_t nrows, int32_t *__restrict__ rowid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) rowid[k] = (int32_t)i;
}

// Per non-zero of CSR(A) (rows = dst, cols = src), one thread each: the
// coefficient dL/dA_u,ij (unit-sphere areas).
__global__ void k_pair_coef(const int32_t *__restrict__ rowid, const int32_t *__restrict__ col, int64_t nnz,
                            int transpose, int normalize, double kappa,
                            const double *__restrict__ a_to, int64_t nf,
                            const double *__restrict__ x, int64_t xc, int64_t xf,
                            const double *__restrict__ gy, int64_t gc, int64_t gf,
                            double *__restrict__ coef) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int64_t i = rowid[k], j = col[k];
    // forward: y lives on dst rows i, x on src cols j; transpose: the reverse
    const int64_t to = transpose ? j : i, from = transpose ? i : j;
    double s = 0.0;
    for (int64_t f = 0; f < nf; ++f) s += gy[to * gc + f * gf] * x[from * xc + f * xf];
    coef[k] = normalize ? kappa * s / a_to[to] : kappa * s;
}

// Scatter target of the clip backward: caller-vertex gradients of both grids.
// The destination cell's slots may go to a block-local accumulator instead
// (lacc_subj or lacc_clip, 3 doubles per slot), flushed once per block.  NC:
// the clip slots of the kernel (clip index NC wraps to 0).
template <int NC>
struct VertexSink {
    const double *subj, *clip;
    int ns, nc;
    const int32_t *subj_vi, *clip_vi;  // caller vertex of each slot of the two cells
    double *g_subj, *g_clip;           // caller-vertex gradients (may be null)
    double *lacc_subj, *lacc_clip;     // LDS accumulators of one side's slots, or null
    double *own_clip;                  // the pair's own 4 x 3 clip-slot sums in LDS, or null
    double *own_subj;                  // the pair's own 4 x 3 subject-slot sums in LDS, or null
    bool own_atomic = false;           // other lanes add to the own sums too (LDS atomics)

    // slot e of the subject, 0 <= e <= ns (an edge's end vertex e + 1 wraps)
    __device__ int wrap(int e) const { return e >= ns ? e - ns : e; }
    __device__ d3 sv(int e) const { return load3(subj + 3 * wrap(e)); }
    __device__ d3 cv(int k) const {
        const int kk = k >= NC ? k - NC : k;
        return load3(clip + 3 * (kk < nc ? kk : nc - 1));
    }
    __device__ static void add(double *g, int32_t v, const d3 &d) {
        if (!g) return;
        atomicAdd(g + 3 * (int64_t)v, d.x);
        atomicAdd(g + 3 * (int64_t)v + 1, d.y);
        atomicAdd(g + 3 * (int64_t)v + 2, d.z);
    }
    __device__ static void add_local(double *l, int slot, const d3 &d) {
        atomicAdd(l + 3 * slot, d.x);
        atomicAdd(l + 3 * slot + 1, d.y);
        atomicAdd(l + 3 * slot + 2, d.z);
    }
    __device__ void add_subj(int e, const d3 &g) const {
        e = wrap(e);
        if (lacc_subj) {
            add_local(lacc_subj, e, tangential(g, sv(e)));
        } else if (own_subj) {  // this pair's own sums
            const d3 t = tangential(g, sv(e));
            if (own_atomic) {
                add_local(own_subj, e, t);
            } else {
                own_subj[3 * e] += t.x;
                own_subj[3 * e + 1] += t.y;
                own_subj[3 * e + 2] += t.z;
            }
        } else {
            add(g_subj, subj_vi[e], tangential(g, sv(e)));
        }
    }
    __device__ void add_clip(int k, const d3 &g) const {
        int kk = k >= NC ? k - NC : k;
        kk = kk < nc ? kk : nc - 1;
        const d3 t = tangential(g, cv(kk));
        if (lacc_clip) {
            add_local(lacc_clip, kk, t);
        } else if (own_clip) {  // this pair's own sums
            if (own_atomic) {
                add_local(own_clip, kk, t);
            } else {
                own_clip[3 * kk] += t.x;
                own_clip[3 * kk + 1] += t.y;
                own_clip[3 * kk + 2] += t.z;
            }
        } else {
            add(g_clip, clip_vi[kk], t);
        }
    }
    __device__ void operator()(int32_t l, const d3 &x, const d3 &g) const {
        if (l < 0) return;
        int32_t t = lab_type(l), a = lab_hi(l), b = lab_lo(l);
        if (t == kLabS) {
            add_subj(l & kLabIndexMask, g);  // the subject index is the low 30 bits
        } else if (t == kLabC) {
            add_clip(b, g);
        } else {
            d3 p1, q1, p2 = cv(b), q2 = cv(b + 1), gp1, gq1, gp2, gq2;
            if (t == kLabXF) { p1 = sv(a); q1 = sv(a + 1); }
            else { p1 = cv(a); q1 = cv(a + 1); }
            if (!crossing_backward(x, g, p1, q1, p2, q2, gp1, gq1, gp2, gq2)) return;
            if (t == kLabXF) { add_subj(a, gp1); add_subj(a + 1, gq1); }
            else { add_clip(a, gp1); add_clip(a + 1, gq1); }
            add_clip(b, gp2);
            add_clip(b + 1, gq2);
        }
    }
};

// One wave per 64 consecutive non-zeros.  CSR rows are destination cells, so a
// wave's pairs share few destination cells (about 12 on the benchmark): their
// slots' gradients are summed in LDS (rows r0 .. r0 + kGradRows - 1 of the
// wave) and reach memory once per wave.  A source cell of at most 4 vertices
// (clipping or clipped) has its slots summed per pair and written once per
// pair; other contributions use global FP64 atomics directly.
constexpr int kGradRows = 16;

// The sink of one pair: its two rings, its slots' accumulators (the wave's
// destination rows in LDS when `lacc` is given and the row is one of them,
// the pair's own source-cell sums in `own`) and the caller-vertex gradients;
// NC: the clip slots of the kernel.
template <int NC>
__device__ __forceinline__ VertexSink<NC> pair_sink(int64_t d, int64_t s, int64_t r0, double *lacc, double *own,
                                                const double *__restrict__ dst_cells, int dst_nv,
                                                const int32_t *__restrict__ dst_vi,
                                                const double *__restrict__ src_cells, int src_nv,
                                                const int32_t *__restrict__ src_vi, int clip_is_dst,
                                                double *__restrict__ g_dst, double *__restrict__ g_src) {
    const double *dp = dst_cells + 3 * (int64_t)dst_nv * d;
    const double *sp = src_cells + 3 * (int64_t)src_nv * s;
    double *mine = (lacc && g_dst && d - r0 < kGradRows) ? lacc + (d - r0) * dst_nv * 3 : nullptr;
    // a source cell of at most 4 vertices sums its slots per pair
    double *own4 = (g_src && src_nv <= 4) ? own : nullptr;
    return clip_is_dst
                       ? VertexSink<NC>{sp, dp, src_nv, dst_nv, src_vi + (int64_t)src_nv * s, dst_vi + (int64_t)dst_nv * d,
                                    g_src, g_dst, nullptr, mine, nullptr, own4}
                       : VertexSink<NC>{dp, sp, dst_nv, src_nv, dst_vi + (int64_t)dst_nv * d, src_vi + (int64_t)src_nv * s,
                                    g_dst, g_src, mine, nullptr, own4, nullptr};
}

// The per-pair sums of the source cell's 4 slots, one atomic per component.
template <int NC>
__device__ __forceinline__ void flush_own(const VertexSink<NC> &sink, int src_nv, double *own, double *g_src) {
    if (!sink.own_clip && !sink.own_subj) return;
    const int32_t *vi = sink.own_clip ? sink.clip_vi : sink.subj_vi;
    for (int kk = 0; kk < src_nv && kk < 4; ++kk) {
        const d3 v = mk3(own[3 * kk], own[3 * kk + 1], own[3 * kk + 2]);
        if (v.x != 0.0 || v.y != 0.0 || v.z != 0.0) VertexSink<NC>::add(g_src, vi[kk], v);
    }
}

// The wave's destination rows from LDS to memory.
__device__ __forceinline__ void flush_rows(const double *lacc, int64_t r0, int64_t nrows, int dst_nv,
                                           const int32_t *__restrict__ dst_vi, double *g_dst) {
    for (int i = threadIdx.x; i < nrows * dst_nv * 3; i += 64) {
        const double v = lacc[i];
        if (v == 0.0) continue;
        const int r = i / (dst_nv * 3), slot = (i / 3) % dst_nv, comp = i % 3;
        atomicAdd(g_dst + 3 * (int64_t)dst_vi[(r0 + r) * dst_nv + slot] + comp, v);
    }
}

// Crossing work of the boundary decomposition, queued per wave in LDS: a
// crossing's gradient (tangential at it) with its label and its pair's lane,
// done 64 at a time by the whole wave at the points where every lane's pass 2
// is in step.  A flush leaves fewer than 64; a lane adds at most 2 between
// flushes.
constexpr int kXQueue = 64 + 2 * 64;

struct XQueue {
    int32_t lab[kXQueue], lane[kXQueue];
    double g[3 * kXQueue];
    int32_t pd[64], ps[64];  // each lane's pair (destination, source cell)
    int n;
};

template <int NC>
struct QueueEmit {
    VertexSink<NC> *sink;
    XQueue *q;
    bool live;
    // the decomposition emits subject vertices, clip vertices and
    // subject x clip crossings (no clip x clip crossings)
    __device__ __forceinline__ void operator()(int32_t l, const d3 &x, const d3 &g) {
        if (!live) return;
        const int32_t t = lab_type(l);
        if (t == kLabXF) {
            const d3 gt = tangential(g, x);
            const int e = atomicAdd(&q->n, 1);
            q->lab[e] = l;
            q->lane[e] = threadIdx.x;
            q->g[3 * e] = gt.x;
            q->g[3 * e + 1] = gt.y;
            q->g[3 * e + 2] = gt.z;
        } else if (t == kLabS) {
            sink->add_subj(l & kLabIndexMask, g);
        } else {
            sink->add_clip(lab_lo(l), g);
        }
    }
};

// The vertex pass by boundary decomposition (edges_plan + edges_emit), one
// wave per 64 consecutive non-zeros: flat loops over the subject edges and
// the clip edges (no clip stages), every lane through the same iterations
// (a declined or empty lane emits nothing), crossings through the queue.  A
// pair the decomposition declines (near a degenerate configuration, or a
// degenerate clip edge) is appended to `list` (wave by wave, in order) for
// k_clip_grad_list.
template <int NC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_clip_grad_edges(
    const int32_t *__restrict__ rowid, const int32_t *__restrict__ col, const double *__restrict__ coef,
    int64_t nnz, const double *__restrict__ dst_cells, int dst_nv, const int32_t *__restrict__ dst_vi,
    const double *__restrict__ src_cells, int src_nv, const int32_t *__restrict__ src_vi, int clip_is_dst,
    double *__restrict__ g_dst, double *__restrict__ g_src, int32_t *__restrict__ list,
    uint32_t *__restrict__ count) {
    extern __shared__ double lacc[];  // kGradRows * dst_nv * 3
    __shared__ double own[64 * 12];
    __shared__ XQueue q;
    const int lane = threadIdx.x;
    const int64_t k0 = (int64_t)blockIdx.x * 64, k = k0 + lane;
    const int64_t r0 = rowid[k0];
    const int nacc = kGradRows * dst_nv * 3;
    for (int i = lane; i < nacc; i += 64) lacc[i] = 0.0;
    double *my_own = own + 12 * lane;
#pragma unroll
    for (int i = 0; i < 12; ++i) my_own[i] = 0.0;
    const bool real = k < nnz;
    const int64_t kk = real ? k : k0;  // lanes past the end run the loops on a valid pair
    const int32_t d = rowid[kk], s = col[kk];
    q.pd[lane] = d;
    q.ps[lane] = s;
    if (lane == 0) q.n = 0;
    __syncthreads();
    const double c = real ? coef[kk] : 0.0;
    VertexSink<NC> sink = pair_sink<NC>(d, s, r0, lacc, my_own, dst_cells, dst_nv, dst_vi, src_cells, src_nv, src_vi,
                                clip_is_dst, g_dst, g_src);
    EdgesPlan<NC> P;
    const bool planned = edges_plan<NC>(sink.clip, sink.nc, sink.subj, sink.ns, P);
    const bool live = c != 0.0 && planned;
    // declined pairs to the list: one atomic per wave, lanes in order
    {
        const bool fb = c != 0.0 && !planned;
        const uint64_t m = __ballot(fb);
        if (m) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(count, (uint32_t)__popcll(m));
            base = __shfl(base, 0);
            if (fb) list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)k;
        }
    }
    auto work = [&](int e) {
        const int L = q.lane[e];
        const int32_t l = q.lab[e];
        const d3 gt = mk3(q.g[3 * e], q.g[3 * e + 1], q.g[3 * e + 2]);
        VertexSink<NC> sk = pair_sink<NC>(q.pd[L], q.ps[L], r0, lacc, own + 12 * L, dst_cells, dst_nv, dst_vi, src_cells,
                                  src_nv, src_vi, clip_is_dst, g_dst, g_src);
        sk.own_atomic = true;
        const int a = lab_hi(l), b = lab_lo(l);
        d3 gp1, gq1, gp2, gq2;
        if (crossing_backward_arc(gt, sk.sv(a), sk.sv(a + 1), sk.cv(b), sk.cv(b + 1), gp1, gq1, gp2, gq2)) {
            sk.add_subj(a, gp1);
            sk.add_subj(a + 1, gq1);
            sk.add_clip(b, gp2);
            sk.add_clip(b + 1, gq2);
        }
    };
    auto flush = [&](bool drain) {
        __syncthreads();
        int n = q.n;
        while (n >= 64) {
            work(n - 64 + lane);
            n -= 64;
        }
        if (drain && lane < n) work(lane);
        __syncthreads();
        if (lane == 0) q.n = drain ? 0 : n;
        __syncthreads();
    };
    QueueEmit<NC> em{&sink, &q, live};
    edges_emit<NC>(P, sink.subj, sink.ns, c, em, [&] { flush(false); });
    flush(true);
    if (live) flush_own(sink, src_nv, my_own, g_src);
    __syncthreads();
    if (!g_dst) return;
    const int64_t klast = min(nnz, k0 + 64) - 1;
    flush_rows(lacc, r0, min((int64_t)kGradRows, (int64_t)rowid[klast] - r0 + 1), dst_nv, dst_vi, g_dst);
}

// The vertex pass through the clip pipeline (clip_area_backward) for the
// pairs the boundary decomposition declines: list[0 .. *count), 64 per wave,
// the waves striding over the list.  The list holds each wave's declined
// pairs in order, so a window of 64 entries mostly shares rows: rows r0 ..
// r0 + kGradRows - 1 of its first entry are summed in LDS, the rest go to
// memory directly.  Two waves per SIMD (256 registers): the second wave hides
// the atomics' and the FP64 chain's latency.  The clip pipeline's edge
// normals live in LDS (6 KB per workgroup), which cuts the register spills
// from 73 to about 30; moving its first vertices and clip vertices there too
// removes the rest but lowers the occupancy LDS allows and ran slower (as the
// whole vertex pass on C3: 8.4 ms with the normals in registers, 7.1 ms as
// here, 9.1 ms with none spilled), and so did one wave per SIMD with no
// spills (8.3 ms against 5.4).
template <int NC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_clip_grad_list(
    const int32_t *__restrict__ rowid, const int32_t *__restrict__ col, const double *__restrict__ coef,
    const double *__restrict__ dst_cells, int dst_nv, const int32_t *__restrict__ dst_vi,
    const double *__restrict__ src_cells, int src_nv, const int32_t *__restrict__ src_vi, int clip_is_dst,
    double *__restrict__ g_dst, double *__restrict__ g_src, const int32_t *__restrict__ list,
    const uint32_t *__restrict__ count, int64_t n_rows) {
    extern __shared__ double lacc[];  // kGradRows * dst_nv * 3
    __shared__ double own[64 * 12];
    __shared__ d3 nrm_lds[64 * NC];  // each thread's NC clip-edge normals
    double *my_own = own + 12 * threadIdx.x;
    const int64_t total = *count;
    const int nacc = kGradRows * dst_nv * 3;
    for (int64_t base = (int64_t)blockIdx.x * 64; base < total; base += (int64_t)gridDim.x * 64) {
        const int64_t r0 = rowid[list[base]];
        for (int i = threadIdx.x; i < nacc; i += 64) lacc[i] = 0.0;
#pragma unroll
        for (int i = 0; i < 12; ++i) my_own[i] = 0.0;
        __syncthreads();
        const int64_t i = base + threadIdx.x;
        if (i < total) {
            const int64_t k = list[i];
            const double c = coef[k];
            const int64_t d = rowid[k];
            if (c != 0.0) {
                VertexSink<NC> sink = pair_sink<NC>(d, col[k], r0, d >= r0 ? lacc : nullptr, my_own, dst_cells, dst_nv,
                                            dst_vi, src_cells, src_nv, src_vi, clip_is_dst, g_dst, g_src);
                clip_area_backward<NC, VertexSink<NC>>(sink.clip, sink.nc, sink.subj, sink.ns, c, sink,
                                                   nrm_lds + NC * threadIdx.x);
                flush_own(sink, src_nv, my_own, g_src);
            }
        }
        __syncthreads();
        if (g_dst) flush_rows(lacc, r0, min((int64_t)kGradRows, n_rows - r0), dst_nv, dst_vi, g_dst);
        __syncthreads();
    }
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

// Workgroups of the list pass (the waves stride over the declined pairs).
constexpr unsigned kListBlocks = 2048;

// d(area of each destination cell)/dV times dL/da_u.  private_v: no vertex
// index is shared by two cells, so each thread owns its cell's entries of g
// and adds without atomics (the kernels before it on the stream have finished).
__global__ void k_area_grad(const double *__restrict__ cells, int nv, const int32_t *__restrict__ vi,
                            int64_t ncells, double kappa, const double *__restrict__ a_to, int64_t nf,
                            const double *__restrict__ y, int64_t yc, int64_t yf,
                            const double *__restrict__ gy, int64_t gc, int64_t gf,
                            double *__restrict__ g, int private_v) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells || a_to[t] == 0.0) return;  // zero-area (ghost) cells have no area term
    double s = 0.0;
    for (int64_t f = 0; f < nf; ++f) s += gy[t * gc + f * gf] * y[t * yc + f * yf];
    double c = -kappa * s / a_to[t];
    if (c == 0.0) return;
    const int32_t *v = vi + (int64_t)nv * t;
    ring_area_backward(cells + 3 * (int64_t)nv * t, nv, c, [&](int slot, const d3 &d) {
        double *gv = g + 3 * (int64_t)v[slot];
        if (private_v) {
            gv[0] += d.x;
            gv[1] += d.y;
            gv[2] += d.z;
        } else {
            atomicAdd(gv, d.x);
            atomicAdd(gv + 1, d.y);
            atomicAdd(gv + 2, d.z);
        }
    });
}

// The same for grids whose cells own their vertices (private_v) with at most
// kStagedNv vertices per cell: a workgroup's 64 cells are one contiguous run of
// vertex data and of g (each cell's caller vertices are its own slots, in some
// order), so both are staged through LDS with coalesced loads and stores
// instead of a thread per cell reading and writing 24*nv bytes apart from its
// neighbours.
constexpr int kStagedNv = 16;
__global__ void __launch_bounds__(64) k_area_grad_private(const double *__restrict__ cells, int nv,
                                                          const int32_t *__restrict__ vi, int64_t ncells,
                                                          double kappa, const double *__restrict__ a_to, int64_t nf,
                                                          const double *__restrict__ y, int64_t yc, int64_t yf,
                                                          const double *__restrict__ gy, int64_t gc, int64_t gf,
                                                          double *__restrict__ g) {
    extern __shared__ double stage[];  // 2 * 64 * nv * 3 doubles: vertices, then g
    double *vin = stage, *gout = stage + 64 * 3 * nv;
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int ncell = (int)min((int64_t)64, ncells - c0);
    const int n = ncell * nv * 3;
    const double *src = cells + 3 * (int64_t)nv * c0;
    double *gb = g + 3 * (int64_t)nv * c0;
    for (int i = threadIdx.x; i < n; i += 64) {
        vin[i] = src[i];
        gout[i] = gb[i];
    }
    __syncthreads();
    const int t = threadIdx.x;
    const int64_t c = c0 + t;
    if (t < ncell && a_to[c] != 0.0) {  // zero-area (ghost) cells have no area term
        double sum = 0.0;
        for (int64_t f = 0; f < nf; ++f) sum += gy[c * gc + f * gf] * y[c * yc + f * yf];
        const double coef = -kappa * sum / a_to[c];
        if (coef != 0.0) {
            const int32_t *v = vi + (int64_t)nv * c;
            const int64_t base = (int64_t)nv * c0;
            ring_area_backward(vin + 3 * nv * t, nv, coef, [&](int slot, const d3 &d) {
                double *o = gout + 3 * (v[slot] - base);
                o[0] += d.x;
                o[1] += d.y;
                o[2] += d.z;
            });
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 64) gb[i] = gout[i];
}

// The same for tables of NV-vertex cells whose slot -> vertex map is the
// identity (cells own their vertices and none was re-oriented): two lanes per
// cell, each taking half of its fan (triangles (v0, v_i, v_{i+1}) below and
// from the middle vertex v_H on) with its slots' gradients in registers (the
// fan's slots are compile-time indices); the two shares of v0 and v_H meet by
// a lane shuffle.  One LDS buffer stages the block's 64 cells' vertices in
// and their gradients out, both with coalesced accesses (two lanes a cell:
// twice the waves one lane a cell had for the same LDS).
template <int NV>
__global__ void __launch_bounds__(128) k_area_grad_table(const double *__restrict__ cells, int64_t ncells,
                                                         double kappa, const double *__restrict__ a_to, int64_t nf,
                                                         const double *__restrict__ y, int64_t yc, int64_t yf,
                                                         const double *__restrict__ gy, int64_t gc, int64_t gf,
                                                         double *__restrict__ g) {
    constexpr int H = NV / 2;
    static_assert(NV >= 4 && NV % 2 == 0, "even rings of at least 4 vertices");
    __shared__ double stage[64 * NV * 3];
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int ncell = (int)min((int64_t)64, ncells - c0);
    const int n = ncell * NV * 3;
    const double *src = cells + 3 * (int64_t)NV * c0;
    double *gb = g + 3 * (int64_t)NV * c0;
    for (int i = threadIdx.x; i < n; i += 128) stage[i] = src[i];
    __syncthreads();
    const int t = threadIdx.x >> 1;  // the cell; lanes 2t and 2t + 1 share it (one wave)
    const bool hi = (threadIdx.x & 1) != 0;
    const int64_t c = c0 + t;
    d3 ga = mk3(0.0, 0.0, 0.0);
    d3 gr[H];  // the lane's slots: 1 .. H (low lane) or H .. NV-1 (high lane)
#pragma unroll
    for (int k = 0; k < H; ++k) gr[k] = mk3(0.0, 0.0, 0.0);
    if (t < ncell && a_to[c] != 0.0) {  // zero-area (ghost) cells have no area term
        double sum = 0.0;
        for (int64_t f = 0; f < nf; ++f) sum += gy[c * gc + f * gf] * y[c * yc + f * yf];
        const double coef = -kappa * sum / a_to[c];
        if (coef != 0.0) {
            // the low lane: triangles 1 .. H-1 (slots 1 .. H), the high lane:
            // triangles H .. NV-2 (slots H .. NV-1)
            if (!hi) ring_area_backward_part<1, H>(stage + 3 * NV * t, coef, ga, gr);
            else ring_area_backward_part<H, NV - 1>(stage + 3 * NV * t, coef, ga, gr);
        }
    }
    // the shared slots: v0 from both lanes, v_H (the low lane's last, the high lane's first)
    const d3 ga_o = mk3(__shfl_xor(ga.x, 1), __shfl_xor(ga.y, 1), __shfl_xor(ga.z, 1));
    const d3 mid = hi ? gr[0] : gr[H - 1];
    const d3 mid_o = mk3(__shfl_xor(mid.x, 1), __shfl_xor(mid.y, 1), __shfl_xor(mid.z, 1));
    __syncthreads();
    if (t < ncell) {
        double *o = stage + 3 * NV * t;
        auto put = [&](int k, const d3 &v) {
            o[3 * k] = v.x;
            o[3 * k + 1] = v.y;
            o[3 * k + 2] = v.z;
        };
        if (!hi) {
            const double *a = o;  // v0, still staged: the low lane writes it only now
            put(0, tangential(ga + ga_o, load3(a)));
#pragma unroll
            for (int k = 1; k < H; ++k) put(k, gr[k - 1]);
            put(H, gr[H - 1] + mid_o);
        } else {
#pragma unroll
            for (int k = H + 1; k < NV; ++k) put(k, gr[k - H]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 128) gb[i] += stage[i];
}

__global__ void k_div_rows(const double *__restrict__ gy, int64_t gc, int64_t gf, int64_t n, int64_t nf,
                           const double *__restrict__ a, double *__restrict__ w) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nf) return;
    int64_t f = t / n, i = t - f * n;
    w[t] = a ? gy[i * gc + f * gf] / a[i] : gy[i * gc + f * gf];
}

void apply(const cr_regridder *R, bool transpose, int dtype, int64_t nf, const void *src, int64_t sc,
           int64_t sf, void *dst, int64_t dc, int64_t df, bool normalize, hipStream_t s,
           int force_variant);
bool apply_ready(const cr_regridder *R, bool transpose, int dtype, int64_t nf, int64_t sf, int64_t df);

// Scratch of the reverse mode: w (n_to x nfields doubles, the field adjoint's
// input), coef (a double per non-zero), rowid (an int32 per non-zero), list
// (the non-zeros the boundary decomposition declines, at most one int32 per
// non-zero) and its count.
struct BackwardScratch {
    double *w = nullptr, *coef = nullptr;
    int32_t *rowid = nullptr, *list = nullptr;
    uint32_t *count = nullptr;
};

// The reverse mode on the given stream.  ws == nullptr: the scratch is
// allocated for the call (stream-ordered) and the stream synchronised at the
// end; otherwise the caller's scratch is used, nothing is allocated, every
// per-direction plan must exist already, and the call returns without
// synchronising.
static void backward(const cr_regridder *R, const cr_grid *dst, const cr_grid *src, int32_t transpose,
                     int32_t normalize, int64_t nfields, const double *x, int64_t xc, int64_t xf, const double *y,
                     int64_t yc, int64_t yf, const double *gy, int64_t gc, int64_t gf, double *gx, int64_t gxc,
                     int64_t gxf, double *g_dst_verts, double *g_src_verts, hipStream_t s,
                     const BackwardScratch *ws) {
    CR_REQUIRE(R && nfields > 0 && x && gy, CR_INVALID_ARGUMENT, "cr_regrid_backward: bad arguments");
    CR_REQUIRE(xc > 0 && gc > 0 && (nfields == 1 || (xf > 0 && gf > 0)), CR_INVALID_ARGUMENT,
               "cr_regrid_backward: bad strides");
    const bool want_verts = g_dst_verts || g_src_verts;
    const Csr &F = R->fwd;
    const int64_t n_to = transpose ? F.ncols : F.nrows;
    const double *a_to = transpose ? R->a_src.p : R->a_dst.p;
    const int B = 256;
    // field gradient: gx = A_dir^T (gy ./ a_to)
    if (gx) {
        CR_REQUIRE(gxc > 0 && (nfields == 1 || gxf > 0), CR_INVALID_ARGUMENT, "bad gx strides");
        DevBuf<double> wb;
        double *w = ws ? ws->w : nullptr;
        if (!ws) {
            wb.alloc_async((size_t)n_to * nfields, s);
            w = wb.p;
        } else {
            CR_REQUIRE(apply_ready(R, !transpose, CR_F64, nfields, n_to, gxf), CR_UNSUPPORTED,
                       "the field adjoint's plan is not made: call cr_regrid_prepare(R, !transpose, CR_F64, "
                       "nfields, n_to, field stride of gx) first");
        }
        hipLaunchKernelGGL(k_div_rows, dim3(grid_for(n_to * nfields, B)), dim3(B), 0, s, gy, gc, gf, n_to,
                           nfields, normalize ? a_to : nullptr, w);
        CR_HIP(hipGetLastError());
        apply(R, !transpose, CR_F64, nfields, w, 1, n_to, gx, gxc, gxf, false, s, -1);
    }
    if (!want_verts) {
        if (!ws) CR_HIP(hipStreamSynchronize(s));
        return;
    }
    CR_REQUIRE(R->area_scale > 0.0, CR_UNSUPPORTED,
               "vertex gradients need a regridder built from grids (cr_regridder_build)");
    CR_REQUIRE(dst && src && dst->ncells == F.nrows && src->ncells == F.ncols, CR_DIMENSION_MISMATCH,
               "cr_regrid_backward: grids do not match the regridder");
    CR_REQUIRE(!dst->planar && !src->planar, CR_UNSUPPORTED,
               "vertex gradients are of spherical intersection areas; these grids are planar");
    const int clip_nv = R->stats.clip_with_dst ? dst->nv : src->nv;
    CR_REQUIRE(clip_nv <= 8, CR_UNSUPPORTED, "vertex gradients need clip cells with at most 8 vertices");
    if (normalize) CR_REQUIRE(y, CR_INVALID_ARGUMENT, "cr_regrid_backward: the forward output y is needed");
    const double kappa = R->area_scale;
    DevBuf<double> coef_b;
    DevBuf<int32_t> rowid_b;
    double *coef = ws ? ws->coef : nullptr;
    int32_t *rowid = ws ? ws->rowid : nullptr;
    if (!ws) {
        coef_b.alloc_async(F.nnz, s);
        rowid_b.alloc_async(F.nnz, s);
        coef = coef_b.p;
        rowid = rowid_b.p;
    }
    hipLaunchKernelGGL(k_pair_rows, dim3(grid_for(F.nrows, B)), dim3(B), 0, s, F.rowptr.p, F.nrows, rowid);
    if (F.nnz)
        hipLaunchKernelGGL(k_pair_coef, dim3(grid_for(F.nnz, B)), dim3(B), 0, s, rowid, F.col.p, F.nnz,
                           (int)(transpose != 0), (int)(normalize != 0), kappa, a_to, nfields, x, xc, xf, gy, gc,
                           gf, coef);
    if (F.nnz) {
        // most pairs by boundary decomposition; the pairs it declines through the clip pipeline
        DevBuf<int32_t> list_b;
        DevBuf<uint32_t> count_b;
        int32_t *list = ws ? ws->list : nullptr;
        uint32_t *count = ws ? ws->count : nullptr;
        if (!ws) {
            list_b.alloc_async(F.nnz, s);
            count_b.alloc_async(1, s);
            list = list_b.p;
            count = count_b.p;
        }
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, count);  // a kernel: no memset node in a graph
        const size_t lds = (size_t)kGradRows * dst->nv * 3 * sizeof(double);
        if (clip_nv <= 4) {
            hipLaunchKernelGGL(k_clip_grad_edges<4>, dim3(grid_for(F.nnz, 64)), dim3(64), lds, s, rowid, F.col.p,
                               coef, F.nnz, dst->cells.p, dst->nv, dst->vidx.p, src->cells.p, src->nv, src->vidx.p,
                               R->stats.clip_with_dst, g_dst_verts, g_src_verts, list, count);
            hipLaunchKernelGGL(k_clip_grad_list<4>, dim3(kListBlocks), dim3(64), lds, s, rowid, F.col.p, coef,
                               dst->cells.p, dst->nv, dst->vidx.p, src->cells.p, src->nv, src->vidx.p,
                               R->stats.clip_with_dst, g_dst_verts, g_src_verts, (const int32_t *)list,
                               (const uint32_t *)count, F.nrows);
        } else {
            hipLaunchKernelGGL(k_clip_grad_edges<8>, dim3(grid_for(F.nnz, 64)), dim3(64), lds, s, rowid, F.col.p,
                               coef, F.nnz, dst->cells.p, dst->nv, dst->vidx.p, src->cells.p, src->nv, src->vidx.p,
                               R->stats.clip_with_dst, g_dst_verts, g_src_verts, list, count);
            hipLaunchKernelGGL(k_clip_grad_list<8>, dim3(kListBlocks), dim3(64), lds, s, rowid, F.col.p, coef,
                               dst->cells.p, dst->nv, dst->vidx.p, src->cells.p, src->nv, src->vidx.p,
                               R->stats.clip_with_dst, g_dst_verts, g_src_verts, (const int32_t *)list,
                               (const uint32_t *)count, F.nrows);
        }
        CR_HIP(hipGetLastError());
    }
    double *g_to = transpose ? g_src_verts : g_dst_verts;
    const cr_grid *gto = transpose ? src : dst;
    if (normalize && g_to) {
        const bool table = gto->private_vertices && gto->flipped == 0;  // slot -> vertex is the identity
        if (table && gto->nv == 4)
            hipLaunchKernelGGL(k_area_grad_table<4>, dim3(grid_for(n_to, 64)), dim3(128), 0, s, gto->cells.p, n_to,
                               kappa, a_to, nfields, y, yc, yf, gy, gc, gf, g_to);
        else if (table && gto->nv == 8)
            hipLaunchKernelGGL(k_area_grad_table<8>, dim3(grid_for(n_to, 64)), dim3(128), 0, s, gto->cells.p, n_to,
                               kappa, a_to, nfields, y, yc, yf, gy, gc, gf, g_to);
        else if (table && gto->nv == 16)
            hipLaunchKernelGGL(k_area_grad_table<16>, dim3(grid_for(n_to, 64)), dim3(128), 0, s, gto->cells.p,
                               n_to, kappa, a_to, nfields, y, yc, yf, gy, gc, gf, g_to);
        else if (gto->private_vertices && gto->nv <= kStagedNv)
            hipLaunchKernelGGL(k_area_grad_private, dim3(grid_for(n_to, 64)), dim3(64),
                               2 * 64 * 3 * gto->nv * sizeof(double), s, gto->cells.p, gto->nv,
                               gto->vidx.p, n_to, kappa, a_to, nfields, y, yc, yf, gy, gc, gf, g_to);
        else
            hipLaunchKernelGGL(k_area_grad, dim3(grid_for(n_to, B)), dim3(B), 0, s, gto->cells.p, gto->nv,
                               gto->vidx.p, n_to, kappa, a_to, nfields, y, yc, yf, gy, gc, gf, g_to,
                               gto->private_vertices);
    }
    CR_HIP(hipGetLastError());
    if (!ws) CR_HIP(hipStreamSynchronize(s));
}

// The allocation-free reverse mode's scratch: one static array per device
// (every GPU the library runs on keeps its copy until the process ends).  Host
// code must use the per-device address from hipGetSymbolAddress, never the
// array's name (that is a host-side stand-in, not a device address).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static unsigned char *scratch_address() {
    static std::mutex mu;
    static unsigned char *ptr[256] = {};
    int dev = 0;
    CR_HIP(hipGetDevice(&dev));
    CR_REQUIRE(dev >= 0 && dev < 256, CR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        CR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CR_REQUIRE(p, CR_INTERNAL, "hipGetSymbolAddress gave no address for the scratch array");
        ptr[dev] = (unsigned char *)p;
    }
    return ptr[dev];
}

}  // namespace cr

using namespace cr;

extern "C" {

cr_status cr_regrid_backward_noalloc(const cr_regridder *R, const cr_grid *dst, const cr_grid *src,
                                     int32_t transpose, int32_t normalize, int64_t nfields, const double *x,
                                     int64_t xc, int64_t xf, const double *y, int64_t yc, int64_t yf,
                                     const double *gy, int64_t gc, int64_t gf, double *gx, int64_t gxc,
                                     int64_t gxf, double *g_dst_verts, double *g_src_verts, void *stream) {
    return guard([&] {
        CR_REQUIRE(R && nfields > 0, CR_INVALID_ARGUMENT, "cr_regrid_backward_noalloc: bad arguments");
        const Csr &F = R->fwd;
        const int64_t n_to = transpose ? F.ncols : F.nrows;
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t wbytes = gx ? up((size_t)n_to * (size_t)nfields * sizeof(double)) : 0;
        const bool verts = g_dst_verts || g_src_verts;
        const size_t cbytes = verts ? up((size_t)F.nnz * sizeof(double)) : 0;
        const size_t rbytes = verts ? up((size_t)F.nnz * sizeof(int32_t)) : 0;
        const size_t lbytes = verts ? up((size_t)F.nnz * sizeof(int32_t)) + 256 : 0;  // list + count
        const size_t need = wbytes + cbytes + rbytes + lbytes;
        CR_REQUIRE(need <= kNoAllocScratchBytes, CR_INVALID_ARGUMENT,
                   "cr_regrid_backward_noalloc: the call needs " + std::to_string(need) +
                       " bytes of scratch, over its maximum of " + std::to_string(kNoAllocScratchBytes) +
                       " (use cr_regrid_backward)");
        unsigned char *base = scratch_address();
        BackwardScratch ws;
        ws.w = wbytes ? (double *)base : nullptr;
        ws.coef = cbytes ? (double *)(base + wbytes) : nullptr;
        ws.rowid = rbytes ? (int32_t *)(base + wbytes + cbytes) : nullptr;
        ws.count = lbytes ? (uint32_t *)(base + wbytes + cbytes + rbytes) : nullptr;
        ws.list = lbytes ? (int32_t *)(base + wbytes + cbytes + rbytes + 256) : nullptr;
        backward(R, dst, src, transpose, normalize, nfields, x, xc, xf, y, yc, yf, gy, gc, gf, gx, gxc, gxf,
                 g_dst_verts, g_src_verts, as_stream(stream), &ws);
    });
}

cr_status cr_regrid_backward(const cr_regridder *R, const cr_grid *dst, const cr_grid *src,
                             int32_t transpose, int32_t normalize, int64_t nfields, const double *x,
                             int64_t xc, int64_t xf, const double *y, int64_t yc, int64_t yf,
                             const double *gy, int64_t gc, int64_t gf, double *gx, int64_t gxc,
                             int64_t gxf, double *g_dst_verts, double *g_src_verts, void *stream) {
    return guard([&] {
        backward(R, dst, src, transpose, normalize, nfields, x, xc, xf, y, yc, yf, gy, gc, gf, gx, gxc, gxf,
                 g_dst_verts, g_src_verts, as_stream(stream), nullptr);
    });
}

}  // extern "C"
