// grids.hip -- grid generators (host entry points) and device grid creation.
#include <algorithm>
#include <cmath>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <thread>

#include <pthread.h>
#include <vector>

#include "gridgen.hpp"
#include "planar.hpp"
#include "sphgeom.hpp"

namespace cr {

// Cells of point blocks: cell (b,i,j) = P[i,j], P[i+1,j], P[i+1,j+1], P[i,j+1]
// (CellBasedGrid getcell, src/trees/grids.jl:71-84), index b*ni*nj + i + j*ni
// (linear_to_cartesian_idx, src/trees/interfaces.jl:236-243).
__global__ void k_points_to_cells(const double *__restrict__ pts, int64_t nblocks, int64_t ni,
                                  int64_t nj, double *__restrict__ cells, int32_t *__restrict__ vidx) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t per = ni * nj;
    if (c >= nblocks * per) return;
    int64_t b = c / per, r = c - b * per;
    int64_t j = r / ni, i = r - j * ni;
    int64_t W = ni + 1;
    int64_t base = b * W * (nj + 1);
    int64_t idx[4] = {i + j * W, i + 1 + j * W, i + 1 + (j + 1) * W, i + (j + 1) * W};
    for (int k = 0; k < 4; ++k) {
        for (int d = 0; d < 3; ++d) cells[(c * 4 + k) * 3 + d] = pts[(base + idx[k]) * 3 + d];
        vidx[c * 4 + k] = (int32_t)(base + idx[k]);
    }
}

// Tripolar grid with a RightCenterFolded top row (treeify of
// ZRegOrthogonalSphericalShellGrid{..., RightCenterFolded},
// ext/ConservativeRegriddingOceananigansExt.jl:119-187): rows 0..ny-2 are
// CellBasedGrid cells of the (nx+1) x (ny+1) point matrix; fold-row cells i and
// nx-1-i are one physical quadrilateral, so only the primaries i < nq and
// nh <= i < nh+nq (nh = nx/2, nq = nh/2) keep their polygon and the others
// are zero-area ghosts with every vertex at P[0, ny-1].
__host__ __device__ inline bool fold_primary(int64_t i, int64_t nh, int64_t nq) {
    return i < nq || (i >= nh && i < nh + nq);
}

__global__ void k_tripolar_cells(const double *__restrict__ pts, int64_t nx, int64_t ny,
                                 double *__restrict__ cells, int32_t *__restrict__ vidx) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nx * ny) return;
    const int64_t j = c / nx, i = c - j * nx, W = nx + 1;
    int64_t idx[4] = {i + j * W, i + 1 + j * W, i + 1 + (j + 1) * W, i + (j + 1) * W};
    if (j == ny - 1 && !fold_primary(i, nx / 2, nx / 4)) idx[0] = idx[1] = idx[2] = idx[3] = (ny - 1) * W;
    for (int k = 0; k < 4; ++k) {
        for (int d = 0; d < 3; ++d) cells[(c * 4 + k) * 3 + d] = pts[idx[k] * 3 + d];
        vidx[c * 4 + k] = (int32_t)idx[k];
    }
}

// (primary, partner) cell pairs of the fold row: primary local r -> partner
// nx-1-r (mirror_fold_partners!, ext/ConservativeRegriddingOceananigansExt.jl:216-240).
__global__ void k_fold_pairs(int64_t nx, int64_t ny, int32_t *__restrict__ pairs) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nh = nx / 2, nq = nh / 2;
    if (r >= 2 * nq) return;
    const int64_t p = r < nq ? r : r + nh - nq, off = (ny - 1) * nx;
    pairs[2 * r] = (int32_t)(off + p);
    pairs[2 * r + 1] = (int32_t)(off + nx - 1 - p);
}

__global__ void k_iota_i32(int32_t *__restrict__ x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (int32_t)i;
}

__global__ void k_healpix_cells(int64_t nside, int order, int step, int64_t first, int64_t count,
                                double *__restrict__ cells) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int nv = 4 * step;
    if (t >= count * nv) return;
    int64_t c = t / nv;
    int k = (int)(t - c * nv);
    int f;
    int64_t ix, iy;
    hpx_cell_fxy(nside, order, first + c, &f, &ix, &iy);
    hpx_pixel_vertex(f, ix, iy, nside, step, k, cells + 3 * t);
}

// Re-orient clockwise cells to counter-clockwise (reverse the ring) and count them.
__global__ void k_orient_cells(double *__restrict__ cells, int32_t *__restrict__ vidx, int64_t ncells,
                               int nv, int *__restrict__ nflipped) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    double *p = cells + 3 * nv * c;
    if (ring_area(p, nv) >= 0.0) return;
    int32_t *vi = vidx + (int64_t)nv * c;
    for (int a = 1, b = nv - 1; a < b; ++a, --b) {
        for (int d = 0; d < 3; ++d) {
            double t = p[3 * a + d];
            p[3 * a + d] = p[3 * b + d];
            p[3 * b + d] = t;
        }
        int32_t t = vi[a];
        vi[a] = vi[b];
        vi[b] = t;
    }
    atomicAdd(nflipped, 1);
}

__global__ void k_cell_areas(const double *__restrict__ cells, int64_t ncells, int nv,
                             double *__restrict__ area) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    area[c] = ring_area(cells + 3 * nv * c, nv);
}

// Planar cell areas (shoelace) from the caller vertices, in the slot order the
// orientation pass left.
__global__ void k_planar_areas(const double *__restrict__ xy, const int32_t *__restrict__ vidx, int64_t ncells,
                               int nv, double *__restrict__ area) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    p2 r[kMaxCellVertices];
    for (int k = 0; k < nv; ++k) r[k] = planar_vertex(xy, vidx, c * nv + k);
    area[c] = planar_ring_area(r, nv);
}

void launch_planar_areas(const cr_grid *g, double *area, hipStream_t s) {
    hipLaunchKernelGGL(k_planar_areas, dim3(grid_for(g->ncells, 256)), dim3(256), 0, s, g->xy.p, g->vidx.p,
                       g->ncells, g->nv, area);
}

void orient_grid(cr_grid *g, hipStream_t s) {
    DevBuf<int> cnt(1, s);
    CR_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_orient_cells, dim3(grid_for(g->ncells, 256)), dim3(256), 0, s, g->cells.p,
                       g->vidx.p, g->ncells, g->nv, cnt.p);
    CR_HIP(hipGetLastError());
    int h = 0;
    CR_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(int), hipMemcpyDeviceToHost, s));
    CR_HIP(hipStreamSynchronize(s));
    g->flipped = h;
}

// Host worker threads of the input checks, started once and asleep between
// calls (starting 16 threads per call cost 0.3 ms of the 0.35 ms unit check of
// a 1 M-point matrix).  A call's range is cut into chunks that the workers and
// the calling thread take from one counter; a call that finds the workers
// busy with another thread's call runs its range alone.  A forked child has
// no workers: its pool is reset to run every call on the calling thread.
class HostPool {
  public:
    static HostPool &get() {
        static HostPool *p = new HostPool();  // never destroyed: the workers sleep until exit
        return *p;
    }
    unsigned size() const { return nworkers_ + 1; }
    void run(int64_t nchunks, const std::function<void(int64_t)> &fn) {
        std::unique_lock<std::mutex> call(call_mu_, std::try_to_lock);
        if (!call.owns_lock() || nworkers_ == 0) {
            for (int64_t c = 0; c < nchunks; ++c) fn(c);
            return;
        }
        auto job = std::make_shared<Job>(fn, nchunks);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = job;
            ++generation_;
        }
        cv_.notify_all();
        work(*job);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return job->done.load() == nchunks; });
        job_.reset();
    }

  private:
    // One call's chunks: every thread that holds the job takes chunks from its
    // own counter, so a worker that wakes late finds them taken and never runs
    // another call's chunks under this call's count.
    struct Job {
        Job(const std::function<void(int64_t)> &f, int64_t n) : fn(&f), n(n) {}
        const std::function<void(int64_t)> *fn;
        const int64_t n;
        std::atomic<int64_t> next{0}, done{0};
    };
    HostPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        nworkers_ = std::max(1u, std::min(16u, hw)) - 1;
        for (unsigned t = 0; t < nworkers_; ++t) std::thread([this] { loop(); }).detach();
        pthread_atfork(nullptr, nullptr, [] { HostPool::get().reset_in_child(); });
    }
    // the child of a fork: the workers did not follow, and a lock held at the
    // fork stays held; fresh locks, no workers
    void reset_in_child() {
        new (&call_mu_) std::mutex();
        new (&mu_) std::mutex();
        new (&cv_) std::condition_variable();
        new (&done_cv_) std::condition_variable();
        new (&job_) std::shared_ptr<Job>();
        nworkers_ = 0;
    }
    void work(Job &job) {
        int64_t finished = 0;
        for (int64_t c; (c = job.next.fetch_add(1)) < job.n;) {
            (*job.fn)(c);
            ++finished;
        }
        if (finished && job.done.fetch_add(finished) + finished == job.n) {
            std::lock_guard<std::mutex> lk(mu_);
            done_cv_.notify_all();
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            std::shared_ptr<Job> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return generation_ != seen; });
                seen = generation_;
                job = job_;
            }
            if (job) work(*job);
        }
    }
    unsigned nworkers_ = 0;
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    std::shared_ptr<Job> job_;
    uint64_t generation_ = 0;
};

template <class F>
void parallel_for(int64_t n, F &&f, int64_t work_per_item = 1) {
    if (n <= 0) return;
    HostPool &pool = HostPool::get();
    if (n * work_per_item < 4096 || pool.size() == 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    const int64_t nchunks = std::min<int64_t>(n, 4 * (int64_t)pool.size());
    const int64_t chunk = (n + nchunks - 1) / nchunks;
    const std::function<void(int64_t)> job = [&](int64_t c) {
        const int64_t a = c * chunk, b = std::min<int64_t>(n, a + chunk);
        for (int64_t i = a; i < b; ++i) f(i);
    };
    pool.run((n + chunk - 1) / chunk, job);
}

// Host checks of a caller's cells before they reach the device.  Every
// coordinate finite and every vertex a unit vector: the points of the
// spherical manifold (UnitSphericalPoint, src/trees/interfaces.jl:125-126;
// CR_INVALID_ARGUMENT otherwise).  No edge between antipodal points: its great
// circle is undefined, so the clip and the cell area are not
// (CR_DEGENERATE_GEOMETRY).  For a grid declared convex, every cell convex: the
// Sutherland-Hodgman clip needs a convex clip polygon (regridder.jl:96-103;
// CR_DEGENERATE_GEOMETRY).  A cell whose vertices all coincide is legal (a
// zero-area ghost that intersects nothing) and is counted in *collapsed.
// get(c, k) returns vertex k of cell c.
constexpr double kUnitTol = 1e-10;      // | |v|^2 - 1 | of a unit vector
constexpr double kConvexTol = 1e-12;    // sine of the angle a vertex may lie outside an edge
constexpr double kCoincident2 = 1e-28;  // squared chord of coincident vertices

// first(x, c): x = min(x, c), so that a report names the lowest offending cell
inline void first(std::atomic<int64_t> &x, int64_t c) {
    int64_t cur = x.load();
    while (c < cur && !x.compare_exchange_weak(cur, c)) {
    }
}

template <class Get>
void validate_cells(int64_t ncells, int nv, bool convex, Get get, const char *what, int64_t *collapsed) {
    constexpr int64_t kNone = INT64_MAX;
    std::atomic<int64_t> bad_unit{kNone}, bad_anti{kNone}, bad_convex{kNone}, ncoll{0};
    parallel_for(ncells, [&](int64_t c) {
        bool any_edge = false;
        double smin = 0.0, smax = 0.0;
        for (int k = 0; k < nv; ++k) {
            const double *p = get(c, k);
            const d3 a = load3(p);
            const double r2 = dot(a, a);
            if (!(fabs(r2 - 1.0) <= kUnitTol)) {  // also false for NaN or infinite coordinates
                first(bad_unit, c);
                return;
            }
        }
        for (int k = 0; k < nv; ++k) {
            const d3 a = load3(get(c, k)), b = load3(get(c, (k + 1) % nv));
            const d3 e = b - a;
            if (dot(e, e) < kCoincident2) continue;  // a repeated vertex (pole corner, padding)
            any_edge = true;
            const d3 s = a + b;
            if (dot(s, s) < kCoincident2) {
                first(bad_anti, c);
                return;
            }
            if (!convex) continue;
            const d3 n = cross(a, e);
            const double nn = dot(n, n);
            if (nn < kDegenerateEdge2) continue;
            const double inv = 1.0 / sqrt(nn);
            for (int j = 0; j < nv; ++j) {
                const double sj = dot(n, load3(get(c, j)) - a) * inv;
                smin = std::min(smin, sj);
                smax = std::max(smax, sj);
            }
        }
        if (!any_edge) ncoll.fetch_add(1);
        // convex: every vertex on one side of every edge (either orientation;
        // cells are turned counter-clockwise on the device)
        if (convex && smin < -kConvexTol && smax > kConvexTol) first(bad_convex, c);
    });
    CR_REQUIRE(bad_unit.load() == kNone, CR_INVALID_ARGUMENT,
               std::string(what) + ": cell " + std::to_string(bad_unit.load()) +
                   " has a vertex that is not a finite unit vector (grids are on the unit sphere)");
    CR_REQUIRE(bad_anti.load() == kNone, CR_DEGENERATE_GEOMETRY,
               std::string(what) + ": cell " + std::to_string(bad_anti.load()) +
                   " has an edge between antipodal points (its great circle is undefined)");
    CR_REQUIRE(bad_convex.load() == kNone, CR_DEGENERATE_GEOMETRY,
               std::string(what) + ": cell " + std::to_string(bad_convex.load()) +
                   " is not convex but the grid is declared convex (the clip needs a convex side)");
    if (collapsed) *collapsed = ncoll.load();
}

// Geometric checks of one cell of a point block (corners in CellBasedGrid
// order): an edge between antipodal points (*anti), no edge at all (all four
// corners coincide: a legal zero-area cell, *collapsed), or a corner outside
// an edge by more than kConvexTol (*nonconvex).  The same tests as
// validate_cells, for host and device.
CR_HD void check_quad(const d3 (&v)[4], bool *anti, bool *collapsed, bool *nonconvex) {
    bool any_edge = false;
    double smin = 0.0, smax = 0.0;
    *anti = *collapsed = *nonconvex = false;
    for (int k = 0; k < 4; ++k) {
        const d3 a = v[k], e = v[(k + 1) & 3] - a;
        if (dot(e, e) < kCoincident2) continue;  // a repeated vertex (pole corner)
        any_edge = true;
        const d3 sum = a + v[(k + 1) & 3];
        if (dot(sum, sum) < kCoincident2) {
            *anti = true;
            return;
        }
        const d3 n = cross(a, e);
        const double nn = dot(n, n);
        if (nn < kDegenerateEdge2) continue;
        const double inv = 1.0 / sqrt(nn);
        for (int q = 0; q < 4; ++q) {
            const double sq = dot(n, v[q] - a) * inv;
            smin = fmin(smin, sq);
            smax = fmax(smax, sq);
        }
    }
    *collapsed = !any_edge;
    *nonconvex = smin < -kConvexTol && smax > kConvexTol;
}

CR_HD void point_cell(const double *pts, int64_t ni, int64_t nj, int64_t c, d3 (&v)[4]) {
    const int64_t W = ni + 1, per = ni * nj, b = c / per, r = c - b * per, i = r % ni, j = r / ni;
    const double *p0 = pts + 3 * (b * W * (nj + 1) + j * W), *p1 = p0 + 3 * W;
    v[0] = load3(p0 + 3 * i);
    v[1] = load3(p0 + 3 * (i + 1));
    v[2] = load3(p1 + 3 * (i + 1));
    v[3] = load3(p1 + 3 * i);
}

// Every point of nblocks point matrices ((ni+1) x (nj+1), i fastest) a finite
// unit vector, each tested once; a bad point is reported as the lowest cell it
// belongs to.
void validate_point_units(const double *pts, int64_t nblocks, int64_t ni, int64_t nj, const char *what) {
    const int64_t W = ni + 1, H = nj + 1, per = ni * nj;
    std::atomic<int64_t> bad_unit{INT64_MAX};
    parallel_for(nblocks * H, [&](int64_t row) {
        const int64_t b = row / H, j = row - b * H;
        const double *p = pts + 3 * row * W;
        for (int64_t i = 0; i < W; ++i) {
            const d3 a = load3(p + 3 * i);
            if (!(fabs(dot(a, a) - 1.0) <= kUnitTol)) {  // also false for NaN or infinite coordinates
                const int64_t ci = std::min(std::max<int64_t>(i - 1, 0), ni - 1);
                const int64_t cj = std::min(std::max<int64_t>(j - 1, 0), nj - 1);
                first(bad_unit, b * per + ci + cj * ni);
            }
        }
    }, W);
    CR_REQUIRE(bad_unit.load() == INT64_MAX, CR_INVALID_ARGUMENT,
               std::string(what) + ": cell " + std::to_string(bad_unit.load()) +
                   " has a vertex that is not a finite unit vector (grids are on the unit sphere)");
}

void raise_point_cells(const char *what, int64_t anti, int64_t nonconvex) {
    CR_REQUIRE(anti == INT64_MAX, CR_DEGENERATE_GEOMETRY,
               std::string(what) + ": cell " + std::to_string(anti) +
                   " has an edge between antipodal points (its great circle is undefined)");
    CR_REQUIRE(nonconvex == INT64_MAX, CR_DEGENERATE_GEOMETRY,
               std::string(what) + ": cell " + std::to_string(nonconvex) +
                   " is not convex but the grid is declared convex (the clip needs a convex side)");
}

// The checks of validate_cells for the cells of point blocks, on the host.
void validate_point_blocks(const double *pts, int64_t nblocks, int64_t ni, int64_t nj, const char *what,
                           int64_t *collapsed) {
    validate_point_units(pts, nblocks, ni, nj, what);
    std::atomic<int64_t> bad_anti{INT64_MAX}, bad_convex{INT64_MAX}, ncoll{0};
    parallel_for(nblocks * ni * nj, [&](int64_t c) {
        d3 v[4];
        point_cell(pts, ni, nj, c, v);
        bool anti, coll, nonconvex;
        check_quad(v, &anti, &coll, &nonconvex);
        if (anti) first(bad_anti, c);
        else if (nonconvex) first(bad_convex, c);
        if (!anti && coll) ncoll.fetch_add(1);
    });
    raise_point_cells(what, bad_anti.load(), bad_convex.load());
    if (collapsed) *collapsed = ncoll.load();
}

// check_quad for every cell of uploaded point blocks: res[0] / res[1] the
// lowest cell with an antipodal edge / that is not convex (all bits set: none).
__global__ void k_check_point_cells(const double *__restrict__ pts, int64_t nblocks, int64_t ni, int64_t nj,
                                    unsigned long long *__restrict__ res) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nblocks * ni * nj) return;
    d3 v[4];
    point_cell(pts, ni, nj, c, v);
    bool anti, coll, nonconvex;
    check_quad(v, &anti, &coll, &nonconvex);
    if (anti) atomicMin(res, (unsigned long long)c);
    else if (nonconvex) atomicMin(res + 1, (unsigned long long)c);
}

// Vertex count of polygon c of a ragged list; a ring whose last index repeats
// its first is closed (the GI.LinearRing convention of Trees.getcell) and
// loses the repeat.
int64_t polygon_len(const int64_t *off, const int32_t *verts, int64_t c) {
    const int64_t a = off[c], b = off[c + 1];
    int64_t n = b - a;
    if (n >= 2 && verts[b - 1] == verts[a]) --n;
    return n;
}

// Validate a ragged polygon list and return its widest polygon.
int32_t polygons_width(int64_t nverts, const int64_t *off, const int32_t *verts, int64_t ncells) {
    CR_REQUIRE(off[0] == 0, CR_INVALID_ARGUMENT, "polygon offsets must start at 0");
    int64_t width = 0;
    for (int64_t c = 0; c < ncells; ++c) {
        CR_REQUIRE(off[c + 1] >= off[c], CR_INVALID_ARGUMENT, "polygon offsets must not decrease");
        const int64_t n = polygon_len(off, verts, c);
        CR_REQUIRE(n >= 3 && n <= kMaxCellVertices, CR_INVALID_ARGUMENT,
                   "polygon " + std::to_string(c) + " has " + std::to_string(n) +
                       " vertices (need 3.." + std::to_string(kMaxCellVertices) + ")");
        width = std::max(width, n);
    }
    std::atomic<int64_t> bad{-1};
    parallel_for(off[ncells], [&](int64_t k) {
        if (verts[k] < 0 || verts[k] >= nverts) bad.store(k);
    });
    CR_REQUIRE(bad.load() < 0, CR_INVALID_ARGUMENT,
               "polygon vertex index out of range at position " + std::to_string(bad.load()));
    return (int32_t)width;
}

// Fixed-width cell table of a ragged polygon list: polygon c occupies slots
// [c*nv, (c+1)*nv); a polygon with fewer than nv vertices repeats its last
// vertex.  The repeats are zero-length edges, which the clip (zero normal:
// every point is inside), the area fan (zero triangles) and the caps ignore.
void polygons_fill(const double *xyz, const int64_t *off, const int32_t *verts, int64_t ncells, int nv,
                   double *cells, int32_t *vidx) {
    parallel_for(ncells, [&](int64_t c) {
        const int64_t n = polygon_len(off, verts, c);
        for (int k = 0; k < nv; ++k) {
            const int32_t v = verts[off[c] + std::min<int64_t>(k, n - 1)];
            for (int d = 0; d < 3; ++d) cells[(c * nv + k) * 3 + d] = xyz[3 * (int64_t)v + d];
            if (vidx) vidx[c * nv + k] = v;
        }
    });
}

}  // namespace cr

using namespace cr;

extern "C" {

cr_status cr_gen_lonlat_points(double *xyz, int64_t ni, int64_t nj, double lon0, double lat0,
                               double dlon, double dlat) {
    return guard([&] {
        CR_REQUIRE(xyz && ni > 0 && nj > 0, CR_INVALID_ARGUMENT, "cr_gen_lonlat_points: bad arguments");
        double lat_end = lat0 + (double)nj * dlat;
        CR_REQUIRE(lat0 >= -90.0 && lat0 <= 90.0 && lat_end >= -90.0 && lat_end <= 90.0,
                   CR_INVALID_ARGUMENT, "latitudes outside [-90, 90]");
        int64_t W = ni + 1;
        parallel_for((nj + 1) * W, [&](int64_t t) {
            int64_t j = t / W, i = t - j * W;
            lonlat_to_xyz(lon0 + (double)i * dlon, lat0 + (double)j * dlat, xyz + 3 * t);
        });
    });
}

cr_status cr_lonlat_to_xyz(const double *lon_deg, const double *lat_deg, int64_t n, double *xyz) {
    return guard([&] {
        CR_REQUIRE(lon_deg && lat_deg && xyz && n >= 0, CR_INVALID_ARGUMENT, "cr_lonlat_to_xyz: bad arguments");
        parallel_for(n, [&](int64_t t) { lonlat_to_xyz(lon_deg[t], lat_deg[t], xyz + 3 * t); });
    });
}

cr_status cr_gen_cubed_sphere_points(double *xyz, int64_t ne) {
    return guard([&] {
        CR_REQUIRE(xyz && ne > 0, CR_INVALID_ARGUMENT, "cr_gen_cubed_sphere_points: bad arguments");
        int64_t W = ne + 1;
        parallel_for(6 * W * W, [&](int64_t t) {
            int64_t f = t / (W * W), r = t - f * W * W;
            int64_t j = r / W, i = r - j * W;
            cube_point((int)f, i, j, ne, xyz + 3 * t);
        });
    });
}

cr_status cr_gen_healpix_cells(double *xyz, int64_t nside, int32_t order, int32_t step,
                               int64_t first, int64_t count) {
    return guard([&] {
        CR_REQUIRE(xyz && nside > 0 && step > 0, CR_INVALID_ARGUMENT, "cr_gen_healpix_cells: bad arguments");
        CR_REQUIRE(order == CR_HEALPIX_NESTED || order == CR_HEALPIX_RING, CR_INVALID_ARGUMENT,
                   "HEALPix order must be NESTED or RING");
        CR_REQUIRE((order == CR_HEALPIX_RING) || ((nside & (nside - 1)) == 0), CR_INVALID_ARGUMENT,
                   "nested HEALPix ordering needs nside to be a power of 2");
        CR_REQUIRE(nside <= (1 << 20), CR_INVALID_ARGUMENT, "nside too large");
        int64_t npix = 12 * nside * nside;
        CR_REQUIRE(first >= 0 && count >= 0 && first + count <= npix, CR_INVALID_ARGUMENT,
                   "HEALPix cell range out of bounds");
        int nv = 4 * step;
        parallel_for(count, [&](int64_t c) {
            int f;
            int64_t ix, iy;
            hpx_cell_fxy(nside, order, first + c, &f, &ix, &iy);
            for (int k = 0; k < nv; ++k) hpx_pixel_vertex(f, ix, iy, nside, step, k, xyz + 3 * (c * nv + k));
        });
    });
}

cr_status cr_healpix_nest2ring(int64_t nside, int64_t nest, int64_t *ring) {
    return guard([&] {
        CR_REQUIRE(nside > 0 && (nside & (nside - 1)) == 0 && ring, CR_INVALID_ARGUMENT, "bad nside");
        CR_REQUIRE(nest >= 0 && nest < 12 * nside * nside, CR_INVALID_ARGUMENT, "pixel out of range");
        int f;
        int64_t ix, iy;
        hpx_nest2fxy(nside, nest, &f, &ix, &iy);
        *ring = hpx_fxy2ring(nside, f, ix, iy);
    });
}

cr_status cr_healpix_ring2nest(int64_t nside, int64_t ring, int64_t *nest) {
    return guard([&] {
        CR_REQUIRE(nside > 0 && (nside & (nside - 1)) == 0 && nest, CR_INVALID_ARGUMENT, "bad nside");
        CR_REQUIRE(ring >= 0 && ring < 12 * nside * nside, CR_INVALID_ARGUMENT, "pixel out of range");
        int f;
        int64_t ix, iy;
        hpx_ring2fxy(nside, ring, &f, &ix, &iy);
        *nest = hpx_fxy2nest(nside, f, ix, iy);
    });
}

cr_status cr_points_to_cells(const double *pts, int64_t nblocks, int64_t ni, int64_t nj,
                             double *cells) {
    return guard([&] {
        CR_REQUIRE(pts && cells && nblocks > 0 && ni > 0 && nj > 0, CR_INVALID_ARGUMENT,
                   "cr_points_to_cells: bad arguments");
        int64_t per = ni * nj, W = ni + 1;
        parallel_for(nblocks * per, [&](int64_t c) {
            int64_t b = c / per, r = c - b * per;
            int64_t j = r / ni, i = r - j * ni;
            const double *P = pts + 3 * b * W * (nj + 1);
            int64_t idx[4] = {i + j * W, i + 1 + j * W, i + 1 + (j + 1) * W, i + (j + 1) * W};
            for (int k = 0; k < 4; ++k)
                for (int d = 0; d < 3; ++d) cells[(c * 4 + k) * 3 + d] = P[idx[k] * 3 + d];
        });
    });
}

cr_status cr_grid_from_cells(cr_grid **grid, const double *cells, int64_t ncells, int32_t nv,
                             int32_t convex, void *stream) {
    return guard([&] {
        CR_REQUIRE(grid && cells && ncells > 0 && nv >= 3 && nv <= kMaxCellVertices, CR_INVALID_ARGUMENT,
                   "cr_grid_from_cells: need ncells > 0 and 3 <= nv <= 64");
        CR_REQUIRE(ncells < (int64_t(1) << 31), CR_UNSUPPORTED, "more than 2^31-1 cells");
        validate_cells(
            ncells, nv, convex != 0, [&](int64_t c, int k) { return cells + 3 * (c * nv + k); },
            "cr_grid_from_cells", nullptr);
        check_device();
        hipStream_t s = as_stream(stream);
        auto *g = new cr_grid();
        try {
            g->ncells = ncells;
            g->nv = nv;
            g->convex = convex ? 1 : 0;
            g->cells.alloc((size_t)ncells * nv * 3);
            CR_HIP(hipMemcpyAsync(g->cells.p, cells, g->cells.bytes(), hipMemcpyHostToDevice, s));
            g->nverts = ncells * nv;
            g->vidx.alloc((size_t)ncells * nv);
            hipLaunchKernelGGL(k_iota_i32, dim3(grid_for(g->nverts, 256)), dim3(256), 0, s, g->vidx.p,
                               g->nverts);
            g->private_vertices = 1;  // orientation only permutes a cell's own slots
            orient_grid(g, s);
        } catch (...) {
            delete g;
            throw;
        }
        *grid = g;
    });
}

cr_status cr_polygons_to_cells(const double *xyz, int64_t nverts, const int64_t *offsets,
                               const int32_t *verts, int64_t ncells, int32_t nv, double *cells,
                               int32_t *vidx) {
    return guard([&] {
        CR_REQUIRE(xyz && offsets && verts && cells && nverts > 0 && ncells > 0, CR_INVALID_ARGUMENT,
                   "cr_polygons_to_cells: bad arguments");
        const int32_t width = polygons_width(nverts, offsets, verts, ncells);
        CR_REQUIRE(nv >= width && nv <= kMaxCellVertices, CR_INVALID_ARGUMENT,
                   "cr_polygons_to_cells: nv is below the widest polygon (" + std::to_string(width) + ")");
        polygons_fill(xyz, offsets, verts, ncells, nv, cells, vidx);
    });
}

cr_status cr_polygons_width(int64_t nverts, const int64_t *offsets, const int32_t *verts, int64_t ncells,
                            int32_t *nv) {
    return guard([&] {
        CR_REQUIRE(offsets && verts && nv && nverts > 0 && ncells > 0, CR_INVALID_ARGUMENT,
                   "cr_polygons_width: bad arguments");
        *nv = polygons_width(nverts, offsets, verts, ncells);
    });
}

// A device grid from a ragged polygon list of unit vectors (checked, then
// oriented counter-clockwise on the device).  `planar_xy` (may be null): the
// caller's planar vertices the unit vectors are the images of, kept on the
// grid for the planar clip and areas.
cr_grid *grid_from_polygon_list(const double *xyz, int64_t nverts, const int64_t *offsets, const int32_t *verts,
                                int64_t ncells, int32_t convex, const double *planar_xy, const char *what,
                                hipStream_t s) {
    CR_REQUIRE(ncells < (int64_t(1) << 31) && nverts < (int64_t(1) << 31), CR_UNSUPPORTED,
               "more than 2^31-1 cells or vertices");
    const int32_t nv = polygons_width(nverts, offsets, verts, ncells);
    std::vector<double> cells((size_t)ncells * nv * 3);
    std::vector<int32_t> vidx((size_t)ncells * nv);
    polygons_fill(xyz, offsets, verts, ncells, nv, cells.data(), vidx.data());
    const double *cp = cells.data();
    validate_cells(
        ncells, nv, convex != 0, [&](int64_t c, int k) { return cp + 3 * (c * nv + k); }, what, nullptr);
    check_device();
    auto *g = new cr_grid();
    try {
        g->ncells = ncells;
        g->nv = nv;
        g->convex = convex ? 1 : 0;
        g->nverts = nverts;
        g->cells.alloc(cells.size());
        g->vidx.alloc(vidx.size());
        CR_HIP(hipMemcpyAsync(g->cells.p, cells.data(), g->cells.bytes(), hipMemcpyHostToDevice, s));
        CR_HIP(hipMemcpyAsync(g->vidx.p, vidx.data(), g->vidx.bytes(), hipMemcpyHostToDevice, s));
        if (planar_xy) {
            g->planar = 1;
            g->xy.alloc((size_t)nverts * 2);
            CR_HIP(hipMemcpyAsync(g->xy.p, planar_xy, g->xy.bytes(), hipMemcpyHostToDevice, s));
        }
        CR_HIP(hipStreamSynchronize(s));  // the host tables end with this scope
        orient_grid(g, s);
    } catch (...) {
        (void)hipStreamSynchronize(s);
        delete g;
        throw;
    }
    return g;
}

cr_status cr_grid_from_polygons(cr_grid **grid, const double *xyz, int64_t nverts, const int64_t *offsets,
                                const int32_t *verts, int64_t ncells, int32_t convex, void *stream) {
    return guard([&] {
        CR_REQUIRE(grid && xyz && offsets && verts && nverts > 0 && ncells > 0, CR_INVALID_ARGUMENT,
                   "cr_grid_from_polygons: bad arguments");
        *grid = grid_from_polygon_list(xyz, nverts, offsets, verts, ncells, convex, nullptr, "cr_grid_from_polygons",
                                       as_stream(stream));
    });
}

cr_status cr_grid_from_planar_polygons(cr_grid **grid, const double *xy, int64_t nverts, const int64_t *offsets,
                                       const int32_t *verts, int64_t ncells, int32_t convex, void *stream) {
    return guard([&] {
        CR_REQUIRE(grid && xy && offsets && verts && nverts > 0 && ncells > 0, CR_INVALID_ARGUMENT,
                   "cr_grid_from_planar_polygons: bad arguments");
        double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
        for (int64_t v = 0; v < nverts; ++v)
            for (int d = 0; d < 2; ++d) {
                const double c = xy[2 * v + d];
                CR_REQUIRE(std::isfinite(c), CR_INVALID_ARGUMENT,
                           "cr_grid_from_planar_polygons: vertex " + std::to_string(v) + " is not finite");
                lo[d] = std::min(lo[d], c);
                hi[d] = std::max(hi[d], c);
            }
        const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]);
        const double h = 0.5 * std::max(hi[0] - lo[0], hi[1] - lo[1]);
        CR_REQUIRE(h > 0.0 && std::isfinite(h), CR_INVALID_ARGUMENT,
                   "cr_grid_from_planar_polygons: the vertices span no area");
        // gnomonic images: (x, y) -> ((x - cx) / h, (y - cy) / h, 1) / norm, the
        // plane tangent at the pole seen from the centre of the sphere; the
        // extent maps to within 55 degrees of the pole
        std::vector<double> xyz((size_t)nverts * 3);
        for (int64_t v = 0; v < nverts; ++v) {
            const double X = (xy[2 * v] - cx) / h, Y = (xy[2 * v + 1] - cy) / h;
            const double r = 1.0 / std::sqrt(X * X + Y * Y + 1.0);
            xyz[3 * v] = X * r;
            xyz[3 * v + 1] = Y * r;
            xyz[3 * v + 2] = r;
        }
        *grid = grid_from_polygon_list(xyz.data(), nverts, offsets, verts, ncells, convex, xy,
                                       "cr_grid_from_planar_polygons", as_stream(stream));
    });
}

cr_status cr_grid_from_points(cr_grid **grid, const double *pts, int64_t nblocks, int64_t ni,
                              int64_t nj, void *stream) {
    return guard([&] {
        CR_REQUIRE(grid && pts && nblocks > 0 && ni > 0 && nj > 0, CR_INVALID_ARGUMENT,
                   "cr_grid_from_points: bad arguments");
        CR_REQUIRE(nblocks * ni * nj < (int64_t(1) << 31), CR_UNSUPPORTED, "more than 2^31-1 cells");
        // the points on the host, the cells' geometry on the device (the same
        // tests as cr_validate_points, without a host pass over every cell)
        validate_point_units(pts, nblocks, ni, nj, "cr_grid_from_points");
        check_device();
        hipStream_t s = as_stream(stream);
        int64_t npts = nblocks * (ni + 1) * (nj + 1);
        DevBuf<double> dpts((size_t)npts * 3, s);  // pool memory, freed in order on s
        CR_HIP(hipMemcpyAsync(dpts.p, pts, dpts.bytes(), hipMemcpyHostToDevice, s));
        {
            DevBuf<unsigned long long> res(2, s);
            CR_HIP(hipMemsetAsync(res.p, 0xff, res.bytes(), s));
            hipLaunchKernelGGL(k_check_point_cells, dim3(grid_for(nblocks * ni * nj, 256)), dim3(256), 0, s, dpts.p,
                               nblocks, ni, nj, res.p);
            CR_HIP(hipGetLastError());
            unsigned long long h[2] = {0, 0};
            CR_HIP(hipMemcpyAsync(h, res.p, sizeof(h), hipMemcpyDeviceToHost, s));
            CR_HIP(hipStreamSynchronize(s));
            auto cell = [](unsigned long long v) { return v == ~0ull ? INT64_MAX : (int64_t)v; };
            raise_point_cells("cr_grid_from_points", cell(h[0]), cell(h[1]));
        }
        auto *g = new cr_grid();
        try {
            g->ncells = nblocks * ni * nj;
            g->nv = 4;
            g->convex = 1;
            g->cells.alloc((size_t)g->ncells * 4 * 3);
            g->nverts = npts;
            g->vidx.alloc((size_t)g->ncells * 4);
            hipLaunchKernelGGL(k_points_to_cells, dim3(grid_for(g->ncells, 256)), dim3(256), 0, s,
                               dpts.p, nblocks, ni, nj, g->cells.p, g->vidx.p);
            CR_HIP(hipGetLastError());
            orient_grid(g, s);
        } catch (...) {
            delete g;
            throw;
        }
        *grid = g;
    });
}

cr_status cr_grid_tripolar(cr_grid **grid, const double *pts, int64_t nx, int64_t ny, void *stream) {
    return guard([&] {
        CR_REQUIRE(grid && pts && nx >= 2 && ny >= 2, CR_INVALID_ARGUMENT,
                   "cr_grid_tripolar: need nx >= 2 and ny >= 2");
        CR_REQUIRE(nx % 2 == 0, CR_INVALID_ARGUMENT, "a RightCenterFolded grid needs an even number of cells in x");
        CR_REQUIRE(nx * ny < (int64_t(1) << 31), CR_UNSUPPORTED, "more than 2^31-1 cells");
        validate_point_blocks(pts, 1, nx, ny, "cr_grid_tripolar", nullptr);
        check_device();
        hipStream_t s = as_stream(stream);
        const int64_t npts = (nx + 1) * (ny + 1);
        DevBuf<double> dpts((size_t)npts * 3, s);
        CR_HIP(hipMemcpyAsync(dpts.p, pts, dpts.bytes(), hipMemcpyHostToDevice, s));
        auto *g = new cr_grid();
        try {
            g->ncells = nx * ny;
            g->nv = 4;
            g->convex = 1;
            g->nverts = npts;
            g->cells.alloc((size_t)g->ncells * 4 * 3);
            g->vidx.alloc((size_t)g->ncells * 4);
            hipLaunchKernelGGL(k_tripolar_cells, dim3(grid_for(g->ncells, 256)), dim3(256), 0, s, dpts.p, nx, ny,
                               g->cells.p, g->vidx.p);
            g->nfold = 2 * (nx / 4);
            if (g->nfold) {
                g->fold.alloc((size_t)g->nfold * 2);
                hipLaunchKernelGGL(k_fold_pairs, dim3(grid_for(g->nfold, 256)), dim3(256), 0, s, nx, ny, g->fold.p);
            }
            CR_HIP(hipGetLastError());
            orient_grid(g, s);
        } catch (...) {
            (void)hipStreamSynchronize(s);
            delete g;
            throw;
        }
        *grid = g;
    });
}

cr_status cr_validate_cells(const double *cells, int64_t ncells, int32_t nv, int32_t convex, int64_t *collapsed) {
    return guard([&] {
        CR_REQUIRE(cells && ncells > 0 && nv >= 3 && nv <= kMaxCellVertices, CR_INVALID_ARGUMENT,
                   "cr_validate_cells: need ncells > 0 and 3 <= nv <= 64");
        validate_cells(
            ncells, nv, convex != 0, [&](int64_t c, int k) { return cells + 3 * (c * nv + k); },
            "cr_validate_cells", collapsed);
    });
}

cr_status cr_validate_points(const double *pts, int64_t nblocks, int64_t ni, int64_t nj, int64_t *collapsed) {
    return guard([&] {
        CR_REQUIRE(pts && nblocks > 0 && ni > 0 && nj > 0, CR_INVALID_ARGUMENT, "cr_validate_points: bad arguments");
        validate_point_blocks(pts, nblocks, ni, nj, "cr_validate_points", collapsed);
    });
}

cr_status cr_grid_fold_count(const cr_grid *grid, int64_t *npairs) {
    return guard([&] {
        CR_REQUIRE(grid && npairs, CR_INVALID_ARGUMENT, "null argument");
        *npairs = grid->nfold;
    });
}

cr_status cr_grid_healpix(cr_grid **grid, int64_t nside, int32_t order, int32_t step, void *stream) {
    return guard([&] {
        CR_REQUIRE(grid && nside > 0 && step > 0 && step <= 16, CR_INVALID_ARGUMENT,
                   "cr_grid_healpix: need nside > 0 and 1 <= step <= 16");
        CR_REQUIRE(order == CR_HEALPIX_NESTED || order == CR_HEALPIX_RING, CR_INVALID_ARGUMENT,
                   "HEALPix order must be NESTED or RING");
        CR_REQUIRE((order == CR_HEALPIX_RING) || ((nside & (nside - 1)) == 0), CR_INVALID_ARGUMENT,
                   "nested HEALPix ordering needs nside to be a power of 2");
        CR_REQUIRE(nside <= 8192, CR_UNSUPPORTED, "nside above 8192");
        check_device();
        hipStream_t s = as_stream(stream);
        auto *g = new cr_grid();
        try {
            g->ncells = 12 * nside * nside;
            g->nv = 4 * step;
            g->convex = step == 1 ? 1 : 0;
            g->cells.alloc((size_t)g->ncells * g->nv * 3);
            int64_t nthr = g->ncells * g->nv;
            hipLaunchKernelGGL(k_healpix_cells, dim3(grid_for(nthr, 256)), dim3(256), 0, s, nside, order,
                               step, int64_t(0), g->ncells, g->cells.p);
            g->nverts = nthr;
            g->vidx.alloc((size_t)nthr);
            hipLaunchKernelGGL(k_iota_i32, dim3(grid_for(nthr, 256)), dim3(256), 0, s, g->vidx.p, nthr);
            g->private_vertices = 1;
            CR_HIP(hipGetLastError());
            CR_HIP(hipStreamSynchronize(s));
        } catch (...) {
            delete g;
            throw;
        }
        *grid = g;
    });
}

cr_status cr_grid_info(const cr_grid *grid, int64_t *ncells, int32_t *nv, int32_t *convex) {
    return guard([&] {
        CR_REQUIRE(grid, CR_INVALID_ARGUMENT, "null grid");
        if (ncells) *ncells = grid->ncells;
        if (nv) *nv = grid->nv;
        if (convex) *convex = grid->convex;
    });
}

cr_status cr_grid_vertex_count(const cr_grid *grid, int64_t *nverts) {
    return guard([&] {
        CR_REQUIRE(grid && nverts, CR_INVALID_ARGUMENT, "null argument");
        *nverts = grid->nverts;
    });
}

cr_status cr_grid_copy_cells(const cr_grid *grid, double *cells) {
    return guard([&] {
        CR_REQUIRE(grid && cells, CR_INVALID_ARGUMENT, "null argument");
        CR_HIP(hipMemcpy(cells, grid->cells.p, grid->cells.bytes(), hipMemcpyDeviceToHost));
    });
}

cr_status cr_grid_is_planar(const cr_grid *grid, int32_t *planar) {
    return guard([&] {
        CR_REQUIRE(grid && planar, CR_INVALID_ARGUMENT, "null argument");
        *planar = grid->planar;
    });
}

cr_status cr_grid_copy_planar_cells(const cr_grid *grid, double *cells_xy) {
    return guard([&] {
        CR_REQUIRE(grid && cells_xy, CR_INVALID_ARGUMENT, "null argument");
        CR_REQUIRE(grid->planar, CR_INVALID_ARGUMENT, "cr_grid_copy_planar_cells: not a planar grid");
        std::vector<double> xy((size_t)grid->nverts * 2);
        std::vector<int32_t> vi((size_t)grid->ncells * grid->nv);
        CR_HIP(hipMemcpy(xy.data(), grid->xy.p, grid->xy.bytes(), hipMemcpyDeviceToHost));
        CR_HIP(hipMemcpy(vi.data(), grid->vidx.p, grid->vidx.bytes(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < vi.size(); ++k) {
            cells_xy[2 * k] = xy[2 * (size_t)vi[k]];
            cells_xy[2 * k + 1] = xy[2 * (size_t)vi[k] + 1];
        }
    });
}

cr_status cr_grid_areas(const cr_grid *grid, double *areas_host) {
    return guard([&] {
        CR_REQUIRE(grid && areas_host, CR_INVALID_ARGUMENT, "null argument");
        DevBuf<double> a(grid->ncells);
        if (grid->planar)
            launch_planar_areas(grid, a.p, 0);
        else
            hipLaunchKernelGGL(k_cell_areas, dim3(grid_for(grid->ncells, 256)), dim3(256), 0, 0,
                               grid->cells.p, grid->ncells, grid->nv, a.p);
        CR_HIP(hipGetLastError());
        CR_HIP(hipMemcpy(areas_host, a.p, a.bytes(), hipMemcpyDeviceToHost));
    });
}

void cr_grid_destroy(cr_grid *grid) { delete grid; }

}  // extern "C"
