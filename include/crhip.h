/*
 * crhip.h -- C ABI of the MI355X (gfx950) conservative-regridding library.
 *
 * The library replaces what sits behind ConservativeRegridding.jl's two hot
 * stages (paths relative to the reference repository root):
 *
 *   build   Regridder(dst, src; normalize, ...)       src/regridder/regridder.jl:105-163
 *             intersection_areas(manifold, threaded, dst_tree, src_tree; ...)
 *                                                     src/regridder/intersection_areas.jl:67-157
 *             areas(manifold, item, tree)             src/regridder/regridder.jl:165-178
 *             normalize!(regridder)                   src/regridder/regridder.jl:54-62
 *   apply   regrid!(dst, R, src; dims=1, normalize)   src/regridder/regrid.jl:63-118
 *             NDSliceLoop batched slices              src/regridder/regrid.jl:303-318
 *             transpose(R)                            src/regridder/regridder.jl:49-50
 *
 * Conventions
 *   - Every function returns a cr_status; cr_last_error() gives a thread-local
 *     message for the last failing call on the calling thread.
 *   - Cell and point coordinates are unit 3-vectors stored xyz-interleaved as
 *     double.  Cell indices are 0-based (the Julia layer adds 1).
 *   - A polygon table is (ncells x nv) vertices, open rings (no repeated closing
 *     vertex), counter-clockwise seen from outside the sphere.  Cells with fewer
 *     than nv vertices repeat their last vertex.
 *   - Rows of A are destination cells and columns are source cells
 *     (intersection_areas.jl:115-121).
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *   - Handles are immutable after creation except through cr_regridder_normalize;
 *     concurrent cr_regrid calls on one handle with distinct outputs are safe.
 */
#ifndef CRHIP_H
#define CRHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CR_OK = 0,
    CR_INVALID_ARGUMENT = 1,   /* Julia: ArgumentError                         */
    CR_DIMENSION_MISMATCH = 2, /* Julia: DimensionMismatch (regrid.jl:289-296) */
    CR_UNSUPPORTED = 3,        /* Julia: MethodError / ErrorException          */
    CR_OUT_OF_MEMORY = 4,
    CR_HIP_ERROR = 5,
    CR_DEGENERATE_GEOMETRY = 6,
    CR_INTERNAL = 7,
    CR_NO_DEVICE = 8
} cr_status;

typedef enum { CR_F64 = 0, CR_F32 = 1 } cr_dtype;
typedef enum { CR_HEALPIX_NESTED = 0, CR_HEALPIX_RING = 1 } cr_healpix_order;
typedef enum { CR_CLIP_AUTO = 0, CR_CLIP_WITH_DST = 1, CR_CLIP_WITH_SRC = 2 } cr_clip_side;

typedef struct cr_grid cr_grid;
typedef struct cr_regridder cr_regridder;

/* Build options (keyword arguments of Regridder, regridder.jl:125-131). */
typedef struct {
    double threshold_rel; /* drop A_ij <= threshold_rel * min(a_src_j, a_dst_i);
                             default 1e-14; 0 keeps every area > 0 like
                             intersection_areas.jl:24 */
    double radius;        /* sphere radius; areas are in radius^2 units
                             (default 6371008.8, GeometryOpsCore's Spherical()) */
    int32_t normalize;    /* Regridder(...; normalize=true): normalize!(R) */
    int32_t clip_side;    /* cr_clip_side; AUTO clips with a convex side */
    int32_t reserved0;
    int32_t reserved1;
} cr_build_opts;

/* Per-build counters and stage times (device event timing, milliseconds). */
typedef struct {
    int64_t n_dst, n_src, nnz, n_candidates;
    double ms_total, ms_bounds, ms_tree, ms_search, ms_clip, ms_assemble;
    int32_t clip_with_dst; /* 1: dst cells were the clip polygons */
    int32_t flipped_cells; /* cells re-oriented to CCW at grid creation */
    double ms_host;        /* host wall time of the whole build call */
    /* Per-pair failure accounting (the reference's per-pair try/catch,
     * regridder.jl:88-101, is disabled; here every case is counted): */
    int64_t n_zero_area_cells;  /* cells of either grid with zero area (they intersect nothing) */
    int64_t n_zero_cell_pairs;  /* candidate pairs dropped because one of their cells has zero area */
    int64_t n_nonfinite_pairs;  /* candidate pairs whose intersection area is not finite: the build
                                   fails with CR_DEGENERATE_GEOMETRY (cr_last_failed_pair) */
    int64_t n_nonfinite_cells;  /* cells with a non-finite area: the build fails likewise */
} cr_build_stats;

/* ------------------------------------------------------------ library -- */
const char *cr_last_error(void);
/* The cells of the last build on this thread that failed with
 * CR_DEGENERATE_GEOMETRY (-1 where a side is not involved, both -1 after a
 * build that did not fail that way): the pair whose intersection area is not
 * finite, or the one cell whose area is not.  The reference's
 * DefaultIntersectionFailureError carries the two polygons
 * (src/regridder/regridder.jl:64-73). */
cr_status cr_last_failed_pair(int64_t *dst_cell, int64_t *src_cell);
const char *cr_version(void);
/* Number of visible HIP devices (0 on a host without a GPU; never fails). */
int32_t cr_device_count(void);

/* ------------------------------------------- grid generators (host only) -- */
/* These replace the grid packages the reference's adapters call into; they
 * need no GPU.  Point matrices are (ni+1) x (nj+1), i fastest, the layout of
 * the Matrix{UnitSphericalPoint} that Trees.treeify turns into a CellBasedGrid
 * (src/trees/interfaces.jl:125-126, src/trees/grids.jl:57-85). */

/* Lon-lat points lon0 + i*dlon, lat0 + j*dlat (degrees), converted with exact
 * degree sincos (poles and multiples of 90 degrees are exact). */
cr_status cr_gen_lonlat_points(double *xyz, int64_t ni, int64_t nj, double lon0_deg,
                               double lat0_deg, double dlon_deg, double dlat_deg);
/* Equiangular gnomonic cubed sphere: 6 faces of (ne+1)^2 points, face-major.
 * Face f covers cells f*ne^2 + i + j*ne (the offsets of
 * ext/ConservativeRegriddingClimaCoreExt.jl:56-63). */
cr_status cr_gen_cubed_sphere_points(double *xyz, int64_t ne);
/* Unit vectors of n (lon, lat) degree pairs, exact at multiples of 90 degrees
 * (GeometryOps UnitSphereFromGeographic, as used by
 * ext/ConservativeRegriddingRingGridsExt.jl:42-44). */
cr_status cr_lonlat_to_xyz(const double *lon_deg, const double *lat_deg, int64_t n, double *xyz);
/* HEALPix cells [first, first+count) as 4*step-vertex polygons, counter-
 * clockwise from the north corner, in nested or ring index order.  Replaces
 * Healpix.boundariesRing (ext/ConservativeRegriddingHealpixExt.jl:76-90). */
cr_status cr_gen_healpix_cells(double *xyz, int64_t nside, int32_t order, int32_t step,
                               int64_t first, int64_t count);
/* HEALPix index conversions (0-based).  Healpix.nest2ring / ring2nest. */
cr_status cr_healpix_nest2ring(int64_t nside, int64_t nest, int64_t *ring);
cr_status cr_healpix_ring2nest(int64_t nside, int64_t ring, int64_t *nest);
/* Cell polygons of point blocks: nblocks blocks of (ni+1)*(nj+1) points; cell
 * (b,i,j) -> index b*ni*nj + i + j*ni with vertices P[i,j], P[i+1,j],
 * P[i+1,j+1], P[i,j+1] (CellBasedGrid getcell, grids.jl:71-84). */
cr_status cr_points_to_cells(const double *pts, int64_t nblocks, int64_t ni, int64_t nj,
                             double *cells);
/* Ragged polygon lists, the flattened form of any grid that implements
 * Trees.ncells / Trees.getcell (docs/src/interfaces/trees.md:34-44; RingGrids,
 * ISEA and user grids): nverts unit vectors xyz (3 per vertex) shared by the
 * cells, and cell c = vertices verts[offsets[c] .. offsets[c+1]-1] in ring
 * order.  A ring whose last index repeats its first is closed and loses the
 * repeat; every cell needs 3..64 vertices.  cr_polygons_width returns the
 * widest cell; cr_polygons_to_cells writes the fixed-width table of nv >=
 * that width (ncells*nv*3 doubles, narrower cells repeat their last vertex,
 * a zero-length edge that clip and area ignore) and, if vidx is not NULL, the
 * vertex index of every slot. */
cr_status cr_polygons_width(int64_t nverts, const int64_t *offsets, const int32_t *verts,
                            int64_t ncells, int32_t *nv);
cr_status cr_polygons_to_cells(const double *xyz, int64_t nverts, const int64_t *offsets,
                               const int32_t *verts, int64_t ncells, int32_t nv, double *cells,
                               int32_t *vidx);

/* Host checks every grid constructor below makes of a caller's cells before
 * anything reaches the device (they need no GPU).  CR_INVALID_ARGUMENT: a
 * coordinate is not finite or a vertex is not a unit vector (the points of the
 * spherical manifold, UnitSphericalPoint: the reference promotes planar
 * inputs to it in Regridder(dst, src), regridder.jl:105-123).
 * CR_DEGENERATE_GEOMETRY: an edge joins antipodal points (no great circle),
 * or, with convex=1, a cell is not convex (the Sutherland-Hodgman clip of
 * regridder.jl:96-103 needs a convex clip polygon).  Cells whose vertices all
 * coincide are legal zero-area cells that intersect nothing; *collapsed (may
 * be NULL) receives their number.  The message names the lowest offending
 * cell.  cr_validate_points checks the cells of point blocks (layout of
 * cr_grid_from_points, always convex). */
cr_status cr_validate_cells(const double *cells, int64_t ncells, int32_t nv, int32_t convex, int64_t *collapsed);
cr_status cr_validate_points(const double *pts, int64_t nblocks, int64_t ni, int64_t nj, int64_t *collapsed);

/* ------------------------------------------------------ device grids -- */
/* Upload a host polygon table.  `convex` declares every cell convex. */
cr_status cr_grid_from_cells(cr_grid **grid, const double *cells, int64_t ncells, int32_t nv,
                             int32_t convex, void *stream);
/* Upload point blocks (a lon-lat Matrix{UnitSphericalPoint} is one block, a
 * cubed sphere six) and expand them to cells on the device.  The points are
 * checked on the host before anything reaches the device, the cells' geometry
 * on the device after the upload; the errors of cr_validate_points, naming the
 * same lowest cell. */
cr_status cr_grid_from_points(cr_grid **grid, const double *pts, int64_t nblocks, int64_t ni,
                              int64_t nj, void *stream);
/* Upload a ragged polygon list (layout of cr_polygons_to_cells) with shared
 * vertices: the generic tree adapter (Trees.treeify of a grid given by
 * Trees.ncells / Trees.getcell, src/trees/interfaces.jl:94-129).  Vertex
 * gradients are indexed by the nverts shared vertices. */
cr_status cr_grid_from_polygons(cr_grid **grid, const double *xyz, int64_t nverts,
                                const int64_t *offsets, const int32_t *verts, int64_t ncells,
                                int32_t convex, void *stream);
/* A grid on the Planar manifold: the ragged polygon list of cr_grid_from_polygons
 * with 2-D vertices (x, y interleaved, nverts of them), for
 * Regridder(GO.Planar(), dst, src) -- the planar DefaultIntersectionOperator
 * (src/regridder/regridder.jl:87-94) on grids given as polygons
 * (GO.SpatialTreeInterface.FlatNoTree(polys) of test/usecases/simple.jl, the
 * planar rectangle matrices of test/regridding.jl:299-376).  cr_regridder_build
 * of two planar grids clips the rings in the plane and measures areas with the
 * shoelace formula (planar units, no radius); a planar grid next to a
 * spherical one fails with CR_INVALID_ARGUMENT ("manifolds must be the same").
 * Rings of either orientation are turned counter-clockwise; `convex` declares
 * every cell convex (one side of a build must be).  The vertices must span a
 * non-empty box and be finite (CR_INVALID_ARGUMENT).  Vertex gradients,
 * cell averages of the example fields and spectral-element weights are
 * spherical and return CR_UNSUPPORTED on planar grids. */
cr_status cr_grid_from_planar_polygons(cr_grid **grid, const double *xy, int64_t nverts,
                                       const int64_t *offsets, const int32_t *verts, int64_t ncells,
                                       int32_t convex, void *stream);
/* 1 if the grid is on the Planar manifold (GOCore.best_manifold), else 0. */
cr_status cr_grid_is_planar(const cr_grid *grid, int32_t *planar);
/* The oriented planar polygon table of a planar grid: ncells*nv*2 doubles. */
cr_status cr_grid_copy_planar_cells(const cr_grid *grid, double *cells_xy);
/* Tripolar grid with a RightCenterFolded top row from its (nx+1) x (ny+1)
 * point matrix (one block, layout of cr_grid_from_points): treeify of
 * ZRegOrthogonalSphericalShellGrid{..., RightCenterFolded}
 * (ext/ConservativeRegriddingOceananigansExt.jl:119-187).  Fold-row cells i
 * and nx-1-i are one physical cell: only i < nx/4 and nx/2 <= i < nx/2+nx/4
 * keep their polygon, the rest are zero-area ghosts that intersect nothing
 * (PaddedTreeWrapper, :65-115); (primary, partner) pairs are kept for
 * cr_mirror_fold_partners.  nx must be even (the reference's assertion). */
cr_status cr_grid_tripolar(cr_grid **grid, const double *pts, int64_t nx, int64_t ny, void *stream);
/* Number of (primary, partner) fold pairs of a grid (0 unless folded). */
cr_status cr_grid_fold_count(const cr_grid *grid, int64_t *npairs);
/* Generate a HEALPix grid on the device (HealpixRootNode{O}(nside)). */
cr_status cr_grid_healpix(cr_grid **grid, int64_t nside, int32_t order, int32_t step,
                          void *stream);
cr_status cr_grid_info(const cr_grid *grid, int64_t *ncells, int32_t *nv, int32_t *convex);
/* Number of caller-visible vertices: points of a point grid ((ni+1)(nj+1) per
 * block), ncells*nv for a polygon table or a HEALPix grid.  Vertex gradients
 * (cr_regrid_backward) are indexed by these. */
cr_status cr_grid_vertex_count(const cr_grid *grid, int64_t *nverts);
/* Download the (re-oriented) polygon table, ncells*nv*3 doubles (for a planar
 * grid: the unit-vector images its candidate search uses). */
cr_status cr_grid_copy_cells(const cr_grid *grid, double *cells);
/* Cell areas on the unit sphere, or in the plane for a planar grid, computed on
 * the device (areas(), regridder.jl:165-178). */
cr_status cr_grid_areas(const cr_grid *grid, double *areas_host);
void cr_grid_destroy(cr_grid *grid);

/* ------------------------------------------------------------- build -- */
void cr_build_opts_default(cr_build_opts *opts);
/* Regridder(dst, src; kwargs...) on the device: candidate search, FP64 clip and
 * area per pair, CSR(A), CSR(A^T), cell areas, optional normalize!. */
cr_status cr_regridder_build(cr_regridder **R, const cr_grid *dst, const cr_grid *src,
                             const cr_build_opts *opts, void *stream);
/* Regridder(intersections, dst_areas, src_areas, ...) from host CSR arrays
 * (the struct's field constructor, regridder.jl:25-36). */
cr_status cr_regridder_from_csr(cr_regridder **R, int64_t n_dst, int64_t n_src, int64_t nnz,
                                const int64_t *rowptr, const int32_t *col, const double *val,
                                const double *dst_areas, const double *src_areas, void *stream);
cr_status cr_regridder_info(const cr_regridder *R, int64_t *n_dst, int64_t *n_src, int64_t *nnz);
cr_status cr_regridder_stats(const cr_regridder *R, cr_build_stats *stats);
/* Host copies of CSR(A) (transpose=0) or CSR(A^T) (transpose=1). */
cr_status cr_regridder_copy_csr(const cr_regridder *R, int32_t transpose, int64_t *rowptr,
                                int32_t *col, double *val);
cr_status cr_regridder_copy_areas(const cr_regridder *R, double *dst_areas, double *src_areas);
/* normalize!(R): A, a_src, a_dst ./= maximum(A) (regridder.jl:54-62). */
cr_status cr_regridder_normalize(cr_regridder *R, void *stream);
void cr_regridder_destroy(cr_regridder *R);

/* ------------------------------------------------------------- apply -- */
/* Fold partners after an apply into a folded grid (the destination grid of
 * this direction): dst[partner] = dst[primary] for every field, the
 * finalize_regridding! step for fields on RightCenterFolded grids
 * (mirror_fold_partners!, ext/ConservativeRegriddingOceananigansExt.jl:207-240).
 * A no-op for grids without folds.  Layout arguments as cr_regrid. */
cr_status cr_mirror_fold_partners(const cr_regridder *R, int32_t transpose, int32_t dtype, int64_t nfields,
                                  void *dst, int64_t dst_cell_stride, int64_t dst_field_stride, void *stream);
/* Its adjoint on an FP64 gradient, in place: g[primary] += g[partner],
 * g[partner] = 0 (no reference counterpart: the pullback of the mirror, for
 * the rrule of regrid into a folded grid). */
cr_status cr_mirror_fold_partners_adjoint(const cr_regridder *R, int32_t transpose, int64_t nfields, double *g,
                                          int64_t cell_stride, int64_t field_stride, void *stream);
/*
 * regrid!(dst, R, src) for nfields fields at once (transpose=1: transpose(R)).
 * src and dst are DEVICE pointers of dtype elements.  Element (cell c, field f)
 * is at ptr[c*cell_stride + f*field_stride].  cells-fastest ([n,64,4] with
 * dims=1) is cell_stride=1, field_stride=n; fields-fastest is cell_stride=k,
 * field_stride=1.  normalize=1 divides by the destination areas
 * (finalize_regridding!, regrid.jl:104-118).  CR_F64 fields: products and
 * sums in FP64 (fused multiply-adds), a true division by the area.  CR_F32
 * fields: FP32 in and out; the default kernel multiplies each FP32 value by A
 * (see cr_regrid_set_variant for whether A is used in FP64 or rounded to FP32)
 * and sums in FP64, adds a row's partial sums in FP64, scales by the FP64
 * reciprocal area and rounds once; variants 5-7 and 18-19 instead sum in FP32
 * with Kahan compensation.  Both directions are gathers with a fixed summation
 * order (no atomics): for a fixed kernel variant the results are bitwise
 * reproducible (cr_regrid_tune may pick another variant, whose last bits can
 * differ).  The first call of a direction makes its per-regridder set-up (see
 * cr_regrid_prepare), which allocates and synchronises.
 */
cr_status cr_regrid(const cr_regridder *R, int32_t transpose, int32_t dtype, int64_t nfields,
                    const void *src, int64_t src_cell_stride, int64_t src_field_stride,
                    void *dst, int64_t dst_cell_stride, int64_t dst_field_stride,
                    int32_t normalize, void *stream);

/*
 * Make the per-direction set-up that cr_regrid makes on its first call with
 * this dtype, field count and field strides (the default FP32 kernel of the
 * direction, its tile plan or the long-row list), without applying anything
 * (no reference counterpart: Julia's mul! has no set-up).  After it, such
 * cr_regrid calls allocate nothing and never synchronise, so they can be
 * captured into a HIP graph; a cr_regrid that would still need set-up during a
 * capture fails with CR_UNSUPPORTED instead.  Synchronises the stream.
 */
cr_status cr_regrid_prepare(const cr_regridder *R, int32_t transpose, int32_t dtype, int64_t nfields,
                            int64_t src_field_stride, int64_t dst_field_stride, void *stream);

/*
 * Reverse mode of regrid!(y, R, x) (transpose=1: of regrid!(y, transpose(R), x)),
 * the kernel behind a ChainRulesCore.rrule for regrid! (no reference
 * counterpart: the reference has no AD).  FP64 device arrays, strides as in
 * cr_regrid.  Given the forward input x, output y and the gradient gy of a loss
 * w.r.t. y:
 *   gx (overwritten, may be NULL)  = dL/dx = A^T (gy ./ a_to)
 *   g_dst_verts, g_src_verts (3 doubles per caller vertex of the grids, see
 *   cr_grid_vertex_count; ACCUMULATED with FP64 atomics, may be NULL) = dL/dV
 *   through every intersection area and the destination cell areas.
 * dst and src must be the grids R was built from (vertex gradients are not
 * available for cr_regridder_from_csr regridders, for planar grids, or when
 * the clipping side's cells have more than 8 vertices: CR_UNSUPPORTED).  y is
 * needed when normalize=1.
 */
cr_status cr_regrid_backward(const cr_regridder *R, const cr_grid *dst, const cr_grid *src,
                             int32_t transpose, int32_t normalize, int64_t nfields, const double *x,
                             int64_t xc, int64_t xf, const double *y, int64_t yc, int64_t yf,
                             const double *gy, int64_t gc, int64_t gf, double *gx, int64_t gxc,
                             int64_t gxf, double *g_dst_verts, double *g_src_verts, void *stream);

/*
 * cr_regrid_backward without allocating (no reference counterpart: for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * loops, harnesses that forbid hipMalloc).  Same arguments and results; its
 * scratch (n_to * nfields doubles for gx; for vertex gradients 16 bytes per
 * non-zero and 256 more (a coefficient, a row and a list slot per non-zero,
 * and the list's count); each part rounded up to 256 bytes; n_to = the output
 * cell count of the direction) lives in one static device array of 64 MiB per
 * GPU, kept until
 * the process ends.  Larger requests fail with CR_INVALID_ARGUMENT.  When gx is
 * wanted, the field adjoint's plan must exist: call
 * cr_regrid_prepare(R, !transpose, CR_F64, nfields, n_to, gx's field stride)
 * first, else CR_UNSUPPORTED.  Does not synchronise the stream.  Calls on one
 * GPU share the array: calls that overlap (other streams, other threads) race.
 */
cr_status cr_regrid_backward_noalloc(const cr_regridder *R, const cr_grid *dst, const cr_grid *src,
                                     int32_t transpose, int32_t normalize, int64_t nfields, const double *x,
                                     int64_t xc, int64_t xf, const double *y, int64_t yc, int64_t yf,
                                     const double *gy, int64_t gc, int64_t gf, double *gx, int64_t gxc,
                                     int64_t gxf, double *g_dst_verts, double *g_src_verts, void *stream);

/*
 * Autotuner for the FP32 cells-fastest batched apply (no reference
 * counterpart: the reference's slice loop has no tunable).  Times every kernel
 * variant with the default's numerics (products and sums in FP64 from the FP64
 * A: variants 8-17, 25 and 28) on the given device fields (dst is overwritten
 * with the regridded result) and stores the fastest in the handle for that
 * direction; later cr_regrid calls of that direction use it.  Variants that
 * round A to FP32 or sum in FP32 are used only when pinned
 * (cr_regrid_set_variant).  Synchronises the stream.
 */
cr_status cr_regrid_tune(cr_regridder *R, int32_t transpose, int64_t nfields, const void *src,
                         int64_t src_cell_stride, int64_t src_field_stride, void *dst,
                         int64_t dst_cell_stride, int64_t dst_field_stride, int32_t *chosen,
                         float *best_ms, void *stream);

/*
 * Apply kernel variants (no reference counterpart).  cr_regrid_variant_count
 * gives the number of variants for a dtype; cr_regrid_set_variant pins the
 * variant cr_regrid uses for one direction of R (-1 restores the default).
 *   CR_F32: 0-4 row kernel with 1, 2, 4, 8, 16 fields per thread; 5-7 staged
 *           tile kernel with 2, 4, 8 fields per LDS stage and FP32 Kahan sums;
 *           8-10 the tile kernel with 2, 4, 8 fields per stage and FP64 sums;
 *           11-14 variant 9 with 4 field chunks, 8 and 16 chunks
 *           swept one at a time, and a single chunk; 15-17 FP64 sums of the
 *           FP64 A on 1024-row tiles (512 threads) with 4, 2 and 4 fields per
 *           stage (17: 4 field chunks); 18-19 FP32 Kahan sums on 1024-row tiles
 *           with 4 fields per stage (19: 4 field chunks); 20-21 FP64 sums of A
 *           rounded to FP32 on 1024-row tiles, 4 fields per stage (21: 4 field
 *           chunks); 22-23 variants 20-21 on a plan whose rows are 32-row
 *           blocks sorted by their smallest column (for rows that are strips
 *           of a row-major grid over a source in another order); 24-26
 *           variants 21, 17 and 21 with 4, 4 and 2 field chunks swept one at a
 *           time; 27 variant 23 swept likewise; 28 variant 15 on the
 *           block-ordered plan.  Variants without a fixed chunk count take as
 *           many chunks as keep 8192 workgroups busy, fewer where re-reading a
 *           tile's slots per chunk would pass a tenth of the launch's field
 *           bytes.  The default is 15 (A used in FP64), or 28 when 15's tiles
 *           fetch over 3 source lines per 32 distinct source cells and 28's
 *           fetch fewer (decided once per direction at its first default
 *           apply).  FP32 fields with unit field strides and >= 64 fields
 *           always use the fields-fastest kernel.
 *   CR_F64: 0 row kernel, 1 staged tile kernel (default).
 */
int32_t cr_regrid_variant_count(int32_t dtype);
cr_status cr_regrid_set_variant(cr_regridder *R, int32_t transpose, int32_t dtype, int32_t variant);

/*
 * Tile plans of the staged apply (no reference counterpart).  The tile kernel
 * runs on a plan of one direction of R, in layout style & 1 (0: rounds, used by
 * FP32 fields; 1: even runs, used by FP64 fields), over the matrix rows in
 * order (style & 2 == 0) or over 32-row blocks sorted by their smallest column
 * (style & 2; FP32 variants 22-23), and geometry `geom` (0: 256 threads, 512
 * rows, 1024 columns per tile; 1: 512, 1024, 2048).  The first cr_regrid that
 * needs a plan makes it on the device; cr_regrid_tile_plan makes it ahead of
 * time (for instance before capturing cr_regrid in a HIP graph) and gives its
 * sizes: sizes[0] tiles, [1] slots, [2] column-list entries, [3] segment-map
 * entries, [4] rows left to the long-row kernel, [5] row-map entries (block
 * order only).  Synchronises the stream.  cr_regrid_copy_tile_plan copies a
 * made plan to host memory: descriptors (72 bytes per tile: row0, ent_off,
 * col_off as int64; nrows, ncols, ent, nseg as int32; 8 uint32 flush masks),
 * slot keys (local column | segment << 16), slot values as FP32 and FP64,
 * column lists, segment maps, long rows (matrix rows) and the row map (plan
 * row -> matrix row, -1 for none), each buffer holding the corresponding
 * sizes[] entries.
 */
cr_status cr_regrid_tile_plan(const cr_regridder *R, int32_t transpose, int32_t style, int32_t geom,
                              int64_t *sizes, void *stream);
cr_status cr_regrid_copy_tile_plan(const cr_regridder *R, int32_t transpose, int32_t style, int32_t geom,
                                   void *desc, uint32_t *key, float *v32, double *v64, int32_t *cols,
                                   uint16_t *rseg, int32_t *huge, int32_t *rowmap);

/* ------------------------------------------------- ESMF weights files -- */
/*
 * save_esmf_weights(path, regridder; src_grid_name, dst_grid_name, src_shape,
 * dst_shape, created_at) of ext/ConservativeRegriddingNCDatasetsExt.jl:15-59:
 * an ESMF offline-weights file with dimensions n_s, n_a, n_b, variables S =
 * A ./ dst_areas[row], row, col (1-based, column-major order of A), frac_a,
 * frac_b, area_a, area_b and the same global attributes.  Written as netCDF
 * classic: cdf_version 0 or 5 = CDF-5 (64-bit data, the default), with row,
 * col and the grid shapes as 64-bit integers like the reference's Int64
 * (findnz of a SparseMatrixCSC{Float64,Int64}); 2 = CDF-2 (64-bit offsets,
 * 32-bit row/col, for readers without CDF-5).  Strings and shapes in
 * cr_esmf_attrs may be NULL / 0 (defaults "source", "destination"; created_at
 * and the shapes are then omitted, as in the reference).
 */
typedef struct {
    const char *src_grid_name;
    const char *dst_grid_name;
    const char *created_at;
    const int64_t *src_shape;
    int32_t src_ndim;
    int32_t dst_ndim;
    const int64_t *dst_shape;
    int32_t cdf_version; /* 0 or 5: CDF-5; 2: CDF-2 */
    int32_t reserved;
} cr_esmf_attrs;
/* The file of this direction of R (transpose=1: transpose(R)); host work only
 * (synchronous copies of A and the areas). */
cr_status cr_regridder_write_esmf(const cr_regridder *R, int32_t transpose, const char *path,
                                  const cr_esmf_attrs *attrs);
/* The same file from host arrays (S, row, col of length n_s; frac_a, area_a of
 * n_a; frac_b, area_b of n_b); needs no GPU. */
cr_status cr_esmf_write(const char *path, int64_t n_s, const double *S, const int64_t *row,
                        const int64_t *col, int64_t n_a, const double *frac_a, const double *area_a,
                        int64_t n_b, const double *frac_b, const double *area_b,
                        const cr_esmf_attrs *attrs);
/*
 * Reading a weights file back (no reference counterpart: the reference only
 * writes; reading lets a build be reused).  Any netCDF classic file (CDF-1,
 * CDF-2 or CDF-5) with the dimensions n_s, n_a, n_b and 1-D variables of those
 * names is accepted, integer or floating types converted.  cr_esmf_read_sizes
 * gives the dimensions; cr_esmf_read fills the host arrays that are not NULL
 * (S, row, col: n_s; frac_a, area_a: n_a; frac_b, area_b: n_b); both need no
 * GPU.  cr_regridder_from_esmf makes a device regridder of it: A_ij =
 * S_k * area_b[i] with i = row_k - 1, j = col_k - 1, dst_areas = area_b,
 * src_areas = area_a (as cr_regridder_from_csr: no vertex gradients).
 */
cr_status cr_esmf_read_sizes(const char *path, int64_t *n_s, int64_t *n_a, int64_t *n_b);
cr_status cr_esmf_read(const char *path, double *S, int64_t *row, int64_t *col, double *frac_a,
                       double *area_a, double *frac_b, double *area_b);
cr_status cr_regridder_from_esmf(cr_regridder **R, const char *path, void *stream);

/* ---------------------------------------------- spectral elements -- */
/* The ClimaCore extension's regridders (ext/ConservativeRegriddingClimaCoreExt.jl)
 * for spectral-element spaces given by their elements: a grid of
 * quadrilaterals whose vertices c1..c4 (ring order, counter-clockwise) are the
 * corners (-1,-1), (1,-1), (1,1), (-1,1) of the NormalizedBilinearMap
 * x(xi, eta) = p / |p|, p the bilinear interpolation of the corners, and nq
 * Gauss-Lobatto-Legendre nodes per direction (2 <= nq <= 8).  Nodes are
 * numbered e*nq^2 + i + j*nq (element-major, i fastest: se_node_positions,
 * :136-145).  tri_degree 0 selects 2*(nq-1) (the reference's default). */

/* reference_rule(degree) (src/regridder/triangle_quadrature.jl:139-156): the
 * Dunavant rule of that degree or the next one up (max 8); bary 3*16 and w 16
 * doubles. */
cr_status cr_se_tri_rule(int32_t degree, double *bary, double *w, int32_t *npts);
/* Gauss-Lobatto-Legendre nodes (ascending) and weights, 2 <= n <= 64. */
cr_status cr_se_gll(int32_t n, double *x, double *w);
/* Node positions (3 per node) and weights W = w_i w_j J(xi_i, eta_j) radius^2
 * (se_node_positions / se_node_weights, :136-158), host outputs, either may be
 * NULL. */
cr_status cr_se_nodes(const cr_grid *elements, int32_t nq, double radius, double *xyz, double *weights);
/* accumulate_principled_b (:248-298) for every non-zero of R (built from
 * cells as destination and elements as source): B[k*nq^2 + i + j*nq] = the
 * integral over (element ∩ cell) of phi_i(xi) phi_j(eta), in R's area units;
 * B is a DEVICE array of nnz*nq^2 doubles in R's device CSR order, whose
 * (cell, element) pairs go to the optional DEVICE arrays pair_cell and
 * pair_elem (nnz int32 each). */
cr_status cr_se_moments(const cr_regridder *R, const cr_grid *cells, const cr_grid *elements, int32_t nq,
                        int32_t tri_degree, double *B, int32_t *pair_cell, int32_t *pair_elem, void *stream);
/* Regridder(cells, se_space): se_to_fv_principled (:424-478).  Rows = cells,
 * columns = nodes, values = the moments; dst areas = cell areas, src areas =
 * node weights (node_weights, host, nelements*nq^2, or NULL for the map's
 * own).  opts: radius and threshold_rel of the pair search (normalize ignored). */
cr_status cr_se_regridder_to_fv(cr_regridder **R, const cr_grid *cells, const cr_grid *elements, int32_t nq,
                                int32_t tri_degree, const double *node_weights, const cr_build_opts *opts,
                                void *stream);
/* Regridder(se_space, cells): fv_to_se_l2_projection (:562-676).  Rows =
 * nodes, columns = cells; per element the covered part of the local mass
 * matrix (rows rescaled to the moments' sums) is inverted into the weights;
 * dst areas = 1, src areas = cell areas.  The per-element solves run on host
 * threads. */
cr_status cr_se_regridder_from_fv(cr_regridder **R, const cr_grid *elements, const cr_grid *cells, int32_t nq,
                                  int32_t tri_degree, const double *node_weights, const cr_build_opts *opts,
                                  void *stream);

/* ------------------------------------------- example-field cell averages -- */
/* The reference's example fields (src/utils/example_data.jl), as functions of
 * (lon, lat) in degrees. */
typedef enum {
    CR_FIELD_LONGITUDE = 0,   /* LongitudeField: lon */
    CR_FIELD_SINUSOID = 1,    /* SinusoidField(; dp_length, coef, coefmult) */
    CR_FIELD_HARMONIC = 2,    /* HarmonicField: 2 + sin(2 lat)^16 cos(16 lon) */
    CR_FIELD_GULF_STREAM = 3, /* GulfStreamField(; dp_length, coef, gf_coef, gf_ori_lon, gf_ori_lat,
                                 gf_end_lon, gf_end_lat, gf_dmp_lon, gf_dmp_lat) */
    CR_FIELD_VORTEX = 4       /* VortexField(; lon0_rad, lat0_rad, r0, d, t) */
} cr_example_field;

/* Cell averages int f dOmega / int dOmega of an example field over every cell
 * of a grid, into the DEVICE array out (ncells doubles): the test helpers'
 * set_field_values! with SphericalPolygonIntegrator (fan triangles from the
 * first vertex, x = p/|p| over each, the symmetric rule of `degree`, 8 by the
 * reference's default;
 * lib/ConservativeRegriddingTestHelpers/src/ConservativeRegriddingTestHelpers.jl:84-128).
 * lon = atan2(y, x), lat = asin(z) in degrees.  params: NULL for the
 * reference's defaults, else nparams values in the keyword order above (3, 9
 * or 5; none for the longitude and harmonic fields).  Zero-area cells get 0. */
cr_status cr_cell_averages(const cr_grid *grid, int32_t field, const double *params, int32_t nparams,
                           int32_t degree, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CRHIP_H */
