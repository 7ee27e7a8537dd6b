"""ctypes front-end of the CPU oracle (oracle/cr_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package.

The functions mirror the reference's CPU path (see the header of cr_oracle.c
for the file:line map):

* :func:`build`      -- ``Regridder(dst, src)`` (regridder.jl:125-163)
* :func:`spmv`       -- ``regrid!`` single field (regrid.jl:63-118)
* :func:`transpose`  -- the CSR of ``transpose(R).intersections`` (regridder.jl:49-50)
* :func:`clip_area`, :func:`polygon_area`, :func:`cell_areas` -- the per-pair
  operator (regridder.jl:96-103) and ``areas`` (regridder.jl:165-178)
* :func:`build_planar`, :func:`planar_clip_area`, :func:`planar_polygon_area`
  -- the same on the Planar manifold (regridder.jl:87-94; cr_oracle_planar.c)
* :func:`tile_plan`  -- the library's own staged-apply layout (cr_tileplan.cpp;
  no reference counterpart), the checker of the device planner
* :func:`dunavant`, :func:`gll`, :func:`lagrange`, :func:`element_map`,
  :func:`inverse_element_map`, :func:`element_jacobian`, :func:`se_moments` --
  the spectral-element weights (cr_oracle_se.c; triangle_quadrature.jl and
  ext/ConservativeRegriddingClimaCoreExt.jl)
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build_library() -> str:
    """Compile oracle/build/liboracle.so with gcc (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build_library()
    L = ctypes.CDLL(_LIB_PATH)
    L.orc_polygon_area.restype = ctypes.c_double
    L.orc_polygon_area.argtypes = [_dp, ctypes.c_int32]
    L.orc_clip_area.restype = ctypes.c_double
    L.orc_clip_area.argtypes = [_dp, ctypes.c_int32, _dp, ctypes.c_int32]
    L.orc_clip_polygon.restype = ctypes.c_int32
    L.orc_clip_polygon.argtypes = [_dp, ctypes.c_int32, _dp, ctypes.c_int32, _dp, ctypes.c_int32]
    L.orc_cell_areas.restype = None
    L.orc_cell_areas.argtypes = [_dp, ctypes.c_int64, ctypes.c_int32, _dp, ctypes.c_int32]
    L.orc_build.restype = ctypes.c_void_p
    L.orc_build.argtypes = [_dp, ctypes.c_int64, ctypes.c_int32, _dp, ctypes.c_int64, ctypes.c_int32,
                            ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64]
    for name in ("orc_result_nnz", "orc_result_candidates", "orc_result_clip_flops", "orc_result_clip_flops_kept"):
        getattr(L, name).restype = ctypes.c_int64
        getattr(L, name).argtypes = [ctypes.c_void_p]
    L.orc_set_flop_counting.restype = None
    L.orc_set_flop_counting.argtypes = [ctypes.c_int32]
    L.orc_clip_area_flops.restype = ctypes.c_int64
    L.orc_clip_area_flops.argtypes = [_dp, ctypes.c_int32, _dp, ctypes.c_int32]
    L.orc_result_status.restype = ctypes.c_int32
    L.orc_result_status.argtypes = [ctypes.c_void_p]
    L.orc_result_times.restype = None
    L.orc_result_times.argtypes = [ctypes.c_void_p, _dp]
    L.orc_result_copy.restype = None
    L.orc_result_copy.argtypes = [ctypes.c_void_p, _i64p, _i32p, _dp, _dp, _dp]
    L.orc_result_free.restype = None
    L.orc_result_free.argtypes = [ctypes.c_void_p]
    L.orc_spmv.restype = None
    L.orc_spmv.argtypes = [_i64p, _i32p, _dp, ctypes.c_int64, _dp, _dp, _dp, ctypes.c_int32]
    L.orc_polygon_area_grad.restype = ctypes.c_double
    L.orc_polygon_area_grad.argtypes = [_dp, ctypes.c_int32, _dp]
    L.orc_clip_area_grad.restype = ctypes.c_double
    L.orc_clip_area_grad.argtypes = [_dp, ctypes.c_int32, _dp, ctypes.c_int32, _dp, _dp]
    L.orc_dunavant.restype = ctypes.c_int32
    L.orc_dunavant.argtypes = [ctypes.c_int32, _dp, _dp]
    L.orc_gll.restype = None
    L.orc_gll.argtypes = [ctypes.c_int32, _dp, _dp]
    L.orc_lagrange.restype = None
    L.orc_lagrange.argtypes = [ctypes.c_int32, _dp, ctypes.c_double, _dp]
    L.orc_element_map.restype = None
    L.orc_element_map.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _dp]
    L.orc_element_jacobian.restype = ctypes.c_double
    L.orc_element_jacobian.argtypes = [_dp, ctypes.c_double, ctypes.c_double]
    L.orc_inverse_element_map.restype = ctypes.c_int32
    L.orc_inverse_element_map.argtypes = [_dp, _dp, _dp, _dp]
    L.orc_se_moments.restype = ctypes.c_int32
    L.orc_se_moments.argtypes = [_dp, _dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp]
    L.orc_transpose.restype = None
    L.orc_transpose.argtypes = [_i64p, _i32p, _dp, ctypes.c_int64, ctypes.c_int64, _i64p, _i32p, _dp]
    L.oracle_tile_plan_make.restype = ctypes.c_void_p
    L.oracle_tile_plan_make.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                        _i64p, _i32p, _dp, _i64p]
    L.oracle_tile_plan_copy.restype = None
    L.oracle_tile_plan_copy.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 7
    L.oracle_tile_plan_free.restype = None
    L.oracle_tile_plan_free.argtypes = [ctypes.c_void_p]
    L.orc_planar_polygon_area.restype = ctypes.c_double
    L.orc_planar_polygon_area.argtypes = [_dp, ctypes.c_int32]
    L.orc_planar_clip_area.restype = ctypes.c_double
    L.orc_planar_clip_area.argtypes = [_dp, ctypes.c_int32, _dp, ctypes.c_int32]
    L.orc_planar_build.restype = ctypes.c_void_p
    L.orc_planar_build.argtypes = [_dp, ctypes.c_int64, ctypes.c_int32, _dp, ctypes.c_int64, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_double]
    for name in ("orc_planar_nnz", "orc_planar_candidates"):
        getattr(L, name).restype = ctypes.c_int64
        getattr(L, name).argtypes = [ctypes.c_void_p]
    L.orc_planar_status.restype = ctypes.c_int32
    L.orc_planar_status.argtypes = [ctypes.c_void_p]
    L.orc_planar_copy.restype = None
    L.orc_planar_copy.argtypes = [ctypes.c_void_p, _i64p, _i32p, _dp, _dp, _dp]
    L.orc_planar_free.restype = None
    L.orc_planar_free.argtypes = [ctypes.c_void_p]
    _lib = L
    return L


def default_threads() -> int:
    """Worker threads: the host's cores, capped at 16 (one GPU's share of a box)."""
    return max(1, min(16, os.cpu_count() or 1))


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def polygon_area(ring) -> float:
    """Signed solid angle of an open ring of unit vectors, shape (n, 3)."""
    r = _f64(ring)
    return lib().orc_polygon_area(_d(r), r.shape[0])


def clip_area_flops(subject, clip) -> int:
    """FP64 operations clip_area(subject, clip) performs (adds, subs, muls,
    divs, sqrts; atan2 counts as one)."""
    S, C = _f64(subject), _f64(clip)
    return int(lib().orc_clip_area_flops(_d(S), S.shape[0], _d(C), C.shape[0]))


def clip_area(subject, clip) -> float:
    """Area of subject ∩ clip (clip convex, both CCW open rings of unit vectors)."""
    s, c = _f64(subject), _f64(clip)
    return lib().orc_clip_area(_d(s), s.shape[0], _d(c), c.shape[0])


def polygon_area_grad(ring):
    """(area, d area / d vertices) of area(normalize(ring)), forward-mode duals."""
    r = _f64(ring)
    g = np.empty_like(r)
    a = lib().orc_polygon_area_grad(_d(r), r.shape[0], _d(g))
    return a, g


def clip_area_grad(subject, clip):
    """(area, d/d subject, d/d clip) of area(normalize(subject) ∩ normalize(clip))."""
    s, c = _f64(subject), _f64(clip)
    gs, gc = np.empty_like(s), np.empty_like(c)
    a = lib().orc_clip_area_grad(_d(s), s.shape[0], _d(c), c.shape[0], _d(gs), _d(gc))
    return a, gs, gc


def clip_polygon(subject, clip) -> np.ndarray:
    s, c = _f64(subject), _f64(clip)
    out = np.zeros((256, 3))
    m = lib().orc_clip_polygon(_d(s), s.shape[0], _d(c), c.shape[0], _d(out), 256)
    if m < 0:
        raise RuntimeError("clip buffer overflow")
    return out[:m].copy()


def dunavant(degree: int):
    """``reference_rule(degree)``: (barycentric points (n, 3), weights (n,))."""
    bary, w = np.empty((16, 3)), np.empty(16)
    n = lib().orc_dunavant(degree, _d(bary), _d(w))
    if n < 0:
        raise ValueError(f"no triangle rule of degree >= {degree}")
    return bary[:n].copy(), w[:n].copy()


def gll(n: int):
    """Gauss-Lobatto-Legendre nodes (ascending) and weights, n points."""
    x, w = np.empty(n), np.empty(n)
    lib().orc_gll(n, _d(x), _d(w))
    return x, w


def lagrange(nodes, t) -> np.ndarray:
    x = _f64(nodes)
    out = np.empty(x.shape[0])
    lib().orc_lagrange(x.shape[0], _d(x), float(t), _d(out))
    return out


def element_map(corners, xi, eta) -> np.ndarray:
    c = _f64(corners)
    out = np.empty(3)
    lib().orc_element_map(_d(c), float(xi), float(eta), _d(out))
    return out


def element_jacobian(corners, xi, eta) -> float:
    c = _f64(corners)
    return lib().orc_element_jacobian(_d(c), float(xi), float(eta))


def inverse_element_map(corners, x):
    c, p = _f64(corners), _f64(x)
    xi, eta = ctypes.c_double(), ctypes.c_double()
    if lib().orc_inverse_element_map(_d(c), _d(p), ctypes.byref(xi), ctypes.byref(eta)) != 0:
        raise ValueError("point outside the element")
    return xi.value, eta.value


def se_moments(element, cell, nq: int, degree: int | None = None) -> np.ndarray:
    """B[i, j] = integral over (element ∩ cell) of phi_i(xi) phi_j(eta) on the unit
    sphere (``accumulate_principled_b``), shape (nq, nq) indexed [i, j]."""
    e, c = _f64(element), _f64(cell)
    nodes, _ = gll(nq)
    B = np.empty(nq * nq)
    m = lib().orc_se_moments(_d(e), _d(c), c.shape[0], nq, 2 * (nq - 1) if degree is None else degree,
                             _d(nodes), _d(B))
    if m < 0:
        raise RuntimeError("se_moments failed (clip overflow or a point outside the element)")
    return B.reshape(nq, nq).T.copy()


def cell_areas(cells, nthreads: int = 0) -> np.ndarray:
    """Areas of a polygon table, shape (ncells, nv, 3)."""
    c = _f64(cells)
    out = np.empty(c.shape[0])
    lib().orc_cell_areas(_d(c), c.shape[0], c.shape[1], _d(out), nthreads or default_threads())
    return out


@dataclass
class OracleCSR:
    n_dst: int
    n_src: int
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    a_dst: np.ndarray
    a_src: np.ndarray
    n_candidates: int
    seconds: float
    threads: int
    clip_flops: int = 0  # FP64 operations of every candidate's clip and area (build(count_flops=True))
    clip_flops_kept: int = 0  # the same over the candidates kept as non-zeros

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def row_index(self) -> np.ndarray:
        return np.repeat(np.arange(self.n_dst, dtype=np.int64), np.diff(self.rowptr))

    def transpose(self) -> "OracleCSR":
        rp, c, v = transpose(self.rowptr, self.col, self.val, self.n_dst, self.n_src)
        return OracleCSR(self.n_src, self.n_dst, rp, c, v, self.a_src, self.a_dst,
                         self.n_candidates, self.seconds, self.threads)

    def apply(self, x, normalize: bool = True, nthreads: int = 1) -> np.ndarray:
        return spmv(self.rowptr, self.col, self.val, x, self.a_dst if normalize else None, nthreads)

    def normalized(self) -> "OracleCSR":
        """normalize!(R): A, a_s, a_d ./= maximum(A)  (regridder.jl:54-62)."""
        m = self.val.max()
        return OracleCSR(self.n_dst, self.n_src, self.rowptr, self.col, self.val / m,
                         self.a_dst / m, self.a_src / m, self.n_candidates, self.seconds, self.threads)


def build(dst_cells, src_cells, clip_is_dst: bool = True, threshold_rel: float = 1e-14,
          nthreads: int = 0, row_range=None, count_flops: bool = False) -> OracleCSR:
    """Regridder build on the unit sphere.  ``dst_cells``/``src_cells`` are
    polygon tables (ncells, nv, 3).  ``clip_is_dst`` picks the convex side.
    ``count_flops``: also count the FP64 operations of every candidate's clip
    and area (``OracleCSR.clip_flops``; slower, not for timing)."""
    d, s = _f64(dst_cells), _f64(src_cells)
    nth = nthreads or default_threads()
    r0, r1 = (0, d.shape[0]) if row_range is None else row_range
    L = lib()
    L.orc_set_flop_counting(1 if count_flops else 0)
    try:
        h = L.orc_build(_d(d), d.shape[0], d.shape[1], _d(s), s.shape[0], s.shape[1],
                        1 if clip_is_dst else 0, threshold_rel, nth, r0, r1)
    finally:
        L.orc_set_flop_counting(0)
    try:
        st = L.orc_result_status(h)
        if st != 0:
            raise RuntimeError(f"oracle build failed (status {st})")
        nnz = L.orc_result_nnz(h)
        rowptr = np.empty(d.shape[0] + 1, np.int64)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        a_dst = np.empty(d.shape[0])
        a_src = np.empty(s.shape[0])
        L.orc_result_copy(h, rowptr.ctypes.data_as(_i64p), col.ctypes.data_as(_i32p), _d(val),
                          _d(a_dst), _d(a_src))
        t = np.zeros(3)
        L.orc_result_times(h, _d(t))
        ncand = L.orc_result_candidates(h)
        flops = L.orc_result_clip_flops(h)
        flops_kept = L.orc_result_clip_flops_kept(h)
    finally:
        L.orc_result_free(h)
    return OracleCSR(d.shape[0], s.shape[0], rowptr, col, val, a_dst, a_src, ncand, float(t[0]), nth, flops,
                     flops_kept)


def planar_polygon_area(ring) -> float:
    """Shoelace area of an open ring ``(k, 2)`` (signed, counter-clockwise > 0)."""
    r = _f64(ring)
    return lib().orc_planar_polygon_area(_d(r), r.shape[0])


def planar_clip_area(subject, clip) -> float:
    """Planar area of ``subject ∩ clip`` (open rings ``(k, 2)``, clip convex and
    counter-clockwise): the planar operator of regridder.jl:87-94."""
    s, c = _f64(subject), _f64(clip)
    return lib().orc_planar_clip_area(_d(s), s.shape[0], _d(c), c.shape[0])


def build_planar(dst_cells, src_cells, clip_is_dst: bool = True, threshold_rel: float = 1e-14) -> OracleCSR:
    """Regridder build on the Planar manifold from planar polygon tables
    ``(ncells, nv, 2)`` (open counter-clockwise rings)."""
    d, s = _f64(dst_cells), _f64(src_cells)
    L = lib()
    h = L.orc_planar_build(_d(d), d.shape[0], d.shape[1], _d(s), s.shape[0], s.shape[1],
                           1 if clip_is_dst else 0, threshold_rel)
    try:
        st = L.orc_planar_status(h)
        if st != 0:
            raise RuntimeError(f"oracle planar build failed (status {st})")
        nnz = L.orc_planar_nnz(h)
        rowptr = np.empty(d.shape[0] + 1, np.int64)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        a_dst = np.empty(d.shape[0])
        a_src = np.empty(s.shape[0])
        L.orc_planar_copy(h, rowptr.ctypes.data_as(_i64p), col.ctypes.data_as(_i32p), _d(val), _d(a_dst),
                          _d(a_src))
        ncand = L.orc_planar_candidates(h)
    finally:
        L.orc_planar_free(h)
    return OracleCSR(d.shape[0], s.shape[0], rowptr, col, val, a_dst, a_src, ncand, 0.0, 1)


def spmv(rowptr, col, val, x, area=None, nthreads: int = 1) -> np.ndarray:
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    val, x = _f64(val), _f64(x)
    n = rowptr.shape[0] - 1
    y = np.empty(n)
    a = _f64(area) if area is not None else None
    lib().orc_spmv(rowptr.ctypes.data_as(_i64p), col.ctypes.data_as(_i32p), _d(val), n, _d(x), _d(y),
                   _d(a) if a is not None else None, nthreads)
    return y


def transpose(rowptr, col, val, nrows: int, ncols: int):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    val = _f64(val)
    trp = np.empty(ncols + 1, np.int64)
    tcol = np.empty(col.shape[0], np.int32)
    tval = np.empty(col.shape[0])
    lib().orc_transpose(rowptr.ctypes.data_as(_i64p), col.ctypes.data_as(_i32p), _d(val), nrows, ncols,
                        trp.ctypes.data_as(_i64p), tcol.ctypes.data_as(_i32p), _d(tval))
    return trp, tcol, tval


# ------------------------------------------------------ spectral elements --
# Tile descriptor of a plan (72 bytes, the library's layout).
TILE_DESC = np.dtype([("row0", "<i8"), ("ent_off", "<i8"), ("col_off", "<i8"), ("nrows", "<i4"),
                      ("ncols", "<i4"), ("ent", "<i4"), ("nseg", "<i4"), ("flush", "<u4", (8,))])
# Tile geometries: threads, rows, columns, segments per tile.
TILE_GEOMS = ((256, 512, 1024, 1024), (512, 1024, 2048, 1536))


def empty_tile_plan(sizes) -> dict:
    """Host arrays for a plan of the given sizes (tiles, slots, column-list
    entries, row-map entries, huge rows)."""
    nt, ns, nc, nr, nh = (int(x) for x in sizes)
    return {"desc": np.zeros(nt, TILE_DESC), "key": np.zeros(ns, np.uint32), "v32": np.zeros(ns, np.float32),
            "v64": np.zeros(ns, np.float64), "cols": np.zeros(nc, np.int32), "rseg": np.zeros(nr, np.uint16),
            "huge": np.zeros(nh, np.int32)}


def tile_plan(style: int, geom: int, rowptr, col, val, ncols: int) -> dict:
    """The staged-apply plan of a CSR matrix (style 0: rounds, 1: even runs;
    geom 0/1: TILE_GEOMS) as host arrays."""
    L = lib()
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    c = np.ascontiguousarray(col, dtype=np.int32)
    v = np.ascontiguousarray(val, dtype=np.float64)
    sizes = np.zeros(5, np.int64)
    h = L.oracle_tile_plan_make(int(style), int(geom), rp.shape[0] - 1, int(ncols), rp.ctypes.data_as(_i64p),
                                c.ctypes.data_as(_i32p), v.ctypes.data_as(_dp), sizes.ctypes.data_as(_i64p))
    try:
        P = empty_tile_plan(sizes)
        L.oracle_tile_plan_copy(h, *(P[k].ctypes.data for k in ("desc", "key", "v32", "v64", "cols", "rseg",
                                                                 "huge")))
    finally:
        L.oracle_tile_plan_free(h)
    return P


def se_node_weights(elements, nq: int) -> np.ndarray:
    """W = w_i w_j J(xi_i, eta_j) on the unit sphere, node order e*nq^2 + i + j*nq."""
    x, w = gll(nq)
    E = _f64(elements)
    out = np.empty((E.shape[0], nq, nq))  # [e, j, i]
    for e in range(E.shape[0]):
        for j in range(nq):
            for i in range(nq):
                out[e, j, i] = w[i] * w[j] * element_jacobian(E[e], x[i], x[j])
    return out.reshape(-1)


def se_to_fv(cells, elements, nq: int, degree: int | None = None, node_weights=None) -> OracleCSR:
    """se_to_fv_principled (ext/ConservativeRegriddingClimaCoreExt.jl:424-478) on
    the unit sphere: rows = cells, columns = nodes e*nq^2 + i + j*nq, values the
    moments (exact zeros dropped); dst areas = cell areas, src areas = W."""
    C, E = _f64(cells), _f64(elements)
    P = build(C, E)
    Q = nq * nq
    rows, cols, vals = [], [], []
    for c in range(P.n_dst):
        for k in range(P.rowptr[c], P.rowptr[c + 1]):
            e = int(P.col[k])
            B = se_moments(E[e], C[c], nq, degree).flatten(order="F")  # index i + j*nq
            nz = np.nonzero(B)[0]
            rows.append(np.full(nz.size, c))
            cols.append(e * Q + nz)
            vals.append(B[nz])
    r, cc, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    order = np.lexsort((cc, r))
    r, cc, v = r[order], cc[order], v[order]
    rp = np.zeros(P.n_dst + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp)
    W = se_node_weights(E, nq) if node_weights is None else _f64(node_weights)
    return OracleCSR(P.n_dst, E.shape[0] * Q, rp, cc.astype(np.int32), v, P.a_dst, W, P.n_candidates, 0.0, 1)


def fv_to_se(elements, cells, nq: int, degree: int | None = None, node_weights=None) -> OracleCSR:
    """fv_to_se_l2_projection (ext/ConservativeRegriddingClimaCoreExt.jl:562-676) on
    the unit sphere: per element, the local mass matrix (Jacobian interpolated
    from W / (w_i w_j), GLL rule of ceil((3nq-2)/2)+1 points), its rows
    rescaled to the element's summed moments, restricted to the covered nodes,
    inverted and applied to each cell's moments; dst areas 1, src areas = cell areas."""
    C, E = _f64(cells), _f64(elements)
    P = build(C, E)
    T = P.transpose()  # rows = elements, cells ascending
    Q = nq * nq
    x, w = gll(nq)
    W = (se_node_weights(E, nq) if node_weights is None else _f64(node_weights)).reshape(-1, nq, nq)  # [e, j, i]
    nm = -(-(3 * nq - 2) // 2) + 1
    xm, wm = gll(nm)
    ph = np.array([lagrange(x, t) for t in xm])  # [p, a]
    rows, cols, vals = [], [], []
    for e in range(E.shape[0]):
        cl = T.col[T.rowptr[e]:T.rowptr[e + 1]]
        if cl.size == 0:
            continue
        Bs = [se_moments(E[e], C[c], nq, degree).flatten(order="F") for c in cl]
        bfull = np.sum(Bs, axis=0)
        cov = np.nonzero(bfull)[0]
        if cov.size == 0:
            continue
        Jn = W[e] / np.outer(w, w)  # [b, a] = W[e, j=b, i=a] / (w_a w_b)
        M = np.zeros((Q, Q))
        for q in range(nm):
            for p in range(nm):
                J = ph[p] @ Jn.T @ ph[q]  # sum_ab Jn[b, a] phi_a(x_p) phi_b(x_q)
                f = np.outer(ph[q], ph[p]).reshape(-1)  # index a + b*nq: phi_a(x_p) phi_b(x_q)
                M += wm[p] * wm[q] * J * np.outer(f, f)
        Mc = M[np.ix_(cov, cov)]
        rs = Mc.sum(axis=1)
        scale = np.where(rs != 0, bfull[cov] / np.where(rs != 0, rs, 1.0), 1.0)
        Mc = Mc * scale[:, None]
        Mi = np.linalg.inv(Mc)
        for c, b in zip(cl, Bs):
            new = Mi @ b[cov]
            nz = new != 0
            rows.append(e * Q + cov[nz])
            cols.append(np.full(int(nz.sum()), c))
            vals.append(new[nz])
    r, cc, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    order = np.lexsort((cc, r))
    r, cc, v = r[order], cc[order], v[order]
    n = E.shape[0] * Q
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp)
    return OracleCSR(n, C.shape[0], rp, cc.astype(np.int32), v, np.ones(n), P.a_dst, P.n_candidates, 0.0, 1)


# ---------------------------------------------------------------- example fields
# Restatement of src/utils/example_data.jl (functions of lon, lat in degrees)
# and of the test helpers' SphericalPolygonIntegrator / set_field_values!
# (lib/ConservativeRegriddingTestHelpers/src/ConservativeRegriddingTestHelpers.jl:84-128).

def _rad(d):
    return np.asarray(d, np.float64) * (np.pi / 180.0)


def oracle_field(name: str, lon, lat, **kw):
    """The example field ``name`` (longitude, sinusoid, harmonic, gulf_stream,
    vortex) at (lon, lat) degrees, with the reference's keyword defaults."""
    lon, lat = np.broadcast_arrays(np.asarray(lon, np.float64), np.asarray(lat, np.float64))
    if name == "longitude":
        return lon.copy()
    if name in ("sinusoid", "gulf_stream"):
        dp = kw.get("dp_length", 1.2 * np.pi)
        coef = kw.get("coef", 2.0)
        base = coef - np.cos(np.pi * np.arccos(np.cos(_rad(lat)) * np.cos(_rad(lon))) / dp)
        if name == "sinusoid":
            return kw.get("coefmult", 1.0) * base
        ori = (kw.get("gf_ori_lon", -80.0), kw.get("gf_ori_lat", 25.0))
        end = (kw.get("gf_end_lon", -1.8), kw.get("gf_end_lat", 50.0))
        dmp = (kw.get("gf_dmp_lon", -25.5), kw.get("gf_dmp_lat", 55.5))
        dr0 = math.sqrt(float(_rad(end[0] - ori[0])) ** 2 + float(_rad(end[1] - ori[1])) ** 2)
        dr1 = math.sqrt(float(_rad(dmp[0] - ori[0])) ** 2 + float(_rad(dmp[1] - ori[1])) ** 2)
        wrapped = lon.copy()
        wrapped[lon > 180.0] -= 360.0
        wrapped[lon < -180.0] += 360.0
        dx, dy = _rad(wrapped - ori[0]), _rad(lat - ori[1])
        dr = np.sqrt(dx * dx + dy * dy)
        dth = np.arctan2(dy, dx)
        dc = np.full(lon.shape, 1.3 * kw.get("gf_coef", 1.0))
        mid = (dr > dr1) & (dr <= dr0)
        dc[mid] *= np.cos(0.5 * np.pi * (dr[mid] - dr1) / (dr0 - dr1))
        dc[dr > dr0] = 0.0
        term = np.maximum(1000.0 * np.sin(0.4 * (0.5 * dr + dth) + 0.007 * np.cos(50.0 * dth) + 0.37 * np.pi), 999.0)
        return base + (term - 999.0) * dc
    if name == "harmonic":
        return 2.0 + np.sin(2.0 * _rad(lat)) ** 16 * np.cos(16.0 * _rad(lon))
    if name == "vortex":
        lon0, lat0 = kw.get("lon0_rad", 5.5), kw.get("lat0_rad", 0.2)
        r0, d, t = kw.get("r0", 3.0), kw.get("d", 5.0), kw.get("t", 6.0)
        # rotate so that (lon0, lat0) becomes the pole
        phi, lam = _rad(lat), _rad(lon) - lon0
        x = math.sin(lat0) * np.cos(phi) * np.cos(lam) - math.cos(lat0) * np.sin(phi)
        y = np.cos(phi) * np.sin(lam)
        z = math.sin(lat0) * np.sin(phi) + math.cos(lat0) * np.cos(phi) * np.cos(lam)
        rlon = np.arctan2(y, x)
        rlon = np.where(rlon < 0.0, rlon + 2.0 * np.pi, rlon)
        rho = r0 * np.cos(np.arcsin(np.clip(z, -1.0, 1.0)))
        vt = 1.5 * math.sqrt(3.0) * np.tanh(rho) / np.cosh(rho) ** 2
        omega = np.divide(vt, rho, out=np.zeros_like(rho), where=rho != 0.0)
        return 2.0 * (1.0 + np.tanh(rho / d * np.sin(rlon - omega * t)))
    raise ValueError(name)


def cell_average(ring, f, degree: int = 8, parts: bool = False):
    """int f dOmega / int dOmega over one cell (open ring of unit vectors, (n, 3))
    by the fan rule of the test helpers: f(lon_deg, lat_deg) vectorised; 0 for a
    cell of zero solid angle.  parts: (int f dOmega, int dOmega) instead."""
    bary, w = dunavant(degree)
    r = _f64(ring)
    A = r[0]
    num = den = 0.0
    for i in range(1, r.shape[0] - 1):
        B, C = r[i], r[i + 1]
        det = abs(float(np.dot(A, np.cross(B, C))))
        p = bary[:, :1] * A + bary[:, 1:2] * B + bary[:, 2:3] * C
        npn = np.linalg.norm(p, axis=1)
        s = p / npn[:, None]
        J = det / npn ** 3
        lon = np.degrees(np.arctan2(s[:, 1], s[:, 0]))
        lat = np.degrees(np.arcsin(np.clip(s[:, 2], -1.0, 1.0)))
        num += float(np.sum(w * f(lon, lat) * J))
        den += float(np.sum(w * J))
    if parts:
        return num, den
    return 0.0 if den == 0.0 else num / den
