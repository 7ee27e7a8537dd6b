"""``Regridder``, ``regrid!``, ``regrid``, ``transpose`` on the MI355X library.

Python mirror of the reference API (Python cannot spell ``regrid!``; the
in-place function is ``regrid_``, PyTorch's trailing-underscore convention):

=====================================  =============================================
reference (Julia)                      here
=====================================  =============================================
``Regridder(dst, src; normalize)``     ``Regridder(dst, src, normalize=...)``
  (src/regridder/regridder.jl:105-163)
``R.intersections``                    ``R.intersections`` (host scipy CSR copy)
``R.dst_areas`` / ``R.src_areas``      same (host numpy copies)
``size(R)``                            ``R.shape`` / ``size(R)``
``transpose(R)`` (regridder.jl:49-50)  ``transpose(R)`` / ``R.T`` -- no copy
``normalize!(R)`` (regridder.jl:54-62) ``normalize_(R)``
``regrid!(dst, R, src; dims=1,``       ``regrid_(dst, R, src, dims=1, normalize=True)``
``normalize=true)`` (regrid.jl:63-118)
``regrid(R, src)`` (regrid.jl:322-330) ``regrid(R, src)``
``areas(manifold, grid, tree)``        ``areas(grid)``
=====================================  =============================================

``dims`` keeps its Julia meaning: the 1-based index of the spatial axis of an
N-D field (``NDSliceLoop{dims}``, regrid.jl:157-318).  All fields of an N-D
array go through ONE device launch instead of a loop of 1-D slices.

Fields may be torch CUDA tensors (used in place, any strides) or host arrays
(numpy / CPU torch), which are staged through device buffers and copied back.
The compute path is the HIP library only; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib as L
from .errors import ArgumentError, DefaultIntersectionFailureError, DimensionMismatch, MethodError
from .grids import Grid, as_grid, as_planar_grid
from .manifolds import EARTH_RADIUS, Manifold, Planar, check_buildable, on_sphere, resolve_manifold


class _Core:
    """Shared state of a Regridder and its transposed views."""

    def __init__(self, handle):
        self.h = ctypes.c_void_p(handle)
        self._destroy = L.lib.cr_regridder_destroy  # kept: module globals may be gone at interpreter exit
        n_dst, n_src, nnz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib.cr_regridder_info(self.h, ctypes.byref(n_dst), ctypes.byref(n_src), ctypes.byref(nnz)))
        self.n_dst, self.n_src, self.nnz = n_dst.value, n_src.value, nnz.value
        self._areas = None
        self._csr = {}
        self.nfold = (0, 0)  # fold pairs of the (destination, source) grid

    def folds(self, transpose: bool) -> int:
        """Fold pairs of the grid this direction writes into."""
        return self.nfold[1] if transpose else self.nfold[0]

    def areas(self):
        if self._areas is None:
            a_dst, a_src = np.empty(self.n_dst), np.empty(self.n_src)
            L.check(L.lib.cr_regridder_copy_areas(self.h, L.dptr(a_dst), L.dptr(a_src)), "areas")
            self._areas = (a_dst, a_src)
        return self._areas

    def csr(self, transpose: bool):
        key = bool(transpose)
        if key not in self._csr:
            nrows = self.n_src if key else self.n_dst
            rp = np.empty(nrows + 1, np.int64)
            col = np.empty(self.nnz, np.int32)
            val = np.empty(self.nnz, np.float64)
            L.check(L.lib.cr_regridder_copy_csr(self.h, 1 if key else 0, L.i64ptr(rp), L.i32ptr(col),
                                                L.dptr(val)), "copy_csr")
            self._csr[key] = (rp, col, val)
        return self._csr[key]

    def invalidate(self):
        self._areas = None
        self._csr = {}

    def __del__(self):
        h, destroy = getattr(self, "h", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self.h = type(h)()  # no module globals: they may be gone at interpreter exit


def _raise_intersection_failure(dg: Grid, sg: Grid):
    """DefaultIntersectionFailureError(p1 = source cell, p2 = destination
    cell, e) for a build that failed on a non-finite area (regridder.jl:64-73)."""
    msg = L.last_error()
    d, s_ = ctypes.c_int64(-1), ctypes.c_int64(-1)
    L.check(L.lib.cr_last_failed_pair(ctypes.byref(d), ctypes.byref(s_)), "last_failed_pair")
    p1 = sg.cells()[s_.value] if s_.value >= 0 else None
    p2 = dg.cells()[d.value] if d.value >= 0 else None
    raise DefaultIntersectionFailureError(p1, p2, L.CRError(L.CR_DEGENERATE_GEOMETRY, msg))


class Regridder:
    """Conservative regridder from ``src`` cells to ``dst`` cells (rows = dst).

    ``Regridder(dst, src; normalize=false, intersection_operator, threaded)`` of
    src/regridder/regridder.jl:125-163.  ``dst``/``src`` are :class:`Grid` objects,
    unit-vector point matrices ``(ni+1, nj+1, 3)``, objects with the Trees
    interface, or planar inputs (planar :class:`Grid` objects, ``(x_edges,
    y_edges)`` tuples, ``(ni+1, nj+1, 2)`` point matrices).  The manifold is
    resolved as ``Regridder(dst, src)`` does (regridder.jl:105-123): a planar
    grid next to a spherical one is promoted to ``Spherical()`` (its
    coordinates read as lon/lat degrees), other mismatches raise
    ``ErrorException``, and two planar grids build on the plane (clip in the
    plane, shoelace areas, no radius: regridder.jl:87-94).  A pair whose
    intersection area is not finite raises ``DefaultIntersectionFailureError``
    with the two cells.  Extra keywords:

    * ``threshold_rel`` -- pairs with area <= threshold_rel * min(cell areas) are
      dropped (default 1e-14; 0 keeps every positive area like the reference)
    * ``manifold`` -- build on this manifold (``Regridder(manifold, dst, src)``)
    * ``radius`` -- sphere radius (areas in radius^2 units); default: the
      manifold's (``Spherical().radius`` = 6 371 008.8)
    * ``clip_side`` -- "auto", "dst" or "src": which side's (convex) cells clip
    * ``triangle_quad_degree`` -- Dunavant rule of the spectral-element weights
      when one side is a :class:`SpectralElementSpace` (default 2 (nq - 1))
    """

    def __init__(self, dst, src, *, normalize: bool = False, intersection_operator=None,
                 threaded=True, threshold_rel: float = 1e-14, radius: float | None = None,
                 manifold: Manifold | None = None, clip_side: str = "auto", triangle_quad_degree=None,
                 stream=None, _core=None, _transposed=False):
        if _core is not None:
            self._core, self._t = _core, _transposed
            return
        if intersection_operator is not None:
            raise MethodError("a custom intersection_operator runs per pair on the CPU; the device build "
                              "computes spherical clip areas only")
        from .se import SpectralElementSpace, build_se_regridder
        is_se = isinstance(dst, SpectralElementSpace) or isinstance(src, SpectralElementSpace)
        if manifold is None and not is_se:
            # Regridder(dst, src): best_manifold of each side, planar promoted (regridder.jl:105-123)
            manifold = resolve_manifold(dst, src)
        m_radius = check_buildable(manifold) if manifold is not None else EARTH_RADIUS
        opts = L.BuildOpts()
        L.lib.cr_build_opts_default(ctypes.byref(opts))
        opts.threshold_rel = float(threshold_rel)
        opts.radius = float(m_radius if radius is None else radius)
        opts.normalize = 1 if normalize else 0
        opts.clip_side = {"auto": 0, "dst": 1, "src": 2}[clip_side]
        if is_se:
            # the ClimaCore extension's constructors (ext/ConservativeRegriddingClimaCoreExt.jl:416-426, 489-508)
            h, self.dst_shape, self.src_shape = build_se_regridder(dst, src, opts, triangle_quad_degree, stream)
            self._core, self._t = _Core(h), False
            return
        if isinstance(manifold, Planar):
            dg, sg = as_planar_grid(dst), as_planar_grid(src)
        else:
            dg, sg = as_grid(on_sphere(dst)), as_grid(on_sphere(src))
        h = ctypes.c_void_p()
        st = L.lib.cr_regridder_build(ctypes.byref(h), dg._h, sg._h, ctypes.byref(opts), _stream(stream))
        if st == L.CR_DEGENERATE_GEOMETRY:
            _raise_intersection_failure(dg, sg)
        L.check(st, "Regridder")
        self._core, self._t = _Core(h.value), False
        self._core.nfold = (dg.nfold, sg.nfold)
        self.dst_shape, self.src_shape = dg.shape, sg.shape

    @classmethod
    def from_csr(cls, intersections, dst_areas, src_areas) -> "Regridder":
        """``Regridder(intersections, dst_areas, src_areas, ...)``: the field
        constructor of the reference struct (regridder.jl:25-36) from any scipy
        sparse matrix of shape (n_dst, n_src)."""
        import scipy.sparse as sp
        A = sp.csr_matrix(intersections)
        A.sort_indices()
        n_dst, n_src = A.shape
        a_dst = np.ascontiguousarray(dst_areas, np.float64)
        a_src = np.ascontiguousarray(src_areas, np.float64)
        if a_dst.shape != (n_dst,) or a_src.shape != (n_src,):
            raise DimensionMismatch("area vectors must match the matrix size")
        rp = np.ascontiguousarray(A.indptr, np.int64)
        col = np.ascontiguousarray(A.indices, np.int32)
        val = np.ascontiguousarray(A.data, np.float64)
        h = ctypes.c_void_p()
        L.check(L.lib.cr_regridder_from_csr(ctypes.byref(h), n_dst, n_src, val.shape[0], L.i64ptr(rp),
                                            L.i32ptr(col), L.dptr(val), L.dptr(a_dst), L.dptr(a_src), None),
                "Regridder.from_csr")
        return cls(None, None, _core=_Core(h.value))

    # -- the reference struct's fields ----------------------------------------
    @property
    def shape(self):
        c = self._core
        return (c.n_src, c.n_dst) if self._t else (c.n_dst, c.n_src)

    @property
    def nnz(self) -> int:
        return self._core.nnz

    @property
    def dst_areas(self) -> np.ndarray:
        a_dst, a_src = self._core.areas()
        return a_src if self._t else a_dst

    @property
    def src_areas(self) -> np.ndarray:
        a_dst, a_src = self._core.areas()
        return a_dst if self._t else a_src

    @property
    def intersections(self):
        """Host copy of the intersection matrix (scipy CSR, rows = this
        direction's destination cells)."""
        import scipy.sparse as sp
        rp, col, val = self._core.csr(self._t)
        return sp.csr_matrix((val, col, rp), shape=self.shape)

    def csr_arrays(self):
        """(rowptr, col, val) host copies of this direction's CSR."""
        return self._core.csr(self._t)

    @property
    def T(self) -> "Regridder":
        return transpose(self)

    @property
    def transposed(self) -> bool:
        return self._t

    def stats(self) -> dict:
        st = L.BuildStats()
        L.check(L.lib.cr_regridder_stats(self._core.h, ctypes.byref(st)))
        return st.as_dict()

    def __repr__(self):
        n2, n1 = self.shape
        return f"{n2}x{n1} Regridder(nnz={self.nnz}{', transposed' if self._t else ''})"


def transpose(R: Regridder) -> Regridder:
    """``transpose(R)``: the backward regridder sharing every array with ``R``
    (regridder.jl:49-50); its src/dst areas are R's dst/src areas."""
    return Regridder(None, None, _core=R._core, _transposed=not R._t)


def size(R: Regridder, *dims):
    s = R.shape
    return s if not dims else s[dims[0] - 1]


def normalize_(R: Regridder) -> Regridder:
    """``normalize!(R)``: A, src and dst areas divided by ``maximum(A)``."""
    L.check(L.lib.cr_regridder_normalize(R._core.h, None), "normalize!")
    R._core.invalidate()
    return R


def areas(grid, radius: float = EARTH_RADIUS) -> np.ndarray:
    """Cell areas of a grid in radius^2 units (regridder.jl:165-178)."""
    return as_grid(grid).unit_areas() * (radius * radius)


# ---------------------------------------------------------------- fields --
def _stream(stream):
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available() and torch.cuda.is_initialized():
                return torch.cuda.current_stream().cuda_stream
        except Exception:
            pass
        return None
    if isinstance(stream, int):
        return stream
    return getattr(stream, "cuda_stream", None)


class _FieldView:
    """A field's device pointer, shape and element strides."""

    def __init__(self, ptr, shape, strides, dtype_code, keepalive=None, writeback=None):
        self.ptr, self.shape, self.strides = ptr, tuple(shape), tuple(strides)
        self.dtype_code, self.keepalive, self.writeback = dtype_code, keepalive, writeback


def _dtype_code(dt) -> int:
    s = str(dt)
    if s.endswith("float64"):
        return L.CR_F64
    if s.endswith("float32"):
        return L.CR_F32
    raise MethodError(f"fields must be float64 or float32, got {s}")


def _shape_of(field):
    try:
        import torch
        if isinstance(field, torch.Tensor):
            return tuple(field.shape), _dtype_code(field.dtype)
    except ImportError:  # pragma: no cover
        pass
    if isinstance(field, np.ndarray):
        return field.shape, _dtype_code(field.dtype)
    raise MethodError(f"no regrid! method for a field of type {type(field).__name__}")


def _extract(field, writable: bool) -> _FieldView:
    """extract_*_arraylike + initialize_regridding! (regrid.jl:75-90, 264-274)."""
    try:
        import torch
    except ImportError:  # pragma: no cover
        torch = None
    if torch is not None and isinstance(field, torch.Tensor):
        code = _dtype_code(field.dtype)
        if field.is_cuda:
            return _FieldView(field.data_ptr(), field.shape, field.stride(), code, keepalive=field)
        dev = field.to("cuda")
        wb = (lambda d, f=field: f.copy_(d.cpu())) if writable else None
        return _FieldView(dev.data_ptr(), dev.shape, dev.stride(), code, keepalive=dev, writeback=wb)
    if isinstance(field, np.ndarray):
        code = _dtype_code(field.dtype)
        import torch as _t
        dev = _t.from_numpy(np.ascontiguousarray(field)).to("cuda")
        wb = (lambda d, f=field: np.copyto(f, d.cpu().numpy())) if writable else None
        return _FieldView(dev.data_ptr(), dev.shape, dev.stride(), code, keepalive=dev, writeback=wb)
    raise MethodError(f"no regrid! method for a field of type {type(field).__name__}")


def _merge_fields(shape_s, strides_s, strides_d, dim):
    """Split the non-spatial axes into one flattened field index (the longest
    run, from the smallest stride, whose src and dst strides both form a dense
    mixed radix) and outer axes that are looped.

    Returns ``(nfields, sf, df, outer_axes)``."""
    axes = [a for a in range(len(shape_s)) if a != dim and shape_s[a] != 1]
    if not axes:
        return 1, 1, 1, []
    axes.sort(key=lambda a: (strides_s[a], strides_d[a]))
    sf, df = strides_s[axes[0]], strides_d[axes[0]]
    es, ed, nf = sf, df, 1
    inner = []
    for a in axes:
        if strides_s[a] != es or strides_d[a] != ed:
            break
        inner.append(a)
        nf *= shape_s[a]
        es *= shape_s[a]
        ed *= shape_s[a]
    outer = [a for a in axes if a not in inner]
    return nf, sf, df, outer


def regrid_(dst_field, R: Regridder, src_field, *, dims: int = 1, normalize: bool = True,
            stream=None, mirror_folds: bool = True):
    """``regrid!(dst_field, R, src_field; dims=1, normalize=true)``.

    d = (A s) ./ a_d (``normalize=False`` returns A s).  For N-D fields ``dims``
    is the 1-based spatial axis; every other axis is a field index and all
    fields run in one launch.  When the destination grid has a folded row
    (``Grid.tripolar``), its ghost cells receive their primary's value, as
    ``finalize_regridding!`` does for Oceananigans fields
    (ext/ConservativeRegriddingOceananigansExt.jl:207-240); ``mirror_folds=False``
    leaves them as a plain vector's finalize does (0, or NaN with normalize).
    Returns ``dst_field``.
    """
    n_out, n_in = R.shape
    src_shape, src_code = _shape_of(src_field)
    dst_shape, dst_code = _shape_of(dst_field)
    ns, nd = len(src_shape), len(dst_shape)
    if ns >= 2 and not (1 <= dims <= ns):
        raise ArgumentError(f"dims={dims} is out of range for a {ns}-dimensional array")
    if nd >= 2 and not (1 <= dims <= nd):
        raise ArgumentError(f"dims={dims} is out of range for a {nd}-dimensional array")
    if ns != nd:
        raise DimensionMismatch(f"source and destination ranks must match; got source rank {ns} "
                                f"and destination rank {nd}")
    if src_code != dst_code:
        raise MethodError("source and destination fields must share a float type")
    dim = 0 if ns == 1 else dims - 1
    other_s = tuple(s for a, s in enumerate(src_shape) if a != dim)
    other_d = tuple(s for a, s in enumerate(dst_shape) if a != dim)
    if other_s != other_d:
        raise DimensionMismatch(f"source and destination non-spatial axes must match; got source axes "
                                f"{other_s} and destination axes {other_d}")
    if src_shape[dim] != n_in or dst_shape[dim] != n_out:
        raise DimensionMismatch(f"regridder of size {R.shape} cannot map a field with {src_shape[dim]} "
                                f"cells to one with {dst_shape[dim]} cells")
    # extract_*_arraylike + initialize_regridding!: device views (host fields are staged)
    src = _extract(src_field, writable=False)
    dst = _extract(dst_field, writable=True)
    sc, dc = src.strides[dim], dst.strides[dim]
    st = _stream(stream)
    nf, sf, df, outer = _merge_fields(src.shape, src.strides, dst.strides, dim)
    transpose_flag = 1 if R._t else 0
    item = 8 if src.dtype_code == L.CR_F64 else 4
    # one launch for all flattened fields; axes that cannot be flattened with
    # them are looped (in the NDSliceLoop's column-major order)
    for idx in np.ndindex(*[src.shape[a] for a in reversed(outer)]):
        idx = idx[::-1]
        so = sum(i * src.strides[a] for i, a in zip(idx, outer))
        do = sum(i * dst.strides[a] for i, a in zip(idx, outer))
        L.check(L.lib.cr_regrid(R._core.h, transpose_flag, src.dtype_code, nf, src.ptr + so * item, sc, sf,
                                dst.ptr + do * item, dc, df, 1 if normalize else 0, st), "regrid!")
        if mirror_folds and R._core.folds(R._t):
            L.check(L.lib.cr_mirror_fold_partners(R._core.h, transpose_flag, src.dtype_code, nf,
                                                  dst.ptr + do * item, dc, df, st), "mirror_fold_partners!")
    if dst.writeback is not None:  # finalize_regridding! into a host field
        dst.writeback(dst.keepalive)
    return dst_field


def regrid(R: Regridder, src_field, **kw):
    """Allocate the destination and ``regrid_`` into it (regrid.jl:322-330)."""
    n_out, _ = R.shape
    dims = kw.get("dims", 1)
    try:
        import torch
        if isinstance(src_field, torch.Tensor):
            shape = list(src_field.shape)
            shape[0 if src_field.ndim == 1 else dims - 1] = n_out
            out = torch.zeros(shape, dtype=src_field.dtype, device=src_field.device)
            return regrid_(out, R, src_field, **kw)
    except ImportError:  # pragma: no cover
        pass
    a = np.asarray(src_field)
    shape = list(a.shape)
    shape[0 if a.ndim == 1 else dims - 1] = n_out
    out = np.zeros(shape, dtype=a.dtype)
    return regrid_(out, R, a, **kw)


def regrid_plan(dst_field, R: Regridder, src_field, *, dims: int = 1, normalize: bool = True,
                mirror_folds: bool = True):
    """Validate once and return a zero-argument callable that re-launches
    ``regrid_(dst_field, R, src_field)`` on the current stream with no host work
    beyond the launch (device fields only).  Suitable for timing loops and for
    capture into a HIP graph (the launch allocates nothing)."""
    import torch
    if not (isinstance(dst_field, torch.Tensor) and dst_field.is_cuda and isinstance(src_field, torch.Tensor)
            and src_field.is_cuda):
        raise MethodError("regrid_plan needs device (CUDA) tensors")
    regrid_(dst_field, R, src_field, dims=dims, normalize=normalize, mirror_folds=mirror_folds)  # validates
    ns = src_field.ndim
    dim = 0 if ns == 1 else dims - 1
    nf, sf, df, outer = _merge_fields(tuple(src_field.shape), src_field.stride(), dst_field.stride(), dim)
    if outer:
        raise MethodError("regrid_plan needs field axes that flatten into one stride")
    code = _dtype_code(src_field.dtype)
    # the per-direction set-up (default kernel, tile plan): afterwards the
    # launch allocates nothing and never synchronises (capturable)
    L.check(L.lib.cr_regrid_prepare(R._core.h, 1 if R._t else 0, code, nf, sf, df,
                                    torch.cuda.current_stream().cuda_stream), "regrid_prepare")
    args = (R._core.h, 1 if R._t else 0, code, nf, src_field.data_ptr(), src_field.stride(dim), sf,
            dst_field.data_ptr(), dst_field.stride(dim), df, 1 if normalize else 0)
    fn = L.lib.cr_regrid
    keep = (R, dst_field, src_field)
    mirror = None
    if mirror_folds and R._core.folds(R._t):
        margs = (R._core.h, 1 if R._t else 0, code, nf, dst_field.data_ptr(), dst_field.stride(dim), df)
        mirror = L.lib.cr_mirror_fold_partners

    def launch():
        stream = torch.cuda.current_stream().cuda_stream
        st = fn(*args, stream)
        if st:
            L.check(st, "regrid!")
        if mirror is not None:
            L.check(mirror(*margs, stream), "mirror_fold_partners!")
        return keep[1]

    return launch


def regrid_tune(dst_field, R: Regridder, src_field, *, dims: int = 1):
    """Autotune the FP32 cells-fastest batched apply for this direction of ``R``
    on these device fields (``dst_field`` receives the regridded result).  The
    fastest kernel variant with the default's numerics (FP64 products and sums
    from the FP64 A) is kept in the regridder for later calls.  Returns
    (variant index, milliseconds per launch)."""
    import torch
    regrid_(dst_field, R, src_field, dims=dims)  # validation
    dim = 0 if src_field.ndim == 1 else dims - 1
    nf, sf, df, outer = _merge_fields(tuple(src_field.shape), src_field.stride(), dst_field.stride(), dim)
    if outer or _dtype_code(src_field.dtype) != L.CR_F32:
        raise MethodError("regrid_tune needs FP32 fields whose field axes flatten into one stride")
    chosen, ms = ctypes.c_int32(-1), ctypes.c_float(0.0)
    L.check(L.lib.cr_regrid_tune(R._core.h, 1 if R._t else 0, nf, src_field.data_ptr(), src_field.stride(dim), sf,
                                 dst_field.data_ptr(), dst_field.stride(dim), df, ctypes.byref(chosen),
                                 ctypes.byref(ms), torch.cuda.current_stream().cuda_stream), "regrid_tune")
    return chosen.value, ms.value


def regrid_variants(dtype="float32") -> int:
    """Number of apply kernel variants for FP32 (``"float32"``) or FP64 fields."""
    return int(L.lib.cr_regrid_variant_count(_dtype_code(dtype)))


def regrid_set_variant(R: Regridder, variant: int, dtype="float32") -> None:
    """Pin the apply kernel variant used for this direction of ``R`` and this
    field dtype (``-1`` restores the default).  FP32: 0-4 row kernel with 1-16
    fields per thread, 5-7 staged tile kernel with 2/4/8 fields per stage and
    Kahan sums, 8-10 the same with FP64 sums, 11-14 variant 9 with other field
    chunkings, 15-17 FP64 sums of the FP64 A on 1024-row tiles, 18-19 Kahan
    sums on 1024-row tiles, 20-21 FP64 sums of A rounded to FP32 on 1024-row
    tiles, 22-23 the same on 32-row blocks sorted by their smallest column,
    24-27 field chunks swept one at a time, 28 variant 15 on the block order
    (default: 15, or 28 when its plan fetches fewer source lines than 15's and
    15's fetch over 3 lines per 32 distinct source cells);
    FP64: 0 row kernel, 1 tile kernel (default)."""
    L.check(L.lib.cr_regrid_set_variant(R._core.h, 1 if R._t else 0, _dtype_code(dtype), int(variant)),
            "regrid_set_variant")


def regrid_backward(R: Regridder, dst_grid, src_grid, x, y, gy, *, normalize: bool = True,
                    field_grad: bool = True, vertex_grad: bool = True, mirror_folds: bool = True):
    """Reverse mode of ``y = regrid(R, x)`` (FP64 device tensors, cells along
    dim 0, fields along dim 1 if 2-D).  Returns ``(gx, g_dst, g_src)``: the
    gradient w.r.t. x and w.r.t. the caller vertices of the two grids R was
    built from (``(grid.nverts, 3)`` tensors; tangential).  For ``transpose(R)``
    the roles of the grids follow the regridder's direction."""
    import torch
    for t in (x, gy) + ((y,) if normalize else ()):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64):
            raise MethodError("regrid_backward needs FP64 CUDA tensors")
    nf = 1 if x.ndim == 1 else x.shape[1]
    if mirror_folds and R._core.folds(R._t):
        # pullback of the fold mirror: a ghost's gradient belongs to its primary
        gy = gy.clone(memory_format=torch.contiguous_format)
        L.check(L.lib.cr_mirror_fold_partners_adjoint(
            R._core.h, 1 if R._t else 0, nf, gy.data_ptr(), gy.stride(0), gy.stride(1) if gy.ndim > 1 else 1,
            torch.cuda.current_stream().cuda_stream), "mirror_fold_partners! adjoint")

    def st(t):
        return (t.stride(0), t.stride(1) if t.ndim > 1 else 1)

    gx = torch.empty_like(x, memory_format=torch.contiguous_format) if field_grad else None
    g_dst = g_src = None
    if vertex_grad:
        g_dst = torch.zeros(dst_grid.nverts, 3, dtype=torch.float64, device=x.device)
        g_src = torch.zeros(src_grid.nverts, 3, dtype=torch.float64, device=x.device)
    (xc, xf), (gc, gf) = st(x), st(gy)
    yc, yf = st(y) if y is not None else (1, 1)
    gxc, gxf = st(gx) if gx is not None else (1, 1)
    L.check(L.lib.cr_regrid_backward(
        R._core.h, dst_grid._h, src_grid._h, 1 if R._t else 0, 1 if normalize else 0, nf,
        x.data_ptr(), xc, xf, y.data_ptr() if y is not None else None, yc, yf, gy.data_ptr(), gc, gf,
        gx.data_ptr() if gx is not None else None, gxc, gxf,
        g_dst.data_ptr() if g_dst is not None else None, g_src.data_ptr() if g_src is not None else None,
        torch.cuda.current_stream().cuda_stream), "regrid_backward")
    return gx, g_dst, g_src
