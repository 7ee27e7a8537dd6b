"""Grids: host-side generators and device-resident cell tables.

The reference turns each grid into a tree with ``Trees.treeify``
(src/trees/interfaces.jl:94-129) and reads cells through ``getcell``.  The
device build needs flat polygon tables instead; a :class:`Grid` owns one in HBM.

Generators (no GPU needed; they run in libcrhip.so's host code):

* :func:`lonlat_points`  -- the ``Matrix{UnitSphericalPoint}`` of a lon-lat grid,
  shape ``(ni+1, nj+1, 3)`` in Julia index order (``P[i, j]``)
* :func:`cubed_sphere_points` -- 6 equiangular gnomonic faces, ``(6, ne+1, ne+1, 3)``
* :func:`healpix_cells`  -- HEALPix pixel boundaries with ``step`` points per side
* :func:`points_to_cells` -- CellBasedGrid ``getcell`` order (src/trees/grids.jl:71-84)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .errors import ArgumentError

NESTED = L.CR_HEALPIX_NESTED
RING = L.CR_HEALPIX_RING


def _order_code(order) -> int:
    if order in (NESTED, "nested", "NESTED", "Nested"):
        return NESTED
    if order in (RING, "ring", "RING", "Ring"):
        return RING
    raise ArgumentError(f"unknown HEALPix ordering {order!r}")


def lonlat_points(ni: int, nj: int, lon0: float = 0.0, lat0: float = -90.0,
                  dlon: float | None = None, dlat: float | None = None) -> np.ndarray:
    """Vertices ``lon0 + i*dlon, lat0 + j*dlat`` (degrees) of an ni x nj cell grid,
    as unit vectors, shape ``(ni+1, nj+1, 3)``.  Defaults cover the globe."""
    dlon = 360.0 / ni if dlon is None else dlon
    dlat = 180.0 / nj if dlat is None else dlat
    out = np.empty((nj + 1, ni + 1, 3))
    L.check(L.lib.cr_gen_lonlat_points(L.dptr(out), ni, nj, lon0, lat0, dlon, dlat), "lonlat_points")
    return out.transpose(1, 0, 2)


def cubed_sphere_points(ne: int) -> np.ndarray:
    """Equiangular cubed-sphere vertices, shape ``(6, ne+1, ne+1, 3)`` indexed
    ``[face, i, j]``; cell (f, i, j) has global index ``f*ne^2 + i + j*ne``."""
    out = np.empty((6, ne + 1, ne + 1, 3))
    L.check(L.lib.cr_gen_cubed_sphere_points(L.dptr(out), ne), "cubed_sphere_points")
    return out.transpose(0, 2, 1, 3)


def healpix_cells(nside: int, order="nested", step: int = 1, first: int = 0,
                  count: int | None = None) -> np.ndarray:
    """HEALPix cells ``[first, first+count)`` as ``(count, 4*step, 3)`` rings,
    counter-clockwise from the north corner."""
    n = 12 * nside * nside
    count = n - first if count is None else count
    out = np.empty((count, 4 * step, 3))
    L.check(L.lib.cr_gen_healpix_cells(L.dptr(out), nside, _order_code(order), step, first, count),
            "healpix_cells")
    return out


def healpix_nest2ring(nside: int, nest: int) -> int:
    r = ctypes.c_int64()
    L.check(L.lib.cr_healpix_nest2ring(nside, nest, ctypes.byref(r)), "nest2ring")
    return r.value


def healpix_ring2nest(nside: int, ring: int) -> int:
    r = ctypes.c_int64()
    L.check(L.lib.cr_healpix_ring2nest(nside, ring, ctypes.byref(r)), "ring2nest")
    return r.value


def _point_blocks(points) -> tuple[np.ndarray, int, int, int]:
    """(nblocks, ni+1, nj+1, 3) or (ni+1, nj+1, 3) in Julia order -> C layout."""
    p = np.asarray(points, dtype=np.float64)
    if p.ndim == 3:
        p = p[None]
    if p.ndim != 4 or p.shape[-1] != 3:
        raise ArgumentError("points must have shape (ni+1, nj+1, 3) or (nblocks, ni+1, nj+1, 3)")
    nb, W, H = p.shape[0], p.shape[1], p.shape[2]
    if W < 2 or H < 2:
        raise ArgumentError("a point matrix needs at least 2 x 2 points")
    c = np.ascontiguousarray(p.transpose(0, 2, 1, 3))  # (nb, nj+1, ni+1, 3): i fastest
    return c, nb, W - 1, H - 1


def points_to_cells(points) -> np.ndarray:
    """Polygon table ``(ncells, 4, 3)`` of point blocks, cell index
    ``b*ni*nj + i + j*ni`` with vertices P[i,j], P[i+1,j], P[i+1,j+1], P[i,j+1]."""
    c, nb, ni, nj = _point_blocks(points)
    out = np.empty((nb * ni * nj, 4, 3))
    L.check(L.lib.cr_points_to_cells(L.dptr(c), nb, ni, nj, L.dptr(out)), "points_to_cells")
    return out


class Grid:
    """A cell table in device memory (the build's view of a treeified grid).

    ``shape`` is the Julia-side array shape of fields on this grid (e.g.
    ``(ni, nj)`` for a lon-lat matrix); ``ncells == prod(shape)``.
    """

    def __init__(self, handle, shape, kind: str):
        self._h = ctypes.c_void_p(handle)
        n, nv, cv = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        L.check(L.lib.cr_grid_info(self._h, ctypes.byref(n), ctypes.byref(nv), ctypes.byref(cv)))
        self.ncells, self.nv, self.convex = n.value, nv.value, bool(cv.value)
        self.shape = tuple(shape)
        self.kind = kind
        nvert = ctypes.c_int64()
        L.check(L.lib.cr_grid_vertex_count(self._h, ctypes.byref(nvert)))
        self.nverts = nvert.value  # caller-visible vertices (gradients are indexed by these)

    # -- constructors ---------------------------------------------------------
    @classmethod
    def from_points(cls, points, stream=None) -> "Grid":
        """``Matrix{UnitSphericalPoint}`` of shape ``(ni+1, nj+1, 3)`` (one block,
        CellBasedGrid) or ``(nblocks, ni+1, nj+1, 3)`` (e.g. 6 cubed-sphere faces)."""
        c, nb, ni, nj = _point_blocks(points)
        h = ctypes.c_void_p()
        L.check(L.lib.cr_grid_from_points(ctypes.byref(h), L.dptr(c), nb, ni, nj, _stream(stream)),
                "Grid.from_points")
        shape = (ni, nj) if nb == 1 else (ni, nj, nb)
        return cls(h.value, shape, "points")

    @classmethod
    def from_cells(cls, cells, convex: bool, stream=None) -> "Grid":
        """Polygon table ``(ncells, nv, 3)``; ``convex`` declares every cell convex."""
        c = np.ascontiguousarray(cells, dtype=np.float64)
        if c.ndim != 3 or c.shape[2] != 3:
            raise ArgumentError("cells must have shape (ncells, nv, 3)")
        h = ctypes.c_void_p()
        L.check(L.lib.cr_grid_from_cells(ctypes.byref(h), L.dptr(c), c.shape[0], c.shape[1],
                                         1 if convex else 0, _stream(stream)), "Grid.from_cells")
        return cls(h.value, (c.shape[0],), "cells")

    @classmethod
    def lonlat(cls, ni: int, nj: int, **kw) -> "Grid":
        return cls.from_points(lonlat_points(ni, nj, **kw))

    @classmethod
    def cubed_sphere(cls, ne: int) -> "Grid":
        return cls.from_points(cubed_sphere_points(ne))

    @classmethod
    def healpix(cls, nside: int, order="nested", step: int = 1, stream=None) -> "Grid":
        """``HealpixRootNode{O}(nside)`` with ``step`` boundary points per side,
        generated on the device."""
        h = ctypes.c_void_p()
        L.check(L.lib.cr_grid_healpix(ctypes.byref(h), nside, _order_code(order), step, _stream(stream)),
                "Grid.healpix")
        return cls(h.value, (12 * nside * nside,), "healpix")

    # -- queries --------------------------------------------------------------
    def cells(self) -> np.ndarray:
        out = np.empty((self.ncells, self.nv, 3))
        L.check(L.lib.cr_grid_copy_cells(self._h, L.dptr(out)), "Grid.cells")
        return out

    def unit_areas(self) -> np.ndarray:
        out = np.empty(self.ncells)
        L.check(L.lib.cr_grid_areas(self._h, L.dptr(out)), "Grid.unit_areas")
        return out

    def __len__(self):
        return self.ncells

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib.cr_grid_destroy(h)
            self._h = ctypes.c_void_p()

    def __repr__(self):
        return f"Grid({self.kind}, ncells={self.ncells}, nv={self.nv}, convex={self.convex})"


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return getattr(stream, "cuda_stream", None)


def as_grid(g) -> Grid:
    """treeify: accept a Grid, a point matrix ``(ni+1, nj+1, 3)`` or point blocks."""
    if isinstance(g, Grid):
        return g
    a = np.asarray(g)
    if a.dtype.kind == "f" and a.shape[-1] == 3 and a.ndim in (3, 4):
        return Grid.from_points(a)
    from .errors import MethodError
    raise MethodError(f"cannot treeify {type(g).__name__}: pass a Grid or a unit-vector point matrix")
