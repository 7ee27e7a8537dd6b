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
* :func:`polygons_to_cells` -- ragged polygon lists (any grid given by
  ``Trees.ncells`` / ``Trees.getcell``) padded to a fixed-width table
* :func:`ring_full_polygons` -- the cells of a RingGrids full grid
  (ext/ConservativeRegriddingRingGridsExt.jl:22-50)
* :func:`lonlat_to_unit` -- degree-exact geographic -> unit vectors

Planar grids (the ``Planar`` manifold: ``GO.SpatialTreeInterface.FlatNoTree`` of
planar polygons, matrices of planar rectangles) are made by
:meth:`Grid.from_planar_polygons`, :meth:`Grid.from_planar_rings` and
:meth:`Grid.from_planar_points`; their ``manifold`` is ``Planar()``.
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


def lonlat_to_unit(lon_deg, lat_deg) -> np.ndarray:
    """Unit vectors of (lon, lat) in degrees, shape ``broadcast.shape + (3,)``;
    exact at multiples of 90 degrees (``UnitSphereFromGeographic``)."""
    lon, lat = np.broadcast_arrays(np.asarray(lon_deg, dtype=np.float64), np.asarray(lat_deg, dtype=np.float64))
    lon, lat = np.ascontiguousarray(lon), np.ascontiguousarray(lat)
    out = np.empty(lon.shape + (3,))
    L.check(L.lib.cr_lonlat_to_xyz(L.dptr(lon), L.dptr(lat), lon.size, L.dptr(out)), "lonlat_to_unit")
    return out


def _ragged(vertices, offsets, indices):
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    idx = np.ascontiguousarray(indices, dtype=np.int32)
    if v.ndim != 2 or v.shape[1] != 3 or v.shape[0] == 0:
        raise ArgumentError("vertices must have shape (nverts, 3)")
    if off.ndim != 1 or off.size < 2 or idx.ndim != 1 or off[-1] != idx.size:
        raise ArgumentError("offsets must be (ncells+1,) with offsets[-1] == len(indices)")
    return v, off, idx


def polygons_width(vertices, offsets, indices) -> int:
    """Widest polygon (vertex count, closing repeats dropped) of a ragged list."""
    v, off, idx = _ragged(vertices, offsets, indices)
    nv = ctypes.c_int32()
    L.check(L.lib.cr_polygons_width(v.shape[0], L.i64ptr(off), L.i32ptr(idx), off.size - 1, ctypes.byref(nv)),
            "polygons_width")
    return nv.value


def polygons_to_cells(vertices, offsets, indices, nv: int | None = None):
    """Fixed-width table of a ragged polygon list: ``(cells (ncells, nv, 3),
    vidx (ncells, nv))``.  Cell c has vertices ``indices[offsets[c]:offsets[c+1]]``
    (a repeated closing index is dropped); narrower cells repeat their last vertex."""
    v, off, idx = _ragged(vertices, offsets, indices)
    nc = off.size - 1
    nv = polygons_width(v, off, idx) if nv is None else int(nv)
    cells = np.empty((nc, nv, 3))
    vidx = np.empty((nc, nv), dtype=np.int32)
    L.check(L.lib.cr_polygons_to_cells(L.dptr(v), v.shape[0], L.i64ptr(off), L.i32ptr(idx), nc, nv,
                                       L.dptr(cells), L.i32ptr(vidx)), "polygons_to_cells")
    return cells, vidx


def ring_full_polygons(latd, lond):
    """Cells of a RingGrids full grid with ring latitudes ``latd`` (degrees, north
    to south) and longitudes ``lond`` (cell centres, equally spaced), as a ragged
    polygon list ``(vertices, offsets, indices)``.

    Follows treeify for ``AbstractFullGrid`` (ext/ConservativeRegriddingRingGridsExt.jl:22-50):
    latitude edges are the midpoints between rings, pinned to +-90 at the poles;
    longitude edges sit half a cell west of each centre; the vertex matrix is
    ``P[i, j] = (lon_edges[i], lat_edges[nlat+1-j])`` (south to north in j) and
    cell ``k = ring*nlon + i`` (rings north to south, 0-based) is cartesian cell
    ``(i, nlat-1-ring)`` of the CellBasedGrid, the ``Reorderer2D`` of the extension."""
    latd = np.asarray(latd, dtype=np.float64)
    lond = np.asarray(lond, dtype=np.float64)
    nlat, nlon = latd.size, lond.size
    if nlat < 1 or nlon < 2:
        raise ArgumentError("a full ring grid needs at least one ring and two longitudes")
    lat_edges = np.empty(nlat + 1)
    lat_edges[0], lat_edges[-1] = 90.0, -90.0
    lat_edges[1:-1] = 0.5 * (latd[:-1] + latd[1:])
    dlon = 360.0 / nlon
    lon_edges = lond[0] - dlon / 2 + np.arange(nlon + 1) * dlon
    W = nlon + 1
    jj, ii = np.meshgrid(np.arange(nlat + 1), np.arange(W), indexing="ij")  # vertex id ii + jj*W
    verts = lonlat_to_unit(lon_edges[ii], lat_edges[nlat - jj]).reshape(-1, 3)
    ring, i = np.meshgrid(np.arange(nlat), np.arange(nlon), indexing="ij")
    j = nlat - 1 - ring
    quad = np.stack([i + j * W, i + 1 + j * W, i + 1 + (j + 1) * W, i + (j + 1) * W], axis=-1)
    idx = quad.reshape(-1).astype(np.int32)
    off = np.arange(0, idx.size + 1, 4, dtype=np.int64)
    return verts, off, idx


def ring_latitudes(kind: str, nlat_half: int) -> np.ndarray:
    """Ring latitudes (degrees, north to south) of RingGrids' full grids:
    ``"gaussian"`` -- 2*nlat_half Gauss-Legendre latitudes (FullGaussianGrid);
    ``"clenshaw"`` -- 2*nlat_half-1 equidistant latitudes without the poles,
    colatitudes j*180/(nlat+1) (FullClenshawGrid)."""
    if nlat_half < 1:
        raise ArgumentError("nlat_half must be positive")
    if kind == "gaussian":
        mu, _ = np.polynomial.legendre.leggauss(2 * nlat_half)
        return np.degrees(np.arcsin(mu[::-1]))
    if kind == "clenshaw":
        nlat = 2 * nlat_half - 1
        return 90.0 - np.arange(1, nlat + 1) * 180.0 / (nlat + 1)
    raise ArgumentError(f"unknown full ring grid {kind!r}")


def ring_longitudes(nlat_half: int) -> np.ndarray:
    """Cell-centre longitudes (degrees) of RingGrids' full grids: 4*nlat_half from 0."""
    nlon = 4 * nlat_half
    return np.arange(nlon) * (360.0 / nlon)


def flatten_tree(tree, share_vertices: bool = True, planar: bool = False):
    """Ragged polygon list ``(vertices, offsets, indices)`` of any object with the
    Trees interface: ``ncells`` (method or attribute) and ``getcell(i)`` (0-based,
    a ring of unit vectors ``(k, 3)`` or lon/lat degrees ``(k, 2)``, open or
    closed).  With ``share_vertices`` bitwise-equal vertices become one.
    ``planar``: the rings are planar ``(k, 2)`` points and stay so."""
    n = tree.ncells() if callable(getattr(tree, "ncells", None)) else int(tree.ncells)
    if n <= 0:
        raise ArgumentError("tree has no cells")
    rings = []
    for i in range(n):
        r = np.asarray(tree.getcell(i), dtype=np.float64)
        if planar:
            if r.ndim != 2 or r.shape[1] != 2:
                raise ArgumentError(f"getcell({i}) of a planar tree must return (k, 2) points")
        elif r.ndim != 2 or r.shape[1] not in (2, 3):
            raise ArgumentError(f"getcell({i}) must return (k, 3) unit vectors or (k, 2) lon/lat degrees")
        elif r.shape[1] == 2:
            r = lonlat_to_unit(r[:, 0], r[:, 1])
        rings.append(r)
    return _ragged_from_rings(rings, share_vertices)


def _ragged_from_rings(rings, share_vertices: bool):
    """``(vertices, offsets, indices)`` of a list of rings (open or closed)."""
    rings = [r[:-1] if r.shape[0] >= 2 and np.array_equal(r[0], r[-1]) else r for r in rings]
    n = len(rings)
    if n == 0:
        raise ArgumentError("no cells")
    lens = np.fromiter((r.shape[0] for r in rings), dtype=np.int64, count=n)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    allv = np.concatenate(rings)
    if share_vertices:
        verts, inv = np.unique(allv, axis=0, return_inverse=True)
        idx = inv.reshape(-1).astype(np.int32)
    else:
        verts, idx = allv, np.arange(allv.shape[0], dtype=np.int32)
    return verts, off, idx


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
        self._destroy = L.lib.cr_grid_destroy  # kept: module globals may be gone at interpreter exit
        n, nv, cv = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        L.check(L.lib.cr_grid_info(self._h, ctypes.byref(n), ctypes.byref(nv), ctypes.byref(cv)))
        self.ncells, self.nv, self.convex = n.value, nv.value, bool(cv.value)
        self.shape = tuple(shape)
        self.kind = kind
        nvert = ctypes.c_int64()
        L.check(L.lib.cr_grid_vertex_count(self._h, ctypes.byref(nvert)))
        self.nverts = nvert.value  # caller-visible vertices (gradients are indexed by these)
        nf = ctypes.c_int64()
        L.check(L.lib.cr_grid_fold_count(self._h, ctypes.byref(nf)))
        self.nfold = nf.value  # (primary, partner) pairs of a folded row
        pl = ctypes.c_int32()
        L.check(L.lib.cr_grid_is_planar(self._h, ctypes.byref(pl)))
        self.planar = bool(pl.value)
        from .manifolds import Planar, Spherical
        self.manifold = Planar() if self.planar else Spherical()  # GOCore.best_manifold

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
    def from_polygons(cls, vertices, offsets, indices, convex: bool = True, stream=None) -> "Grid":
        """Ragged polygon list with shared vertices: cell c is the ring
        ``vertices[indices[offsets[c]:offsets[c+1]]]`` (3..64 vertices; a repeated
        closing index is dropped).  Vertex gradients are indexed by ``vertices``."""
        v, off, idx = _ragged(vertices, offsets, indices)
        h = ctypes.c_void_p()
        L.check(L.lib.cr_grid_from_polygons(ctypes.byref(h), L.dptr(v), v.shape[0], L.i64ptr(off),
                                            L.i32ptr(idx), off.size - 1, 1 if convex else 0, _stream(stream)),
                "Grid.from_polygons")
        return cls(h.value, (off.size - 1,), "polygons")

    @classmethod
    def from_planar_polygons(cls, vertices, offsets, indices, convex: bool = True, stream=None) -> "Grid":
        """A grid on the Planar manifold: ragged polygon list of planar points,
        cell c = ``vertices[indices[offsets[c]:offsets[c+1]]]`` with ``vertices``
        of shape ``(nverts, 2)`` (3..64 vertices a cell, either orientation; a
        repeated closing index is dropped).  Regridders between planar grids
        clip in the plane (the planar operator of regridder.jl:87-94)."""
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        if v.ndim != 2 or v.shape[1] != 2 or v.shape[0] == 0:
            raise ArgumentError("planar vertices must have shape (nverts, 2)")
        if off.ndim != 1 or off.size < 2 or idx.ndim != 1 or off[-1] != idx.size:
            raise ArgumentError("offsets must be (ncells+1,) with offsets[-1] == len(indices)")
        h = ctypes.c_void_p()
        L.check(L.lib.cr_grid_from_planar_polygons(ctypes.byref(h), L.dptr(v), v.shape[0], L.i64ptr(off),
                                                   L.i32ptr(idx), off.size - 1, 1 if convex else 0,
                                                   _stream(stream)), "Grid.from_planar_polygons")
        return cls(h.value, (off.size - 1,), "planar")

    @classmethod
    def from_planar_rings(cls, rings, convex: bool = True, share_vertices: bool = True, stream=None) -> "Grid":
        """A planar grid from a list of rings, each ``(k, 2)`` (open or closed):
        the polygons of a ``FlatNoTree`` (test/usecases/simple.jl:10-26)."""
        rs = [np.asarray(r, dtype=np.float64) for r in rings]
        if any(r.ndim != 2 or r.shape[1] != 2 for r in rs):
            raise ArgumentError("every planar ring must have shape (k, 2)")
        return cls.from_planar_polygons(*_ragged_from_rings(rs, share_vertices), convex=convex, stream=stream)

    @classmethod
    def from_planar_points(cls, points, stream=None) -> "Grid":
        """A planar point matrix ``(ni+1, nj+1, 2)`` (``P[i, j]``): cell
        ``i + j*ni`` has corners P[i,j], P[i+1,j], P[i+1,j+1], P[i,j+1], the
        CellBasedGrid order of the spherical point matrices."""
        p = np.asarray(points, dtype=np.float64)
        if p.ndim != 3 or p.shape[2] != 2 or p.shape[0] < 2 or p.shape[1] < 2:
            raise ArgumentError("planar points must have shape (ni+1, nj+1, 2)")
        W, H = p.shape[0], p.shape[1]
        ni, nj = W - 1, H - 1
        verts = np.ascontiguousarray(p.transpose(1, 0, 2).reshape(-1, 2))  # vertex i + j*(ni+1)
        i, j = np.meshgrid(np.arange(ni), np.arange(nj), indexing="xy")  # (nj, ni): cell i + j*ni
        v00 = (i + j * W).reshape(-1)
        idx = np.stack([v00, v00 + 1, v00 + 1 + W, v00 + W], axis=1).astype(np.int32).reshape(-1)
        off = np.arange(0, 4 * ni * nj + 1, 4, dtype=np.int64)
        g = cls.from_planar_polygons(verts, off, idx, convex=True, stream=stream)
        g.shape = (ni, nj)
        return g

    @classmethod
    def from_tree(cls, tree, convex: bool = True, share_vertices: bool = True, stream=None) -> "Grid":
        """The generic tree adapter: flatten any object with the Trees interface,
        ``ncells`` (method or attribute) and ``getcell(i)`` (0-based, a ring of
        unit vectors ``(k, 3)`` or of lon/lat degrees ``(k, 2)``, open or closed),
        as ``Trees.treeify`` + ``Trees.getcell`` do for the reference's grids
        (src/trees/interfaces.jl:94-129, docs/src/interfaces/trees.md:34-44).
        With ``share_vertices`` bitwise-equal vertices become one vertex, so
        vertex gradients sum over the cells that share it."""
        verts, off, idx = flatten_tree(tree, share_vertices)
        g = cls.from_polygons(verts, off, idx, convex=convex, stream=stream)
        g.kind = "tree"
        return g

    @classmethod
    def from_lonlat_cells(cls, cells_deg, convex: bool = True, stream=None) -> "Grid":
        """Polygon table of lon/lat degrees ``(ncells, k, 2)`` (e.g. the E,S,W,N
        faces of a reduced ring grid, open or closed rings, either orientation)."""
        c = np.asarray(cells_deg, dtype=np.float64)
        if c.ndim != 3 or c.shape[2] != 2:
            raise ArgumentError("cells must have shape (ncells, k, 2)")
        if c.shape[1] >= 4 and np.array_equal(c[:, 0], c[:, -1]):
            c = c[:, :-1]
        return cls.from_cells(lonlat_to_unit(c[..., 0], c[..., 1]), convex=convex, stream=stream)

    @classmethod
    def ring_full(cls, latd, lond, stream=None) -> "Grid":
        """A RingGrids full grid (``treeify(::Spherical, ::AbstractFullGrid)``):
        cells ring by ring from the north, see :func:`ring_full_polygons`."""
        latd = np.asarray(latd, dtype=np.float64)
        g = cls.from_polygons(*ring_full_polygons(latd, lond), convex=True, stream=stream)
        g.kind = "ringgrid"
        return g

    @classmethod
    def tripolar(cls, points, stream=None) -> "Grid":
        """Tripolar grid with a RightCenterFolded top row from its point matrix
        ``(nx+1, ny+1, 3)`` (the treeify of ext/ConservativeRegriddingOceananigansExt.jl:119-187):
        fold-row cells i and nx-1-i (0-based) are one physical cell, so only
        ``i < nx/4`` and ``nx/2 <= i < nx/2 + nx/4`` keep their polygon; the
        others are zero-area ghosts, filled from their primary after every
        ``regrid_`` into this grid (``mirror_fold_partners!``)."""
        c, nb, nx, ny = _point_blocks(points)
        if nb != 1:
            raise ArgumentError("a tripolar grid is one point matrix (nx+1, ny+1, 3)")
        h = ctypes.c_void_p()
        L.check(L.lib.cr_grid_tripolar(ctypes.byref(h), L.dptr(c), nx, ny, _stream(stream)), "Grid.tripolar")
        return cls(h.value, (nx, ny), "tripolar")

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
        """The oriented polygon table: ``(ncells, nv, 3)`` unit vectors, or
        ``(ncells, nv, 2)`` planar points for a planar grid."""
        if self.planar:
            out = np.empty((self.ncells, self.nv, 2))
            L.check(L.lib.cr_grid_copy_planar_cells(self._h, L.dptr(out)), "Grid.cells")
            return out
        out = np.empty((self.ncells, self.nv, 3))
        L.check(L.lib.cr_grid_copy_cells(self._h, L.dptr(out)), "Grid.cells")
        return out

    def unit_areas(self) -> np.ndarray:
        """Cell areas on the unit sphere (planar areas for a planar grid)."""
        out = np.empty(self.ncells)
        L.check(L.lib.cr_grid_areas(self._h, L.dptr(out)), "Grid.unit_areas")
        return out

    def __len__(self):
        return self.ncells

    def __del__(self):
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self._h = type(h)()  # no module globals: they may be gone at interpreter exit

    def __repr__(self):
        return f"Grid({self.kind}, ncells={self.ncells}, nv={self.nv}, convex={self.convex})"


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return getattr(stream, "cuda_stream", None)


def as_planar_grid(g) -> Grid:
    """treeify on the Planar manifold: a planar Grid, a ``(lon_edges, lat_edges)``
    style tuple of edge vectors (rectangles), a planar point matrix
    ``(ni+1, nj+1, 2)``, a list of ``(k, 2)`` rings or a tree whose ``getcell``
    gives ``(k, 2)`` rings."""
    from .errors import ErrorException, MethodError
    from .manifolds import _is_edge_tuple
    if isinstance(g, Grid):
        if not g.planar:
            raise ErrorException("Destination and source manifolds must be the same.  Got Spherical() and Planar().")
        return g
    if _is_edge_tuple(g):
        x, y = (np.asarray(v, dtype=np.float64) for v in g)
        if x.size < 2 or y.size < 2:
            raise MethodError("a regular grid needs at least two edges in each direction")
        X, Y = np.meshgrid(x, y, indexing="ij")
        return Grid.from_planar_points(np.stack([X, Y], axis=-1))
    if hasattr(g, "getcell") and hasattr(g, "ncells"):
        verts, off, idx = flatten_tree(g, planar=True)
        gr = Grid.from_planar_polygons(verts, off, idx)
        gr.kind = "planar tree"
        return gr
    if isinstance(g, (list, tuple)):
        return Grid.from_planar_rings(g)
    a = np.asarray(g)
    if a.dtype.kind in "fi" and a.ndim == 3 and a.shape[-1] == 2:
        return Grid.from_planar_points(a)
    raise MethodError(f"cannot treeify {type(g).__name__} on the Planar manifold")


def as_grid(g) -> Grid:
    """treeify: accept a Grid, an object with the Trees interface (``ncells`` and
    ``getcell``), a point matrix ``(ni+1, nj+1, 3)`` or point blocks."""
    if isinstance(g, Grid):
        return g
    if hasattr(g, "getcell") and hasattr(g, "ncells"):
        return Grid.from_tree(g)
    a = np.asarray(g)
    if a.dtype.kind == "f" and a.shape[-1] == 3 and a.ndim in (3, 4):
        return Grid.from_points(a)
    from .errors import MethodError
    raise MethodError(f"cannot treeify {type(g).__name__}: pass a Grid or a unit-vector point matrix")
