"""Manifolds and the resolution ``Regridder(dst, src)`` makes before it builds.

The reference asks each grid for ``GOCore.best_manifold`` and promotes a
planar grid to the sphere when the other grid is spherical; any other
mismatch is an error (src/regridder/regridder.jl:105-123).  A device grid
knows its manifold (``Grid.manifold``: ``Planar()`` for the planar
constructors, ``Spherical()`` otherwise); the other planar inputs are a tuple
of edge vectors (``RegularGrid``, src/trees/interfaces.jl:128-129,
src/trees/grids.jl:96-115), a planar point matrix ``(ni+1, nj+1, 2)`` and
objects that say ``manifold = Planar()``.  Promoted to the sphere, planar
coordinates are read as (lon, lat) degrees.  Two planar grids build on the
plane: the device clips in the plane where the reference calls
``FosterHormannClipping(Planar())`` (regridder.jl:87-94).
"""
from __future__ import annotations

import numpy as np

from .errors import ErrorException, MethodError

#: GeometryOpsCore's default ``Spherical()`` radius (mean Earth radius, metres)
EARTH_RADIUS = 6371008.8


class Manifold:
    def __eq__(self, other):
        return type(self) is type(other) and vars(self) == vars(other)

    def __hash__(self):
        return hash((type(self).__name__,) + tuple(sorted(vars(self).items())))

    def __repr__(self):
        args = ", ".join(f"{k} = {v}" for k, v in vars(self).items())
        return f"{type(self).__name__}({args})"


class Planar(Manifold):
    pass


class Spherical(Manifold):
    def __init__(self, radius: float = EARTH_RADIUS):
        self.radius = float(radius)


class Geodesic(Manifold):
    """An ellipsoid (GeometryOpsCore's ``Geodesic``): no grid here lives on one,
    and a regridder cannot mix it with another manifold."""

    def __init__(self, semimajor_axis: float = 6378137.0, inv_flattening: float = 298.257223563):
        self.semimajor_axis = float(semimajor_axis)
        self.inv_flattening = float(inv_flattening)


def _is_edge_tuple(g) -> bool:
    if not (isinstance(g, tuple) and len(g) == 2):
        return False
    try:
        return all(np.asarray(v).ndim == 1 and np.asarray(v).dtype.kind in "fiu" for v in g)
    except Exception:  # noqa: BLE001 - anything not array-like is not an edge tuple
        return False


def _is_planar_points(grid) -> bool:
    if isinstance(grid, (list, tuple)) or hasattr(grid, "getcell"):
        return False
    try:
        a = np.asarray(grid)
    except Exception:  # noqa: BLE001 - not array-like
        return False
    return a.dtype.kind in "fi" and a.ndim == 3 and a.shape[-1] == 2


def best_manifold(grid) -> Manifold:
    """``GOCore.best_manifold(grid)`` for the grid forms the host layer accepts."""
    m = getattr(grid, "manifold", None)
    if isinstance(m, Manifold):
        return m
    if _is_edge_tuple(grid) or _is_planar_points(grid):
        return Planar()
    return Spherical()


def resolve_manifold(dst, src) -> Manifold:
    """The manifold ``Regridder(dst, src)`` builds on (regridder.jl:105-123)."""
    md, ms = best_manifold(dst), best_manifold(src)
    if md == ms:
        return md
    # implicit promotion to the sphere
    if isinstance(md, Planar) and isinstance(ms, Spherical):
        return Spherical()
    if isinstance(md, Spherical) and isinstance(ms, Planar):
        return Spherical()
    raise ErrorException(f"Destination and source manifolds must be the same.  Got {md!r} and {ms!r}.")


def on_sphere(grid):
    """A grid made ready for the spherical build: an edge tuple (lon, lat
    degrees) becomes the point matrix of its lon-lat cells (the RegularGrid's
    cell (i, j) = (x[i], y[j]), (x[i+1], y[j]), (x[i+1], y[j+1]), (x[i], y[j+1]));
    a planar point matrix or planar device grid is read as lon/lat degrees;
    everything else is returned as it is."""
    from .grids import Grid, lonlat_to_unit
    if isinstance(grid, Grid) and grid.planar:
        return Grid.from_lonlat_cells(grid.cells(), convex=grid.convex)
    if _is_planar_points(grid):
        a = np.asarray(grid, dtype=np.float64)
        return lonlat_to_unit(a[..., 0], a[..., 1])
    if not _is_edge_tuple(grid):
        return grid
    x, y = (np.asarray(v, dtype=np.float64) for v in grid)
    if x.size < 2 or y.size < 2:
        raise MethodError("a regular grid needs at least two edges in each direction")
    lon, lat = np.meshgrid(x, y, indexing="ij")  # (ni+1, nj+1)
    return lonlat_to_unit(lon, lat)


def check_buildable(manifold: Manifold) -> float:
    """The radius of a manifold the device build supports (1 on the plane:
    planar areas carry the coordinates' units)."""
    if isinstance(manifold, Spherical):
        return manifold.radius
    if isinstance(manifold, Planar):
        return 1.0
    raise MethodError(f"Regridder on {manifold!r} is not implemented by the device library")
