"""MI355X-native conservative regridding (the two hot stages of
ConservativeRegridding.jl): the device ``Regridder`` build and ``regrid!``.

Loaded from its folder by path (the folder name is not a Python identifier):
``__graft_entry__.load_package()`` returns this module.
"""
from ._lib import CRError, LIB_PATH, lib as _lib_handle
from .errors import ArgumentError, DefaultIntersectionFailureError, DimensionMismatch, ErrorException, MethodError
from .manifolds import Geodesic, Planar, Spherical, best_manifold, resolve_manifold
from .grids import (Grid, NESTED, RING, as_grid, as_planar_grid, cubed_sphere_points, healpix_cells,
                    flatten_tree, healpix_nest2ring, healpix_ring2nest, lonlat_points, lonlat_to_unit, points_to_cells,
                    polygons_to_cells, polygons_width, ring_full_polygons, ring_latitudes, ring_longitudes)
from .esmf import read_esmf_weights, regridder_from_esmf, save_esmf_weights, write_esmf_arrays
from .se import SpectralElementSpace, se_moments
from .fields import (ExampleFieldFunction, GulfStreamField, HarmonicField, LongitudeField, SinusoidField, VortexField,
                     cell_averages)
from .regridder import (EARTH_RADIUS, Regridder, areas, normalize_, regrid, regrid_, regrid_plan,
                        regrid_tune, regrid_backward, regrid_variants, regrid_set_variant,
                        size, transpose)

__all__ = [
    "Regridder", "regrid_", "regrid", "regrid_plan", "regrid_tune", "regrid_backward", "regrid_variants",
    "regrid_set_variant", "transpose", "normalize_", "areas", "size", "save_esmf_weights", "write_esmf_arrays", "read_esmf_weights", "regridder_from_esmf",
    "SpectralElementSpace", "se_moments",
    "ExampleFieldFunction", "LongitudeField", "SinusoidField", "HarmonicField", "GulfStreamField", "VortexField",
    "cell_averages",
    "Grid", "NESTED", "RING", "as_grid", "as_planar_grid", "lonlat_points", "cubed_sphere_points", "healpix_cells",
    "healpix_nest2ring", "healpix_ring2nest", "points_to_cells", "lonlat_to_unit", "polygons_to_cells",
    "polygons_width", "flatten_tree", "ring_full_polygons", "ring_latitudes", "ring_longitudes",
    "ArgumentError", "DimensionMismatch", "MethodError", "ErrorException", "DefaultIntersectionFailureError",
    "CRError", "EARTH_RADIUS", "LIB_PATH", "Planar", "Spherical", "Geodesic", "best_manifold", "resolve_manifold",
    "device_count", "version",
]


def device_count() -> int:
    return int(_lib_handle.cr_device_count())


def version() -> str:
    return _lib_handle.cr_version().decode()
