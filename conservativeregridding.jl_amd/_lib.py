"""ctypes binding of libcrhip.so (the gfx950 C ABI declared in include/crhip.h).

The library is required: importing this module raises if libcrhip.so has not
been built (``python -c "import __graft_entry__ as g; g.build()"``).  There is
no CPU fallback for any compute entry point.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcrhip.so")

c_i32, c_i64, c_dbl, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)

CR_OK = 0
CR_INVALID_ARGUMENT = 1
CR_DIMENSION_MISMATCH = 2
CR_UNSUPPORTED = 3
CR_OUT_OF_MEMORY = 4
CR_HIP_ERROR = 5
CR_DEGENERATE_GEOMETRY = 6
CR_INTERNAL = 7
CR_NO_DEVICE = 8

CR_F64, CR_F32 = 0, 1
CR_HEALPIX_NESTED, CR_HEALPIX_RING = 0, 1
CR_CLIP_AUTO, CR_CLIP_WITH_DST, CR_CLIP_WITH_SRC = 0, 1, 2


class BuildOpts(ctypes.Structure):
    _fields_ = [("threshold_rel", c_dbl), ("radius", c_dbl), ("normalize", c_i32),
                ("clip_side", c_i32), ("reserved0", c_i32), ("reserved1", c_i32)]


class BuildStats(ctypes.Structure):
    _fields_ = [("n_dst", c_i64), ("n_src", c_i64), ("nnz", c_i64), ("n_candidates", c_i64),
                ("ms_total", c_dbl), ("ms_bounds", c_dbl), ("ms_tree", c_dbl), ("ms_search", c_dbl),
                ("ms_clip", c_dbl), ("ms_assemble", c_dbl), ("clip_with_dst", c_i32),
                ("flipped_cells", c_i32), ("ms_host", c_dbl), ("n_zero_area_cells", c_i64),
                ("n_zero_cell_pairs", c_i64), ("n_nonfinite_pairs", c_i64), ("n_nonfinite_cells", c_i64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class EsmfAttrs(ctypes.Structure):
    _fields_ = [("src_grid_name", ctypes.c_char_p), ("dst_grid_name", ctypes.c_char_p),
                ("created_at", ctypes.c_char_p), ("src_shape", ctypes.POINTER(c_i64)), ("src_ndim", c_i32),
                ("dst_ndim", c_i32), ("dst_shape", ctypes.POINTER(c_i64)), ("cdf_version", c_i32),
                ("reserved", c_i32)]


# name -> (restype, argtypes); exactly the entry points of include/crhip.h
SIGNATURES = {
    "cr_last_error": (ctypes.c_char_p, []),
    "cr_version": (ctypes.c_char_p, []),
    "cr_last_failed_pair": (c_i32, [_i64p, _i64p]),
    "cr_device_count": (c_i32, []),
    "cr_gen_lonlat_points": (c_i32, [_dp, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_dbl]),
    "cr_gen_cubed_sphere_points": (c_i32, [_dp, c_i64]),
    "cr_lonlat_to_xyz": (c_i32, [_dp, _dp, c_i64, _dp]),
    "cr_gen_healpix_cells": (c_i32, [_dp, c_i64, c_i32, c_i32, c_i64, c_i64]),
    "cr_healpix_nest2ring": (c_i32, [c_i64, c_i64, _i64p]),
    "cr_healpix_ring2nest": (c_i32, [c_i64, c_i64, _i64p]),
    "cr_points_to_cells": (c_i32, [_dp, c_i64, c_i64, c_i64, _dp]),
    "cr_polygons_width": (c_i32, [c_i64, _i64p, _i32p, c_i64, _i32p]),
    "cr_polygons_to_cells": (c_i32, [_dp, c_i64, _i64p, _i32p, c_i64, c_i32, _dp, _i32p]),
    "cr_validate_cells": (c_i32, [_dp, c_i64, c_i32, c_i32, _i64p]),
    "cr_validate_points": (c_i32, [_dp, c_i64, c_i64, c_i64, _i64p]),
    "cr_grid_from_cells": (c_i32, [ctypes.POINTER(c_vp), _dp, c_i64, c_i32, c_i32, c_vp]),
    "cr_grid_from_polygons": (c_i32, [ctypes.POINTER(c_vp), _dp, c_i64, _i64p, _i32p, c_i64, c_i32, c_vp]),
    "cr_grid_from_planar_polygons": (c_i32, [ctypes.POINTER(c_vp), _dp, c_i64, _i64p, _i32p, c_i64, c_i32, c_vp]),
    "cr_grid_is_planar": (c_i32, [c_vp, _i32p]),
    "cr_grid_copy_planar_cells": (c_i32, [c_vp, _dp]),
    "cr_grid_from_points": (c_i32, [ctypes.POINTER(c_vp), _dp, c_i64, c_i64, c_i64, c_vp]),
    "cr_grid_healpix": (c_i32, [ctypes.POINTER(c_vp), c_i64, c_i32, c_i32, c_vp]),
    "cr_grid_tripolar": (c_i32, [ctypes.POINTER(c_vp), _dp, c_i64, c_i64, c_vp]),
    "cr_se_tri_rule": (c_i32, [c_i32, _dp, _dp, _i32p]),
    "cr_se_gll": (c_i32, [c_i32, _dp, _dp]),
    "cr_se_nodes": (c_i32, [c_vp, c_i32, c_dbl, _dp, _dp]),
    "cr_se_moments": (c_i32, [c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp]),
    "cr_se_regridder_to_fv": (c_i32, [ctypes.POINTER(c_vp), c_vp, c_vp, c_i32, c_i32, _dp,
                                      ctypes.POINTER(BuildOpts), c_vp]),
    "cr_se_regridder_from_fv": (c_i32, [ctypes.POINTER(c_vp), c_vp, c_vp, c_i32, c_i32, _dp,
                                        ctypes.POINTER(BuildOpts), c_vp]),
    "cr_grid_fold_count": (c_i32, [c_vp, _i64p]),
    "cr_mirror_fold_partners": (c_i32, [c_vp, c_i32, c_i32, c_i64, c_vp, c_i64, c_i64, c_vp]),
    "cr_mirror_fold_partners_adjoint": (c_i32, [c_vp, c_i32, c_i64, c_vp, c_i64, c_i64, c_vp]),
    "cr_grid_info": (c_i32, [c_vp, _i64p, _i32p, _i32p]),
    "cr_grid_vertex_count": (c_i32, [c_vp, _i64p]),
    "cr_grid_copy_cells": (c_i32, [c_vp, _dp]),
    "cr_grid_areas": (c_i32, [c_vp, _dp]),
    "cr_grid_destroy": (None, [c_vp]),
    "cr_build_opts_default": (None, [ctypes.POINTER(BuildOpts)]),
    "cr_regridder_build": (c_i32, [ctypes.POINTER(c_vp), c_vp, c_vp, ctypes.POINTER(BuildOpts), c_vp]),
    "cr_regridder_from_csr": (c_i32, [ctypes.POINTER(c_vp), c_i64, c_i64, c_i64, _i64p, _i32p, _dp,
                                      _dp, _dp, c_vp]),
    "cr_regridder_info": (c_i32, [c_vp, _i64p, _i64p, _i64p]),
    "cr_regridder_stats": (c_i32, [c_vp, ctypes.POINTER(BuildStats)]),
    "cr_regridder_copy_csr": (c_i32, [c_vp, c_i32, _i64p, _i32p, _dp]),
    "cr_regridder_copy_areas": (c_i32, [c_vp, _dp, _dp]),
    "cr_regridder_normalize": (c_i32, [c_vp, c_vp]),
    "cr_regridder_destroy": (None, [c_vp]),
    "cr_regrid": (c_i32, [c_vp, c_i32, c_i32, c_i64, c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_i32,
                          c_vp]),
    "cr_regrid_backward": (c_i32, [c_vp, c_vp, c_vp, c_i32, c_i32, c_i64, c_vp, c_i64, c_i64, c_vp, c_i64, c_i64,
                                   c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp]),
    "cr_regrid_backward_noalloc": (c_i32, [c_vp, c_vp, c_vp, c_i32, c_i32, c_i64, c_vp, c_i64, c_i64, c_vp, c_i64,
                                           c_i64, c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp]),
    "cr_cell_averages": (c_i32, [c_vp, c_i32, _dp, c_i32, c_i32, c_vp, c_vp]),
    "cr_regrid_prepare": (c_i32, [c_vp, c_i32, c_i32, c_i64, c_i64, c_i64, c_vp]),
    "cr_regrid_tune": (c_i32, [c_vp, c_i32, c_i64, c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, _i32p,
                               ctypes.POINTER(ctypes.c_float), c_vp]),
    "cr_regrid_variant_count": (c_i32, [c_i32]),
    "cr_regridder_write_esmf": (c_i32, [c_vp, c_i32, ctypes.c_char_p, ctypes.POINTER(EsmfAttrs)]),
    "cr_esmf_write": (c_i32, [ctypes.c_char_p, c_i64, _dp, _i64p, _i64p, c_i64, _dp, _dp, c_i64, _dp, _dp,
                              ctypes.POINTER(EsmfAttrs)]),
    "cr_esmf_read_sizes": (c_i32, [ctypes.c_char_p, _i64p, _i64p, _i64p]),
    "cr_esmf_read": (c_i32, [ctypes.c_char_p, _dp, _i64p, _i64p, _dp, _dp, _dp, _dp]),
    "cr_regridder_from_esmf": (c_i32, [ctypes.POINTER(c_vp), ctypes.c_char_p, c_vp]),
    "cr_regrid_set_variant": (c_i32, [c_vp, c_i32, c_i32, c_i32]),
    "cr_regrid_tile_plan": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp]),
    "cr_regrid_copy_tile_plan": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                         c_vp]),
}


def _load():
    # PyTorch bundles its own HIP runtime under the same soname
    # (libamdhip64.so.7) as the one libcrhip.so links; whichever is loaded
    # first serves the whole process.  Load torch's first, so that torch and
    # this library share one runtime and torch keeps the runtime it was built
    # with (loading ours first leaves torch without devices).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built: the HIP library is required (no CPU fallback). "
            "Run __graft_entry__.build() first.")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()


class CRError(RuntimeError):
    """A failing library call (status + message from cr_last_error)."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


def last_error() -> str:
    m = lib.cr_last_error()
    return m.decode() if m else ""


def check(status: int, what: str = ""):
    if status == CR_OK:
        return
    from .errors import ArgumentError, DimensionMismatch, MethodError
    msg = f"{what}: {last_error()}" if what else last_error()
    if status == CR_INVALID_ARGUMENT:
        raise ArgumentError(msg)
    if status == CR_DIMENSION_MISMATCH:
        raise DimensionMismatch(msg)
    if status == CR_UNSUPPORTED:
        raise MethodError(msg)
    raise CRError(status, msg)


def dptr(a):
    return a.ctypes.data_as(_dp)


def i64ptr(a):
    return a.ctypes.data_as(_i64p)


def i32ptr(a):
    return a.ctypes.data_as(_i32p)
