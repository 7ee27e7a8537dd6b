"""ESMF offline-weights files: ``save_esmf_weights`` of
ext/ConservativeRegriddingNCDatasetsExt.jl:15-59, written by the library as
netCDF classic (64-bit offsets) with no netCDF library."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib as L
from .errors import ArgumentError


def _attrs(src_grid_name, dst_grid_name, src_shape, dst_shape, created_at):
    keep = []

    def shape(sh):
        if sh is None:
            return None, 0
        a = np.ascontiguousarray(sh, dtype=np.int64)
        if a.ndim != 1:
            raise ArgumentError("grid shapes must be 1-D sequences of integers")
        keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.shape[0]

    sp, sn = shape(src_shape)
    dp, dn = shape(dst_shape)
    at = L.EsmfAttrs(str(src_grid_name).encode(), str(dst_grid_name).encode(),
                     None if created_at is None else str(created_at).encode(), sp, sn, dn, dp)
    return at, keep


def save_esmf_weights(path, regridder, *, src_grid_name: str = "source", dst_grid_name: str = "destination",
                      src_shape=None, dst_shape=None, created_at=None) -> str:
    """Write ``regridder`` (or ``transpose(regridder)``) as an ESMF weights file:
    S = A ./ dst_areas[row], 1-based row/col in column-major order, frac_a,
    frac_b, area_a, area_b and the reference's global attributes
    ("normalization" = "destarea").  Returns ``path``."""
    path = os.fspath(path)
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    at, keep = _attrs(src_grid_name, dst_grid_name, src_shape, dst_shape, created_at)
    L.check(L.lib.cr_regridder_write_esmf(regridder._core.h, 1 if regridder._t else 0, path.encode(),
                                          ctypes.byref(at)), "save_esmf_weights")
    del keep
    return path


def write_esmf_arrays(path, S, row, col, frac_a, area_a, frac_b, area_b, *, src_grid_name: str = "source",
                      dst_grid_name: str = "destination", src_shape=None, dst_shape=None,
                      created_at=None) -> str:
    """The same file from host arrays (row/col 1-based); needs no GPU."""
    path = os.fspath(path)
    S = np.ascontiguousarray(S, np.float64)
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    fa, aa = np.ascontiguousarray(frac_a, np.float64), np.ascontiguousarray(area_a, np.float64)
    fb, ab = np.ascontiguousarray(frac_b, np.float64), np.ascontiguousarray(area_b, np.float64)
    if not (S.shape == row.shape == col.shape and fa.shape == aa.shape and fb.shape == ab.shape):
        raise ArgumentError("write_esmf_arrays: array lengths disagree")
    at, keep = _attrs(src_grid_name, dst_grid_name, src_shape, dst_shape, created_at)
    L.check(L.lib.cr_esmf_write(path.encode(), S.shape[0], L.dptr(S), L.i32ptr(row), L.i32ptr(col), fa.shape[0],
                                L.dptr(fa), L.dptr(aa), fb.shape[0], L.dptr(fb), L.dptr(ab), ctypes.byref(at)),
            "write_esmf_arrays")
    del keep
    return path
