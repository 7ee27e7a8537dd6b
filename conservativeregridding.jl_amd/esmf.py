"""ESMF offline-weights files: ``save_esmf_weights`` of
ext/ConservativeRegriddingNCDatasetsExt.jl:15-59, written by the library as
netCDF classic with no netCDF library (CDF-5 by default: 64-bit row/col like
the reference's Int64; CDF-2 on request), and read back (``read_esmf_weights``,
``regridder_from_esmf``) so that a build can be reused."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib as L
from .errors import ArgumentError


def _attrs(src_grid_name, dst_grid_name, src_shape, dst_shape, created_at, cdf_version=5):
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
    if cdf_version not in (2, 5):
        raise ArgumentError("cdf_version must be 2 or 5")
    at = L.EsmfAttrs(str(src_grid_name).encode(), str(dst_grid_name).encode(),
                     None if created_at is None else str(created_at).encode(), sp, sn, dn, dp, int(cdf_version), 0)
    return at, keep


def save_esmf_weights(path, regridder, *, src_grid_name: str = "source", dst_grid_name: str = "destination",
                      src_shape=None, dst_shape=None, created_at=None, cdf_version: int = 5) -> str:
    """Write ``regridder`` (or ``transpose(regridder)``) as an ESMF weights file:
    S = A ./ dst_areas[row], 1-based row/col in column-major order, frac_a,
    frac_b, area_a, area_b and the reference's global attributes
    ("normalization" = "destarea").  ``cdf_version`` 5 (64-bit row/col) or 2
    (32-bit).  Returns ``path``."""
    path = os.fspath(path)
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    at, keep = _attrs(src_grid_name, dst_grid_name, src_shape, dst_shape, created_at, cdf_version)
    L.check(L.lib.cr_regridder_write_esmf(regridder._core.h, 1 if regridder._t else 0, path.encode(),
                                          ctypes.byref(at)), "save_esmf_weights")
    del keep
    return path


def write_esmf_arrays(path, S, row, col, frac_a, area_a, frac_b, area_b, *, src_grid_name: str = "source",
                      dst_grid_name: str = "destination", src_shape=None, dst_shape=None,
                      created_at=None, cdf_version: int = 5) -> str:
    """The same file from host arrays (row/col 1-based); needs no GPU."""
    path = os.fspath(path)
    S = np.ascontiguousarray(S, np.float64)
    row = np.ascontiguousarray(row, np.int64)
    col = np.ascontiguousarray(col, np.int64)
    fa, aa = np.ascontiguousarray(frac_a, np.float64), np.ascontiguousarray(area_a, np.float64)
    fb, ab = np.ascontiguousarray(frac_b, np.float64), np.ascontiguousarray(area_b, np.float64)
    if not (S.shape == row.shape == col.shape and fa.shape == aa.shape and fb.shape == ab.shape):
        raise ArgumentError("write_esmf_arrays: array lengths disagree")
    at, keep = _attrs(src_grid_name, dst_grid_name, src_shape, dst_shape, created_at, cdf_version)
    L.check(L.lib.cr_esmf_write(path.encode(), S.shape[0], L.dptr(S), L.i64ptr(row), L.i64ptr(col), fa.shape[0],
                                L.dptr(fa), L.dptr(aa), fb.shape[0], L.dptr(fb), L.dptr(ab), ctypes.byref(at)),
            "write_esmf_arrays")
    del keep
    return path


def read_esmf_weights(path) -> dict:
    """The arrays of an ESMF weights file (any netCDF classic version: CDF-1, 2
    or 5): ``S``, ``row``, ``col`` (1-based, int64), ``frac_a``, ``area_a``,
    ``frac_b``, ``area_b``.  Needs no GPU."""
    p = os.fspath(path).encode()
    n_s, n_a, n_b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    L.check(L.lib.cr_esmf_read_sizes(p, ctypes.byref(n_s), ctypes.byref(n_a), ctypes.byref(n_b)), "read_esmf_weights")
    out = {"S": np.empty(n_s.value), "row": np.empty(n_s.value, np.int64), "col": np.empty(n_s.value, np.int64),
           "frac_a": np.empty(n_a.value), "area_a": np.empty(n_a.value), "frac_b": np.empty(n_b.value),
           "area_b": np.empty(n_b.value)}
    L.check(L.lib.cr_esmf_read(p, L.dptr(out["S"]), L.i64ptr(out["row"]), L.i64ptr(out["col"]),
                               L.dptr(out["frac_a"]), L.dptr(out["area_a"]), L.dptr(out["frac_b"]),
                               L.dptr(out["area_b"])), "read_esmf_weights")
    return out


def regridder_from_esmf(path, stream=None):
    """A device ``Regridder`` from an ESMF weights file: A_ij = S * area_b[i],
    dst_areas = area_b, src_areas = area_a (no vertex gradients)."""
    from .regridder import Regridder, _Core, _stream
    h = ctypes.c_void_p()
    L.check(L.lib.cr_regridder_from_esmf(ctypes.byref(h), os.fspath(path).encode(), _stream(stream)),
            "regridder_from_esmf")
    return Regridder(None, None, _core=_Core(h.value))
