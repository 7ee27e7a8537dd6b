"""The reference's example fields and their cell averages on the device.

``LongitudeField``, ``SinusoidField``, ``HarmonicField``, ``GulfStreamField`` and
``VortexField`` mirror src/utils/example_data.jl: callables of (lon, lat) in
degrees (numpy arrays or scalars), with the same keyword parameters and
defaults.  ``cell_averages(grid, field)`` is the test helpers'
``set_field_values!`` (lib/ConservativeRegriddingTestHelpers/src/
ConservativeRegriddingTestHelpers.jl:84-128): the average of the field over
every cell of a device grid, integrated on the device with the degree-8
symmetric triangle rule over the cell's fan (``cr_cell_averages``).
"""
from __future__ import annotations

import ctypes
from dataclasses import astuple, dataclass

import numpy as np

from . import _lib as L

FIELD_LONGITUDE, FIELD_SINUSOID, FIELD_HARMONIC, FIELD_GULF_STREAM, FIELD_VORTEX = 0, 1, 2, 3, 4


class ExampleFieldFunction:
    """Base of the example fields (``ExampleFieldFunction``)."""
    kind: int = -1

    def params(self):
        return ()


class LongitudeField(ExampleFieldFunction):
    """The longitude itself (example_data.jl:14-24)."""
    kind = FIELD_LONGITUDE

    def __call__(self, lon, lat):
        return np.asarray(lon, dtype=np.float64) + 0.0 * np.asarray(lat)


def _fcos(coef, dp_length, lon, lat):
    return coef - np.cos(np.pi * np.arccos(np.cos(np.radians(lon)) * np.cos(np.radians(lat))) / dp_length)


@dataclass
class SinusoidField(ExampleFieldFunction):
    """coefmult * (coef - cos(pi * acos(cos(lon) cos(lat)) / dp_length)) (example_data.jl:26-41)."""
    dp_length: float = 1.2 * np.pi
    coef: float = 2.0
    coefmult: float = 1.0
    kind = FIELD_SINUSOID

    def params(self):
        return astuple(self)

    def __call__(self, lon, lat):
        return self.coefmult * _fcos(self.coef, self.dp_length, np.asarray(lon, np.float64),
                                     np.asarray(lat, np.float64))


class HarmonicField(ExampleFieldFunction):
    """2 + sin(2 lat)^16 cos(16 lon) (example_data.jl:43-54)."""
    kind = FIELD_HARMONIC

    def __call__(self, lon, lat):
        return 2.0 + np.sin(2.0 * np.radians(lat)) ** 16 * np.cos(16.0 * np.radians(lon))


@dataclass
class GulfStreamField(ExampleFieldFunction):
    """The OASIS fcos field with a Gulf-Stream-like current (example_data.jl:56-117)."""
    dp_length: float = 1.2 * np.pi
    coef: float = 2.0
    gf_coef: float = 1.0
    gf_ori_lon: float = -80.0
    gf_ori_lat: float = 25.0
    gf_end_lon: float = -1.8
    gf_end_lat: float = 50.0
    gf_dmp_lon: float = -25.5
    gf_dmp_lat: float = 55.5
    kind = FIELD_GULF_STREAM

    def params(self):
        return astuple(self)

    def __call__(self, lon, lat):
        lon, lat = np.broadcast_arrays(np.asarray(lon, np.float64), np.asarray(lat, np.float64))
        c = np.pi / 180.0
        dr0 = np.hypot((self.gf_end_lon - self.gf_ori_lon) * c, (self.gf_end_lat - self.gf_ori_lat) * c)
        dr1 = np.hypot((self.gf_dmp_lon - self.gf_ori_lon) * c, (self.gf_dmp_lat - self.gf_ori_lat) * c)
        r = _fcos(self.coef, self.dp_length, lon, lat)
        plon = np.where(lon > 180.0, lon - 360.0, np.where(lon < -180.0, lon + 360.0, lon))
        dx, dy = (plon - self.gf_ori_lon) * c, (lat - self.gf_ori_lat) * c
        dr, dth = np.hypot(dx, dy), np.arctan2(dy, dx)
        dc = np.where(dr > dr0, 0.0,
                      np.where(dr > dr1, 1.3 * self.gf_coef * np.cos(np.pi * 0.5 * (dr - dr1) / (dr0 - dr1)),
                               1.3 * self.gf_coef))
        bump = np.maximum(1000.0 * np.sin(0.4 * (0.5 * dr + dth) + 0.007 * np.cos(50.0 * dth) + 0.37 * np.pi),
                          999.0) - 999.0
        return r + bump * dc


@dataclass
class VortexField(ExampleFieldFunction):
    """Spiral vortices at the poles of a rotated sphere (example_data.jl:119-156)."""
    lon0_rad: float = 5.5
    lat0_rad: float = 0.2
    r0: float = 3.0
    d: float = 5.0
    t: float = 6.0
    kind = FIELD_VORTEX

    def params(self):
        return astuple(self)

    def __call__(self, lon, lat):
        lon, lat = np.asarray(lon, np.float64), np.asarray(lat, np.float64)
        sc, cc = np.sin(self.lat0_rad), np.cos(self.lat0_rad)
        sl, cl = np.sin(np.radians(lat)), np.cos(np.radians(lat))
        trm = cl * np.cos(np.radians(lon) - self.lon0_rad)
        X = sc * trm - cc * sl
        Y = cl * np.sin(np.radians(lon) - self.lon0_rad)
        Z = sc * sl + cc * trm
        dlon = np.arctan2(Y, X)
        dlon = np.where(dlon < 0.0, dlon + 2.0 * np.pi, dlon)
        dlat = np.arcsin(np.clip(Z, -1.0, 1.0))
        rho = self.r0 * np.cos(dlat)
        vt = 3.0 * np.sqrt(3.0) / 2.0 / np.cosh(rho) ** 2 * np.tanh(rho)
        with np.errstate(invalid="ignore", divide="ignore"):
            omega = np.where(rho == 0.0, 0.0, vt / np.where(rho == 0.0, 1.0, rho))
        return 2.0 * (1.0 + np.tanh(rho / self.d * np.sin(dlon - omega * self.t)))


def cell_averages(grid, field: ExampleFieldFunction, degree: int = 8, out=None, stream=None):
    """Average of ``field`` over every cell of ``grid`` (FP64 device tensor of
    ``grid.ncells``): ``set_field_values!`` on the device.  Zero-area cells get 0."""
    import torch
    if not isinstance(field, ExampleFieldFunction) or field.kind < 0:
        raise TypeError("cell_averages takes one of the example fields")
    if out is None:
        out = torch.empty(grid.ncells, dtype=torch.float64, device="cuda")
    if out.dtype != torch.float64 or not out.is_cuda or out.numel() != grid.ncells or not out.is_contiguous():
        raise ValueError("out must be a contiguous FP64 device tensor of ncells values")
    p = np.asarray(field.params(), dtype=np.float64)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    L.check(L.lib.cr_cell_averages(grid._h, field.kind, p.ctypes.data_as(L._dp) if p.size else None, int(p.size),
                                   int(degree), ctypes.c_void_p(out.data_ptr()), s), "cell_averages")
    return out
