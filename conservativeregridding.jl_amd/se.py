"""Spectral-element spaces: the ClimaCore extension's regridders
(ext/ConservativeRegriddingClimaCoreExt.jl).

A :class:`SpectralElementSpace` is a grid of quadrilateral elements (each
element's vertices c1..c4 are the corners (-1,-1), (1,-1), (1,1), (-1,1) of the
normalised bilinear map) with ``nq`` Gauss-Lobatto-Legendre nodes per
direction.  Fields on it are flat node vectors, element-major with i fastest
(``se_node_positions``).  ``Regridder(cells, space)`` is the principled
SE -> FV map, ``Regridder(space, cells)`` the per-element L2 projection
FV -> SE; SE -> SE is not supported, as in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .errors import ArgumentError, MethodError
from .grids import Grid, _stream


class SpectralElementSpace:
    """Elements (a :class:`Grid` of quadrilaterals or cubed-sphere point blocks
    ``(6, ne+1, ne+1, 3)``), ``nq`` nodes per direction and optional node
    weights W (``Spaces.weighted_jacobian``; default: the map's own,
    w_i w_j J(xi_i, eta_j) radius^2)."""

    def __init__(self, elements, nq: int = 4, node_weights=None):
        if not isinstance(elements, Grid):
            elements = Grid.from_points(np.asarray(elements, dtype=np.float64))
        if elements.nv != 4:
            raise ArgumentError("spectral elements must be quadrilaterals")
        if not 2 <= nq <= 8:
            raise ArgumentError("nq must be between 2 and 8")
        self.elements, self.nq = elements, int(nq)
        self.nelements = elements.ncells
        self.nnodes = self.nelements * self.nq * self.nq
        self._weights = None
        if node_weights is not None:
            w = np.ascontiguousarray(node_weights, dtype=np.float64).reshape(-1)
            if w.shape[0] != self.nnodes:
                raise ArgumentError(f"node_weights must have {self.nnodes} entries")
            self._weights = w

    @classmethod
    def cubed_sphere(cls, ne: int, nq: int = 4, **kw) -> "SpectralElementSpace":
        from .grids import cubed_sphere_points
        return cls(Grid.from_points(cubed_sphere_points(ne)), nq, **kw)

    @property
    def shape(self):
        return (self.nnodes,)

    def node_points(self) -> np.ndarray:
        """Unit vectors of the nodes, ``(nnodes, 3)`` (se_node_positions)."""
        out = np.empty((self.nnodes, 3))
        L.check(L.lib.cr_se_nodes(self.elements._h, self.nq, 1.0, L.dptr(out), None), "se_node_positions")
        return out

    def node_weights(self, radius: float = 1.0) -> np.ndarray:
        """W of every node (se_node_weights): the caller's, or the map's own at ``radius``."""
        if self._weights is not None:
            return self._weights.copy()
        out = np.empty(self.nnodes)
        L.check(L.lib.cr_se_nodes(self.elements._h, self.nq, float(radius), None, L.dptr(out)), "se_node_weights")
        return out

    def _weights_ptr(self):
        return L.dptr(self._weights) if self._weights is not None else None

    def __repr__(self):
        return f"SpectralElementSpace(nelements={self.nelements}, nq={self.nq})"


def tri_rule(degree: int):
    """``reference_rule(degree)``: (barycentric points (n, 3), weights (n,))."""
    bary, w, n = np.empty(48), np.empty(16), ctypes.c_int32()
    L.check(L.lib.cr_se_tri_rule(int(degree), L.dptr(bary), L.dptr(w), ctypes.byref(n)), "reference_rule")
    return bary[:3 * n.value].reshape(-1, 3).copy(), w[:n.value].copy()


def gll(n: int):
    """Gauss-Lobatto-Legendre nodes (ascending) and weights."""
    x, w = np.empty(n), np.empty(n)
    L.check(L.lib.cr_se_gll(int(n), L.dptr(x), L.dptr(w)), "gll")
    return x, w


def se_moments(R, cells: Grid, space: SpectralElementSpace, triangle_quad_degree: int = 0, stream=None):
    """``accumulate_principled_b`` for every non-zero of ``R`` (built as
    ``Regridder(cells, space.elements)``): ``(B, cell, element)`` device tensors,
    B ``(nnz, nq, nq)`` indexed ``[k, j, i]`` (node i + j*nq) in R's area units,
    and the cell and element of each pair k."""
    import torch
    nq, n = space.nq, max(R.nnz, 1)
    out = torch.empty((n, nq, nq), dtype=torch.float64, device="cuda")
    pc = torch.empty(n, dtype=torch.int32, device="cuda")
    pe = torch.empty(n, dtype=torch.int32, device="cuda")
    L.check(L.lib.cr_se_moments(R._core.h, cells._h, space.elements._h, nq, int(triangle_quad_degree),
                                out.data_ptr(), pc.data_ptr(), pe.data_ptr(),
                                _stream(stream) or torch.cuda.current_stream().cuda_stream),
            "accumulate_principled_b")
    return out[:R.nnz], pc[:R.nnz], pe[:R.nnz]


def build_se_regridder(dst, src, opts, triangle_quad_degree, stream):
    """The handle of Regridder(dst, src) when one side is a spectral-element space."""
    se_dst, se_src = isinstance(dst, SpectralElementSpace), isinstance(src, SpectralElementSpace)
    if se_dst and se_src:
        raise MethodError("SE → SE regridding is not supported. Use SE → FV → SE via two regridders.")
    deg = 0 if triangle_quad_degree is None else int(triangle_quad_degree)
    h = ctypes.c_void_p()
    if se_src:
        from .grids import as_grid
        cells = as_grid(dst)
        L.check(L.lib.cr_se_regridder_to_fv(ctypes.byref(h), cells._h, src.elements._h, src.nq, deg,
                                            src._weights_ptr(), ctypes.byref(opts), _stream(stream)),
                "Regridder(cells, space)")
        return h.value, cells.shape, src.shape
    from .grids import as_grid
    cells = as_grid(src)
    L.check(L.lib.cr_se_regridder_from_fv(ctypes.byref(h), dst.elements._h, cells._h, dst.nq, deg,
                                          dst._weights_ptr(), ctypes.byref(opts), _stream(stream)),
            "Regridder(space, cells)")
    return h.value, dst.shape, cells.shape
