"""Host-side semantics of regrid! (no GPU): the argument checks of the
reference's N-D pipeline, test/regridding.jl:127-291, run before any device call.

A stand-in core with the sizes of the reference test's regridder (9 dst cells,
4 src cells) is enough: validation never touches the library handle.
"""
import numpy as np
import pytest


class _Core:
    n_dst, n_src, nnz = 9, 4, 0


@pytest.fixture()
def R(cr):
    return cr.Regridder(None, None, _core=_Core())


def test_transpose_swaps_sizes_and_shares_state(cr, R):
    T = cr.transpose(R)
    assert T.shape == (4, 9) and R.shape == (9, 4)
    assert T._core is R._core
    assert cr.transpose(T).shape == R.shape
    assert cr.size(R) == (9, 4) and cr.size(R, 1) == 9


@pytest.mark.parametrize("dims", [0, 3])
def test_dims_out_of_range_is_argument_error(cr, R, dims):
    with pytest.raises(cr.ArgumentError):
        cr.regrid_(np.zeros((9, 3)), R, np.ones((4, 3)), dims=dims)


@pytest.mark.parametrize("dst,src,kw", [
    (np.zeros((9, 3, 1)), np.ones((4, 3)), {}),
    (np.zeros((9, 4)), np.ones((4, 3)), {}),
    (np.zeros((2, 9, 3)), np.ones((2, 4, 4)), {"dims": 2}),
    (np.zeros(8), np.ones(4), {}),
    (np.zeros(9), np.ones(5), {}),
])
def test_axis_mismatch_is_dimension_mismatch(cr, R, dst, src, kw):
    with pytest.raises(cr.DimensionMismatch):
        cr.regrid_(dst, R, src, **kw)


def test_non_array_field_is_method_error(cr, R):
    with pytest.raises(cr.MethodError):
        cr.regrid_(np.zeros(9), R, [1.0, 2.0, 3.0, 4.0])
    with pytest.raises(cr.MethodError):
        cr.regrid_(np.zeros(9, dtype=np.int64), R, np.ones(4, dtype=np.int64))
    with pytest.raises(cr.MethodError):
        cr.regrid_(np.zeros(9, dtype=np.float32), R, np.ones(4))


def test_custom_intersection_operator_is_rejected(cr):
    with pytest.raises(cr.MethodError):
        cr.Regridder(None, None, intersection_operator=lambda a, b: 1.0)


def test_field_merge_rules():
    from importlib import import_module
    m = import_module("crhip.regridder")._merge_fields
    # [n, 64, 4] cells fastest (Julia column-major, dims=1): one field stride
    n = 10
    assert m((n, 64, 4), (1, n, 64 * n), (1, 7, 64 * 7), 0) == (256, n, 7, [])
    # fields fastest
    assert m((n, 8), (8, 1), (8, 1), 0) == (8, 1, 1, [])
    # spatial axis in the middle of a C-ordered (2, n, 3) array (dims=2): the
    # 3 contiguous fields go in one launch, the outer axis is looped
    assert m((2, n, 3), (3 * n, 3, 1), (3 * n, 3, 1), 1) == (3, 1, 1, [0])
    # a gap between the field axes: loop the outer one
    assert m((n, 4, 4), (1, n, 8 * n), (1, n, 4 * n), 0) == (4, n, n, [2])
    # a single field
    assert m((n,), (1,), (1,), 0) == (1, 1, 1, [])
