"""The C ABI library: it loads on a CPU-only host, exports every entry point
declared in include/crhip.h, and its Python binding declares each of them."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "crhip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_symbols()
    for must in ("cr_regridder_build", "cr_regrid", "cr_grid_healpix", "cr_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(cr):
    lib = ctypes.CDLL(cr.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header(cr):
    from importlib import import_module
    sig = import_module("crhip._lib").SIGNATURES
    assert sorted(sig) == declared_symbols()


def test_library_is_gfx950_code_object(cr):
    data = open(cr.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_device_count_never_fails(cr):
    assert cr.device_count() >= 0


def test_no_device_is_an_error_not_a_crash(cr):
    if cr.device_count() > 0:
        pytest.skip("a GPU is visible")
    import numpy as np
    with pytest.raises(cr.CRError) as e:
        cr.Grid.from_cells(np.eye(3)[None], convex=True)
    assert e.value.status == 8  # CR_NO_DEVICE


def test_last_error_is_thread_local_text(cr):
    from importlib import import_module
    L = import_module("crhip._lib")
    st = L.lib.cr_gen_cubed_sphere_points(None, 0)
    assert st == L.CR_INVALID_ARGUMENT
    assert "bad arguments" in L.last_error()


def test_tile_plan_entry_points_reject_bad_arguments(cr):
    """Host-side argument checks of the tile-plan entry points (no device call)."""
    L = cr._lib
    sizes = (ctypes.c_int64 * 6)()
    assert L.lib.cr_regrid_tile_plan(None, 0, 0, 0, sizes, None) == L.CR_INVALID_ARGUMENT
    assert L.lib.cr_regrid_copy_tile_plan(None, 0, 0, 0, None, None, None, None, None, None, None, None) \
        == L.CR_INVALID_ARGUMENT
    assert "null" in L.last_error()
