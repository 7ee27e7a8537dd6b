"""Device tile plans (plan.hip) against the oracle's sequential restatement
(oracle/cr_tileplan.cpp), for every layout and geometry and both directions.

Synthetic matrices (regridding-like bands; ragged rows with repeats, huge rows
and empty rows; wide rows whose tiles start off the 32-row grid; no non-zeros)
are uploaded as given, so the oracle plans the same row order as the device:
the plans must agree bit for bit (descriptors, slot keys and values, column
lists, row maps, long rows).  A device-built C48 -> 1 degree regridder keeps
its rows in candidate order, which the host copy sorts: there the tile cuts,
descriptors, column lists, row maps and long rows must agree exactly, and the
device plan, run through the kernel's contract on the host, must give A x."""
import numpy as np
import pytest

from plan_tools import (EVEN, ROUNDS, assert_same_plan, block_order, device_plan, emulate, raw_regridder,
                        synthetic_csr)

pytestmark = pytest.mark.gpu

COMBOS = [(ROUNDS, 0), (ROUNDS, 1), (EVEN, 0), (EVEN, 1)]  # (style, geom)


@pytest.mark.parametrize("kind", ["banded", "ragged", "wide", "empty"])
def test_device_plan_matches_oracle_synthetic(cr, orc, kind):
    nrows, ncols, rp, col, val = synthetic_csr(kind, seed=5)
    h = raw_regridder(cr._lib, nrows, ncols, rp, col, val)
    trp, tcol, tval = orc.transpose(rp, col, val, nrows, ncols)
    try:
        for transpose, (r, c, v, nc) in enumerate(((rp, col, val, ncols), (trp, tcol, tval, nrows))):
            for style, geom in COMBOS:
                print(f"[case] plan {kind} transpose={transpose} style={style} geom={geom}", flush=True)
                dev = device_plan(cr._lib, h, transpose, style, geom, orc.empty_tile_plan)
                assert_same_plan(dev, orc.tile_plan(style, geom, r, c, v, nc))
    finally:
        cr._lib.lib.cr_regridder_destroy(h)


def test_device_plan_c48(cr, orc):
    R = cr.Regridder(cr.Grid.from_points(cr.lonlat_points(360, 180)),
                     cr.Grid.from_points(cr.cubed_sphere_points(48)), radius=1.0)
    for transpose in (0, 1):
        rp, col, val = R._core.csr(bool(transpose))  # rows sorted by column
        ncols = R._core.n_dst if transpose else R._core.n_src
        x = np.random.default_rng(7).random(ncols)
        want = np.zeros(rp.shape[0] - 1)
        np.add.at(want, np.repeat(np.arange(rp.shape[0] - 1), np.diff(rp)), val * x[col])
        for style, geom in COMBOS:
            print(f"[case] plan c48 transpose={transpose} style={style} geom={geom}", flush=True)
            dev = device_plan(cr._lib, R._core.h, transpose, style, geom, orc.empty_tile_plan)
            ref = orc.tile_plan(style, geom, rp, col, val, ncols)
            for k in ("desc", "rseg", "huge", "cols"):
                assert np.array_equal(dev[k], ref[k]) if k != "cols" else all(
                    np.array_equal(dev[k][d["col_off"]:d["col_off"] + d["ncols"]],
                                   ref[k][d["col_off"]:d["col_off"] + d["ncols"]]) for d in ref["desc"]), k
            y = emulate(dev, style, rp, col, val, x, orc.TILE_GEOMS[geom][0])
            np.testing.assert_allclose(y, want, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("kind", ["banded", "ragged", "wide"])
def test_device_block_plan_matches_oracle(cr, orc, kind):
    """Block-ordered plans: the device's row map equals the restated block
    order, and its plan equals the oracle's plan of the block-ordered matrix
    bit for bit (long rows as matrix rows)."""
    nrows, ncols, rp, col, val = synthetic_csr(kind, seed=6)
    h = raw_regridder(cr._lib, nrows, ncols, rp, col, val)
    try:
        rowmap, rp2, idx = block_order(rp, col)
        for style, geom in COMBOS:
            print(f"[case] block plan {kind} style={style} geom={geom}", flush=True)
            dev = device_plan(cr._lib, h, 0, style, geom, orc.empty_tile_plan, blocks=True)
            np.testing.assert_array_equal(dev["rowmap"], rowmap)
            ref = orc.tile_plan(style, geom, rp2, col[idx], val[idx], ncols)
            ref["huge"] = rowmap[ref["huge"]]
            assert_same_plan({k: v for k, v in dev.items() if k != "rowmap"}, ref)
    finally:
        cr._lib.lib.cr_regridder_destroy(h)
