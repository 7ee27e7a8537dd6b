"""The oracle's restatement of the tile plans (oracle/cr_tileplan.cpp), pinned
by running the staged kernel's contract on it (tests/plan_tools.emulate):
y = A x to rounding for random x, every segment stored once, every non-zero in
one slot, plus the budgets and the 32-row tile ends of the layout."""
import numpy as np
import pytest

from plan_tools import EVEN, ROUNDS, emulate, synthetic_csr

KINDS = ["banded", "ragged", "wide", "empty"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("style", [ROUNDS, EVEN])
@pytest.mark.parametrize("geom", [0, 1])
def test_oracle_plan_applies_the_matrix(orc, kind, style, geom):
    nrows, ncols, rp, col, val = synthetic_csr(kind, seed=3)
    P = orc.tile_plan(style, geom, rp, col, val, ncols)
    threads, rows, cols, segs = orc.TILE_GEOMS[geom]
    x = np.random.default_rng(1).random(ncols)
    y = emulate(P, style, rp, col, val, x, threads)
    ref = np.zeros(nrows)
    np.add.at(ref, np.repeat(np.arange(nrows), np.diff(rp)), val * x[col])
    np.testing.assert_allclose(y, ref, rtol=1e-13, atol=1e-13)
    n = np.diff(rp)
    np.testing.assert_array_equal(P["huge"], np.flatnonzero(n > cols))
    d = P["desc"]
    assert np.all(d["nrows"] >= 1) and np.all(d["nrows"] <= rows)
    assert np.all(d["ncols"] <= cols) and np.all(d["ent"] <= 16)
    assert np.all(d["nseg"] <= segs)
    if style == ROUNDS:
        assert np.all(d["flush"][:, threads // 64:] == 0)
    else:
        assert np.all(d["flush"] == 0)
    # a tile ends on the 32-row grid unless it is the last one or shorter than
    # the distance to the grid line below its end
    ends = d["row0"] + d["nrows"]
    off_grid = (ends % 32 != 0) & (ends < nrows)
    assert np.all((ends[off_grid] // 32) * 32 <= d["row0"][off_grid])
    if kind == "wide":
        assert np.any(d["row0"] % 32 != 0), "the wide case starts tiles off the grid"
    if kind == "empty":
        assert np.all(d["ent"] == 0) and np.all(d["nseg"] == 0)


def test_oracle_plan_of_a_regridding_matrix(orc, cr):
    """A real pattern: the oracle's C24 -> 2 degree matrix, both directions."""
    src = cr.points_to_cells(cr.cubed_sphere_points(24))
    dst = cr.points_to_cells(cr.lonlat_points(180, 90))
    ref = orc.build(dst, src)
    for M in (ref, ref.transpose()):
        rp, col, val, ncols = M.rowptr, M.col, M.val, M.n_src
        x = np.random.default_rng(2).random(ncols)
        want = M.apply(x, normalize=False)
        for style, geom in ((ROUNDS, 1), (EVEN, 0)):
            P = orc.tile_plan(style, geom, rp, col, val, ncols)
            y = emulate(P, style, rp, col, val, x, orc.TILE_GEOMS[geom][0])
            np.testing.assert_allclose(y, want, rtol=1e-13, atol=1e-15)
