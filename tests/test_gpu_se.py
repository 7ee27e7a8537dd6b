"""Spectral-element weights on the device vs the oracle (oracle/cr_oracle_se.c
and oracle.se_to_fv / fv_to_se).

Per-pair moments of accumulate_principled_b, the SE -> FV principled
regridder and the FV -> SE L2 projection (ext/ConservativeRegriddingClimaCoreExt.jl)
on cubed-sphere elements against lon-lat cells, plus the reference's own
checks of them (test/usecases/climacore.jl: constants exact to 1e-10,
conservation to 1e-12; test/extensions/climacore.jl: node weights sum to the
sphere's area).
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _csr_close(R, ref, tol):
    rp, col, val = R.csr_arrays()
    np.testing.assert_array_equal(rp, ref.rowptr)
    np.testing.assert_array_equal(col, ref.col)
    scale = np.repeat(np.maximum.reduceat(np.abs(ref.val), ref.rowptr[:-1][np.diff(ref.rowptr) > 0]),
                      np.diff(ref.rowptr)[np.diff(ref.rowptr) > 0])
    err = np.max(np.abs(val - ref.val) / scale)
    assert err < tol, err
    return err


@pytest.mark.parametrize("nq", [3, 4])
def test_moments_match_oracle(cr, orc, nq):
    _case("se_moments", nq=nq)
    space = cr.SpectralElementSpace.cubed_sphere(6, nq)
    # no lon-lat edge on an element edge (meridians off 45 deg, no edge on the
    # equator): where edges coincide, rounding decides whether the clipped
    # polygon carries a near-duplicate vertex, which moves the fan's centre
    P_ll = cr.lonlat_points(48, 25, lon0=1.7)
    cells = cr.Grid.from_points(P_ll)
    P = cr.Regridder(cells, space.elements, radius=1.0)
    B, pc, pe = cr.se_moments(P, cells, space)
    B, pc, pe = B.cpu().numpy(), pc.cpu().numpy(), pe.cpu().numpy()  # B[k, j, i]
    assert B.shape == (P.nnz, nq, nq)
    # the pairs are exactly R's non-zeros
    rp, col, _ = P.csr_arrays()
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    np.testing.assert_array_equal(np.sort(pc.astype(np.int64) * 10**6 + pe), np.sort(rows * 10**6 + col))
    E = cr.points_to_cells(cr.cubed_sphere_points(6))
    C = cr.points_to_cells(P_ll)
    rng = np.random.default_rng(2)
    worst = 0.0
    for k in rng.choice(P.nnz, 300, replace=False):
        ref = orc.se_moments(E[pe[k]], C[pc[k]], nq)  # [i, j]
        got = B[k].T
        assert abs(got.sum() - ref.sum()) < 1e-13 * abs(ref.sum()) + 1e-18
        worst = max(worst, np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print(f"  worst {worst:.2e}", flush=True)
    assert worst < 1e-11
    # coincident edges (lon-lat meridians on the cube's face edges, an edge on
    # the equator): sums still agree to rounding, single moments to the size of
    # a near-duplicate vertex's pull on the fan centre
    cells2 = cr.Grid.lonlat(48, 24)
    P2 = cr.Regridder(cells2, space.elements, radius=1.0)
    B2, pc2, pe2 = (t.cpu().numpy() for t in cr.se_moments(P2, cells2, space))
    C2 = cr.points_to_cells(cr.lonlat_points(48, 24))
    for k in rng.choice(P2.nnz, 200, replace=False):
        ref = orc.se_moments(E[pe2[k]], C2[pc2[k]], nq)
        assert abs(B2[k].sum() - ref.sum()) < 1e-13 * abs(ref.sum()) + 1e-18
        assert np.max(np.abs(B2[k].T - ref)) < 1e-3 * np.max(np.abs(ref))


def test_node_points_and_weights(cr, orc):
    _case("se_nodes")
    space = cr.SpectralElementSpace.cubed_sphere(4, 4)
    X = space.node_points()
    np.testing.assert_allclose(np.linalg.norm(X, axis=1), 1.0, atol=1e-15)
    W = space.node_weights()
    E = cr.points_to_cells(cr.cubed_sphere_points(4))
    np.testing.assert_allclose(W, orc.se_node_weights(E, 4), rtol=1e-13)
    assert abs(W.sum() - 4 * math.pi) < 1e-5 * 4 * math.pi
    x, _ = orc.gll(4)
    e, i, j = 37, 2, 1
    np.testing.assert_allclose(X[e * 16 + i + 4 * j], orc.element_map(E[e], x[i], x[j]), atol=1e-15)
    R6 = space.node_weights(radius=6.0)
    np.testing.assert_allclose(R6, 36.0 * W, rtol=1e-14)


@pytest.mark.parametrize("radius", [1.0, 6371008.8])
def test_se_to_fv(cr, orc, radius):
    _case("se_to_fv", radius=radius)
    space = cr.SpectralElementSpace.cubed_sphere(4, 4)
    P_ll = cr.lonlat_points(36, 19, lon0=2.3)  # no edge on an element edge
    cells = cr.Grid.from_points(P_ll)
    R = cr.Regridder(cells, space, radius=radius)
    assert R.shape == (36 * 19, space.nnodes)
    ref = orc.se_to_fv(cr.points_to_cells(P_ll), cr.points_to_cells(cr.cubed_sphere_points(4)), 4)
    r2 = radius * radius
    ref.val *= r2
    ref.a_dst *= r2
    ref.a_src *= r2
    _csr_close(R, ref, 1e-11)
    np.testing.assert_allclose(R.dst_areas, ref.a_dst, rtol=1e-12)
    np.testing.assert_allclose(R.src_areas, ref.a_src, rtol=1e-12)
    one = torch.ones(space.nnodes, dtype=torch.float64, device=DEV)
    d = cr.regrid(R, one).cpu().numpy()
    assert np.max(np.abs(d - 1.0)) < 1e-10
    A = R.intersections
    src_int = float(np.asarray(A.sum(axis=0)).sum())
    assert abs(src_int - 4 * math.pi * r2) < 1e-12 * 4 * math.pi * r2
    assert abs(np.sum(d * R.dst_areas) - src_int) < 1e-12 * src_int
    # FP32 batched through the same matrix
    X = torch.rand(space.nnodes, 8, dtype=torch.float64, device=DEV)
    y64 = cr.regrid(R, X).cpu().numpy()
    y32 = cr.regrid(R, X.float()).cpu().numpy()
    assert np.max(np.abs(y32 - y64)) < 1e-6 * np.max(np.abs(y64))


def test_fv_to_se(cr, orc):
    _case("fv_to_se")
    space = cr.SpectralElementSpace.cubed_sphere(4, 4)
    P_ll = cr.lonlat_points(36, 19, lon0=2.3)  # no edge on an element edge
    cells = cr.Grid.from_points(P_ll)
    R = cr.Regridder(space, cells, radius=1.0)
    assert R.shape == (space.nnodes, 36 * 19)
    ref = orc.fv_to_se(cr.points_to_cells(cr.cubed_sphere_points(4)), cr.points_to_cells(P_ll), 4)
    err = _csr_close(R, ref, 1e-9)
    print(f"  fv_to_se err {err:.2e}", flush=True)
    np.testing.assert_array_equal(R.dst_areas, 1.0)
    one = torch.ones(36 * 19, dtype=torch.float64, device=DEV)
    assert np.max(np.abs(cr.regrid(R, one).cpu().numpy() - 1.0)) < 1e-10
    # SE -> FV -> SE keeps the latitude integral (rtol 1e-2)
    S = cr.Regridder(cells, space, radius=1.0)
    lat = torch.tensor(np.arcsin(space.node_points()[:, 2]), device=DEV)
    W = torch.tensor(space.node_weights(), device=DEV)
    back = cr.regrid(R, cr.regrid(S, lat))
    tot = float((lat.abs() * W).sum())
    assert abs(float((back * W).sum() - (lat * W).sum())) < 1e-2 * tot


def test_se_errors(cr):
    _case("se_errors")
    space = cr.SpectralElementSpace.cubed_sphere(2, 3)
    with pytest.raises(cr.MethodError):
        cr.Regridder(space, space)
    with pytest.raises(cr.ArgumentError):
        cr.SpectralElementSpace.cubed_sphere(2, 9)
    with pytest.raises(cr.ArgumentError):
        cr.SpectralElementSpace(cr.Grid.healpix(2, step=2), 4)
