"""Reverse mode of regrid! (field and vertex gradients) vs the CPU oracle.

Loss L = ||regrid(R(V), s) - t||^2 with R built from a jittered cubed sphere
(source) and a lon-lat grid (destination) -- the differentiation configuration
at a size the oracle rebuilds in milliseconds.  Checks:
  * dL/ds against 2 A^T ((d - t) ./ a_d) from the oracle's CSR;
  * dL/dV against central finite differences (h = 1e-6) of the oracle-built
    loss, on random interior vertices (perturb, renormalise, rebuild);
  * dL/dV against forward-mode dual numbers through the oracle's clip and area
    (orc.clip_area_grad / orc.polygon_area_grad), assembled pair by pair.
Vertices on exactly shared edges are avoided: the intersection area is not
differentiable there (one-sided derivatives differ).
"""
import numpy as np
import pytest
import torch

from gpu_common import jittered_cubed_sphere

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def setup(cr, orc):
    ne = 12
    P = jittered_cubed_sphere(cr, ne, seed=11, frac=0.05)
    D = cr.lonlat_points(60, 30, lon0=0.5, lat0=-90.0, dlon=6.0, dlat=6.0)
    rng = np.random.default_rng(5)
    n_src, n_dst = 6 * ne * ne, 60 * 30
    s = rng.random(n_src)
    t = rng.random(n_dst)
    return P, D, s, t


def oracle_loss(cr, orc, P, D, s, t):
    ref = orc.build(cr.points_to_cells(D), cr.points_to_cells(P))
    d = ref.apply(s)
    return float(np.sum((d - t) ** 2)), ref, d


def test_field_and_vertex_gradients(cr, orc, setup):
    _case("grad_c4_small")
    P, D, s, t = setup
    src, dst = cr.Grid.from_points(P), cr.Grid.from_points(D)
    R = cr.Regridder(dst, src, radius=1.0)
    st = torch.tensor(s, device=DEV)
    d = cr.regrid(R, st)
    gy = 2.0 * (d - torch.tensor(t, device=DEV))
    gx, g_dst, g_src = cr.regrid_backward(R, dst, src, st, d, gy)
    L0, ref, d_ref = oracle_loss(cr, orc, P, D, s, t)
    # field gradient: 2 A^T ((d - t) ./ a_d)
    w = 2.0 * (d_ref - t) / ref.a_dst
    T = ref.transpose()
    gx_ref = orc.spmv(T.rowptr, T.col, T.val, w)
    assert np.max(np.abs(gx.cpu().numpy() - gx_ref)) <= 1e-12 * np.max(np.abs(gx_ref))

    g_src = g_src.cpu().numpy()
    g_dst = g_dst.cpu().numpy()
    ne = P.shape[1] - 1
    # central FD on random interior source vertices (points array index -> caller vertex)
    rng = np.random.default_rng(3)
    h = 1e-6
    worst = 0.0
    for _ in range(24):
        f, i, j = rng.integers(0, 6), rng.integers(1, ne), rng.integers(1, ne)
        v = f * (ne + 1) ** 2 + j * (ne + 1) + i  # point blocks are i fastest
        for k in range(3):
            Pp, Pm = P.copy(), P.copy()
            Pp[f, i, j, k] += h
            Pm[f, i, j, k] -= h
            Pp[f, i, j] /= np.linalg.norm(Pp[f, i, j])
            Pm[f, i, j] /= np.linalg.norm(Pm[f, i, j])
            fd = (oracle_loss(cr, orc, Pp, D, s, t)[0] - oracle_loss(cr, orc, Pm, D, s, t)[0]) / (2 * h)
            scale = max(abs(fd), np.abs(g_src[v]).max(), 1e-3)
            worst = max(worst, abs(g_src[v, k] - fd) / scale)
    assert worst < 1e-5, worst
    # and interior destination vertices (the area term acts here too)
    worst = 0.0
    W = D.shape[0]
    for _ in range(12):
        i, j = rng.integers(1, D.shape[0] - 1), rng.integers(2, D.shape[1] - 2)
        v = j * W + i
        for k in range(3):
            Dp, Dm = D.copy(), D.copy()
            Dp[i, j, k] += h
            Dm[i, j, k] -= h
            Dp[i, j] /= np.linalg.norm(Dp[i, j])
            Dm[i, j] /= np.linalg.norm(Dm[i, j])
            fd = (oracle_loss(cr, orc, P, Dp, s, t)[0] - oracle_loss(cr, orc, P, Dm, s, t)[0]) / (2 * h)
            scale = max(abs(fd), np.abs(g_dst[v]).max(), 1e-3)
            worst = max(worst, abs(g_dst[v, k] - fd) / scale)
    assert worst < 1e-5, worst


def test_vertex_gradient_matches_dual_numbers(cr, orc, setup):
    """Forward-mode duals through the oracle's clip: dL/dV of a source vertex =
    sum over the pairs of its cells of dL/dA_ij * dA_ij/dV."""
    _case("grad_dual")
    P, D, s, t = setup
    src, dst = cr.Grid.from_points(P), cr.Grid.from_points(D)
    R = cr.Regridder(dst, src, radius=1.0)
    st = torch.tensor(s, device=DEV)
    d = cr.regrid(R, st)
    gy = 2.0 * (d - torch.tensor(t, device=DEV))
    _, _, g_src = cr.regrid_backward(R, dst, src, st, d, gy, field_grad=False)
    g_src = g_src.cpu().numpy()
    L0, ref, d_ref = oracle_loss(cr, orc, P, D, s, t)
    scells, dcells = cr.points_to_cells(P), cr.points_to_cells(D)
    coef = 2.0 * (d_ref - t) / ref.a_dst  # dL/dA_ij = coef_i * s_j
    T = ref.transpose()  # rows = source cells
    ne = P.shape[1] - 1
    rng = np.random.default_rng(8)
    for _ in range(10):
        f, i, j = rng.integers(0, 6), rng.integers(1, ne), rng.integers(1, ne)
        v = f * (ne + 1) ** 2 + j * (ne + 1) + i
        g = np.zeros(3)
        # the 4 cells around point (i, j) of face f and the slot the point has in each
        for (ci, cj, slot) in ((i - 1, j - 1, 2), (i, j - 1, 3), (i, j, 0), (i - 1, j, 1)):
            c = f * ne * ne + cj * ne + ci
            for k in range(T.rowptr[c], T.rowptr[c + 1]):
                row = T.col[k]
                _, gsub, _ = orc.clip_area_grad(scells[c], dcells[row])
                g += coef[row] * s[c] * gsub[slot]
        np.testing.assert_allclose(g_src[v], g, rtol=1e-9, atol=1e-12 * np.abs(g).max())


def test_autograd_function(cr, setup):
    _case("autograd")
    from importlib import import_module
    ad = import_module("crhip.autodiff")
    P, D, s, t = setup
    Pt = torch.tensor(P, requires_grad=True)
    Dt = torch.tensor(D, requires_grad=True)
    st = torch.tensor(s, device=DEV, requires_grad=True)
    y = ad.regrid_ad(Dt, Pt, st)
    loss = ((y - torch.tensor(t, device=DEV)) ** 2).sum()
    loss.backward()
    assert st.grad is not None and Pt.grad.shape == Pt.shape and Dt.grad.shape == Dt.shape
    # interior source vertex gradients are tangential
    g = Pt.grad.numpy()
    radial = np.abs(np.sum(g * P, axis=-1))
    assert radial.max() <= 1e-10 * np.abs(g).max()


def test_cell_table_gradients_sum_to_point_gradients(cr, setup):
    """A destination given as a table of cells owns one vertex per cell slot
    (no vertex shared between cells: its area term adds without atomics); the
    slot gradients summed over the slots that are one point of the point grid
    give that grid's point gradient."""
    _case("grad_cell_table")
    P, D, s, t = setup
    src = cr.Grid.from_points(P)
    ni, nj = D.shape[0] - 1, D.shape[1] - 1
    grads = []
    for dst in (cr.Grid.from_points(D), cr.Grid.from_cells(cr.points_to_cells(D), convex=True)):
        R = cr.Regridder(dst, src, radius=1.0)
        st = torch.tensor(s, device=DEV)
        d = cr.regrid(R, st)
        gy = 2.0 * (d - torch.tensor(t, device=DEV))
        _, g_dst, _ = cr.regrid_backward(R, dst, src, st, d, gy, field_grad=False)
        grads.append(g_dst.cpu().numpy().reshape(-1, 3))
    g_pts, g_slots = grads
    # slot k of cell (i, j) is point (i, j), (i+1, j), (i+1, j+1), (i, j+1)
    i, j = np.meshgrid(np.arange(ni), np.arange(nj), indexing="xy")
    i, j = i.reshape(-1), j.reshape(-1)  # cell index i + j * ni
    corner = np.stack([i + j * (ni + 1), i + 1 + j * (ni + 1), i + 1 + (j + 1) * (ni + 1),
                       i + (j + 1) * (ni + 1)], axis=1).reshape(-1)
    summed = np.zeros_like(g_pts)
    np.add.at(summed, corner, g_slots)
    scale = np.abs(g_pts).max()
    np.testing.assert_allclose(summed, g_pts, rtol=0, atol=1e-10 * scale)


@pytest.mark.parametrize("step", [1, 2, 4])
def test_healpix_table_gradients_match_polygon_list(cr, setup, step):
    """A HEALPix destination (4*step-gons whose slots are their own vertices:
    its area term runs one lane per cell with the slot gradients in
    registers) against the same cells given as a polygon list without shared
    vertices (the area term through atomics): vertex and field gradients agree
    to 1e-12, the source's too (the clip pass is the same)."""
    _case("grad_healpix_table", step=step)
    P, _, _, _ = setup
    src = cr.Grid.from_points(P)
    hp = cr.Grid.healpix(8, "nested", step=step)
    cells = hp.cells()
    nv = cells.shape[1]
    as_list = cr.Grid.from_polygons(cells.reshape(-1, 3), np.arange(0, cells.shape[0] * nv + 1, nv),
                                    np.arange(cells.shape[0] * nv, dtype=np.int32), convex=False)
    rng = np.random.default_rng(8)
    s = torch.tensor(rng.random(src.ncells), device=DEV)
    out = []
    for dst in (hp, as_list):
        R = cr.Regridder(dst, src, radius=1.0)
        d = cr.regrid(R, s)
        gy = 2.0 * (d - 0.5)
        gx, g_dst, g_src = cr.regrid_backward(R, dst, src, s, d, gy)
        out.append([v.cpu().numpy().reshape(v.shape[0], -1) for v in (gx, g_dst, g_src)])
    for a, b in zip(*out):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


@pytest.fixture(scope="module")
def c4_full(cr):
    """The differentiation configuration at its size: jittered C96 -> 1 deg."""
    P = jittered_cubed_sphere(cr, 96, seed=4, frac=0.05)
    D = cr.lonlat_points(360, 180)
    rng = np.random.default_rng(21)
    return P, D, rng.random(6 * 96 * 96), rng.random(360 * 180)


def test_c4_full_size_gradients(cr, orc, c4_full):
    """C96 -> 1 deg: field gradient vs 2 A^T((d - t) ./ a_d) to 1e-12; vertex
    gradients of 200 random interior source vertices vs the dual-number pair
    assembly to 1e-9; central differences of the oracle-built loss on 10 of
    them to 1e-5."""
    _case("grad_c4_full", src="C96 jittered seed 4", dst="1deg", vertices=200, fd_vertices=10)
    P, D, s, t = c4_full
    src, dst = cr.Grid.from_points(P), cr.Grid.from_points(D)
    R = cr.Regridder(dst, src, radius=1.0)
    st = torch.tensor(s, device=DEV)
    d = cr.regrid(R, st)
    gy = 2.0 * (d - torch.tensor(t, device=DEV))
    gx, _, g_src = cr.regrid_backward(R, dst, src, st, d, gy)
    g_src = g_src.cpu().numpy()
    L0, ref, d_ref = oracle_loss(cr, orc, P, D, s, t)
    coef = 2.0 * (d_ref - t) / ref.a_dst
    T = ref.transpose()
    gx_ref = orc.spmv(T.rowptr, T.col, T.val, coef)
    assert np.max(np.abs(gx.cpu().numpy() - gx_ref)) <= 1e-12 * np.max(np.abs(gx_ref))

    scells, dcells = cr.points_to_cells(P), cr.points_to_cells(D)
    ne = P.shape[1] - 1
    rng = np.random.default_rng(12)
    picks = set()
    while len(picks) < 200:
        picks.add((int(rng.integers(0, 6)), int(rng.integers(1, ne)), int(rng.integers(1, ne))))
    picks = sorted(picks)
    worst = 0.0
    for f, i, j in picks:
        v = f * (ne + 1) ** 2 + j * (ne + 1) + i
        g = np.zeros(3)
        for (ci, cj, slot) in ((i - 1, j - 1, 2), (i, j - 1, 3), (i, j, 0), (i - 1, j, 1)):
            c = f * ne * ne + cj * ne + ci
            for k in range(T.rowptr[c], T.rowptr[c + 1]):
                row = T.col[k]
                _, gsub, _ = orc.clip_area_grad(scells[c], dcells[row])
                g += coef[row] * s[c] * gsub[slot]
        err = np.max(np.abs(g_src[v] - g)) / max(np.abs(g).max(), 1e-300)
        worst = max(worst, err)
        np.testing.assert_allclose(g_src[v], g, rtol=1e-9, atol=1e-12 * np.abs(g).max())
    print(f"c4 dual-number assembly: 200 vertices, worst rel err {worst:.2e}", flush=True)

    h = 1e-6
    worst = 0.0
    for f, i, j in picks[::20]:
        v = f * (ne + 1) ** 2 + j * (ne + 1) + i
        for k in range(3):
            Pp, Pm = P.copy(), P.copy()
            Pp[f, i, j, k] += h
            Pm[f, i, j, k] -= h
            Pp[f, i, j] /= np.linalg.norm(Pp[f, i, j])
            Pm[f, i, j] /= np.linalg.norm(Pm[f, i, j])
            fd = (oracle_loss(cr, orc, Pp, D, s, t)[0] - oracle_loss(cr, orc, Pm, D, s, t)[0]) / (2 * h)
            scale = max(abs(fd), np.abs(g_src[v]).max(), 1e-3)
            worst = max(worst, abs(g_src[v, k] - fd) / scale)
    print(f"c4 central differences: 10 vertices, worst rel err {worst:.2e}", flush=True)
    assert worst < 1e-5, worst
