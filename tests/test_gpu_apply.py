"""Device regrid! (forward and transpose(R)) vs the CPU oracle and the README
golden vectors; batched FP32 vs FP64; the N-D / strided / host-array pipeline
of src/regridder/regrid.jl (test/regridding.jl:45-203 semantics)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEV = "cuda"


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def c2(cr, orc):
    sp_, cp = cr.lonlat_points(360, 180), cr.cubed_sphere_points(48)
    R = cr.Regridder(cr.Grid.from_points(cp), cr.Grid.from_points(sp_), radius=1.0)
    rp, col, val = R.csr_arrays()
    return R, sp.csr_matrix((val, col, rp), shape=R.shape)


def test_readme_golden_vectors(cr):
    _case("readme")
    d = json.load(open(os.path.join(GOLDEN, "readme_example.json")))
    R = cr.Regridder.from_csr(sp.csr_matrix(np.array(d["A_dense"])), d["dst_areas"], d["src_areas"])
    orange = torch.tensor(d["orange"], dtype=torch.float64, device=DEV)
    blue = torch.zeros(4, dtype=torch.float64, device=DEV)
    cr.regrid_(blue, R, orange)
    a_s, a_d = np.array(d["src_areas"]), np.array(d["dst_areas"])
    o, b = orange.cpu().numpy(), blue.cpu().numpy()
    assert (o * a_s).sum() / a_s.sum() == (b * a_d).sum() / a_d.sum()
    tilde = torch.zeros(5, dtype=torch.float64, device=DEV)
    T = cr.transpose(R)
    assert T.src_areas is R.dst_areas and T.dst_areas is R.src_areas
    cr.regrid_(tilde, T, blue)
    np.testing.assert_allclose(tilde.cpu().numpy(), d["orange_tilde"], rtol=2e-16, atol=0)


def test_fp64_forward_and_transpose_match_oracle(cr, orc, c2):
    _case("fp64_c2")
    R, A = c2
    rng = np.random.default_rng(0)
    s = rng.random(A.shape[1])
    ref = orc.spmv(A.indptr, A.indices, A.data, s, R.dst_areas)
    d = torch.zeros(A.shape[0], dtype=torch.float64, device=DEV)
    cr.regrid_(d, R, torch.tensor(s, device=DEV))
    got = d.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    # conservation, checked on every run
    assert abs((got * R.dst_areas).sum() - (s * R.src_areas).sum()) <= 1e-12 * (s * R.src_areas).sum()
    trp, tcol, tval = orc.transpose(A.indptr, A.indices, A.data, A.shape[0], A.shape[1])
    back_ref = orc.spmv(trp, tcol, tval, got, R.src_areas)
    back = torch.zeros(A.shape[1], dtype=torch.float64, device=DEV)
    cr.regrid_(back, cr.transpose(R), d)
    assert np.max(np.abs(back.cpu().numpy() - back_ref)) <= 1e-12 * np.max(np.abs(back_ref))


def test_normalize_false_returns_raw_product(cr, c2):
    _case("raw")
    R, A = c2
    s = np.random.default_rng(1).random(A.shape[1])
    d = torch.zeros(A.shape[0], dtype=torch.float64, device=DEV)
    cr.regrid_(d, R, torch.tensor(s, device=DEV), normalize=False)
    np.testing.assert_allclose(d.cpu().numpy(), A @ s, rtol=1e-13)


def _fields(n, k, seed):
    rng = np.random.default_rng(seed)
    base = rng.random(k)[None, :] + 0.01 * rng.random((n, k))
    return base


@pytest.mark.parametrize("layout", ["cells_fastest", "fields_fastest"])
def test_fp32_batched_matches_fp64(cr, c2, layout):
    _case("fp32_batched", layout=layout)
    R, A = c2
    n_dst, n_src = A.shape
    k = 256
    x64 = _fields(n_src, k, 3)  # (n_src, k)
    if layout == "cells_fastest":  # Julia [n, 64, 4] with dims=1
        x = torch.tensor(x64.T.reshape(4, 64, n_src), dtype=torch.float32, device=DEV).permute(2, 1, 0)
        y = torch.zeros(4, 64, n_dst, dtype=torch.float32, device=DEV).permute(2, 1, 0)
        cr.regrid_(y, R, x)
        got = y.permute(2, 1, 0).reshape(k, n_dst).T.cpu().numpy().astype(np.float64)
        xin = x.permute(2, 1, 0).reshape(k, n_src).T.cpu().numpy().astype(np.float64)
    else:
        x = torch.tensor(x64, dtype=torch.float32, device=DEV)
        y = torch.zeros(n_dst, k, dtype=torch.float32, device=DEV)
        cr.regrid_(y, R, x)
        got = y.cpu().numpy().astype(np.float64)
        xin = x.cpu().numpy().astype(np.float64)
    ref = (A @ xin) / R.dst_areas[:, None]
    rel = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert rel <= 1e-6, rel
    # conservation drift of every field
    tot_s = (xin * R.src_areas[:, None]).sum(0)
    tot_d = (got * R.dst_areas[:, None]).sum(0)
    assert np.max(np.abs(tot_d - tot_s) / np.abs(tot_s)) <= 1e-7


def test_fp64_batched_equals_loop_of_single_fields(cr, c2):
    _case("fp64_batched")
    R, A = c2
    n_dst, n_src = A.shape
    x = torch.tensor(_fields(n_src, 8, 5), device=DEV)  # (n_src, 8): fields fastest
    y = torch.zeros(n_dst, 8, dtype=torch.float64, device=DEV)
    cr.regrid_(y, R, x)
    for f in range(8):
        yf = torch.zeros(n_dst, dtype=torch.float64, device=DEV)
        cr.regrid_(yf, R, x[:, f].contiguous())
        assert torch.equal(yf, y[:, f])


def test_strided_views_match_dense(cr, c2):
    """Dense/strided dispatch (test/regridding.jl:45-125): a strided view gives
    the dense result; here no temp buffers are needed at all."""
    _case("strided")
    R, A = c2
    n_dst, n_src = A.shape
    s = torch.tensor(np.random.default_rng(2).random(n_src), device=DEV)
    ref = torch.zeros(n_dst, dtype=torch.float64, device=DEV)
    cr.regrid_(ref, R, s)
    big_src = torch.zeros(2 * n_src, dtype=torch.float64, device=DEV)
    big_src[::2] = s
    big_dst = torch.zeros(2 * n_dst, dtype=torch.float64, device=DEV)
    cr.regrid_(big_dst[::2], R, big_src[::2])
    assert torch.equal(big_dst[::2], ref)
    assert torch.all(big_dst[1::2] == 0)


def test_host_arrays_are_staged(cr, c2):
    _case("host")
    R, A = c2
    s = np.random.default_rng(3).random(A.shape[1])
    d = np.zeros(A.shape[0])
    out = cr.regrid_(d, R, s)
    assert out is d
    np.testing.assert_allclose(d, (A @ s) / R.dst_areas, rtol=1e-13)
    d2 = cr.regrid(R, s)
    np.testing.assert_array_equal(d2, d)


@pytest.mark.parametrize("shape_src,shape_dst,dims", [
    ((None, 3), (None, 3), 1),
    ((3, None), (3, None), 2),
    ((None, 3, 2), (None, 3, 2), 1),
    ((2, None, 3), (2, None, 3), 2),
    ((3, 2, None), (3, 2, None), 3),
])
def test_nd_constant_fields_stay_constant(cr, c2, shape_src, shape_dst, dims):
    """regrid! on N-D arrays with dims (test/regridding.jl:127-203)."""
    _case("nd", dims=dims, shape=shape_src)
    R, A = c2
    n_dst, n_src = A.shape
    ss = tuple(n_src if v is None else v for v in shape_src)
    ds = tuple(n_dst if v is None else v for v in shape_dst)
    for make in (lambda sh: torch.ones(sh, dtype=torch.float64, device=DEV), lambda sh: np.ones(sh)):
        src = make(ss)
        dst = torch.zeros(ds, dtype=torch.float64, device=DEV) if isinstance(src, torch.Tensor) else np.zeros(ds)
        cr.regrid_(dst, R, src, dims=dims)
        arr = dst.cpu().numpy() if isinstance(dst, torch.Tensor) else dst
        np.testing.assert_allclose(arr, 1.0, atol=1e-12)


def test_regional_source_leaves_uncovered_cells_zero(cr):
    """A source that covers part of the sphere: covered destination cells of a
    constant field are <= 1, uncovered ones are 0 (test/usecases/constant_field.jl:41-77)."""
    _case("regional")
    src_pts = cr.lonlat_points(40, 20, lon0=10.0, lat0=-20.0, dlon=1.0, dlat=2.0)
    R = cr.Regridder(cr.Grid.from_points(cr.lonlat_points(72, 36)), cr.Grid.from_points(src_pts), radius=1.0)
    d = cr.regrid(R, torch.ones(800, dtype=torch.float64, device=DEV)).cpu().numpy()
    covered = np.asarray(R.intersections.sum(1)).ravel() > 0
    assert np.all(d[~covered] == 0.0)
    assert np.all(d[covered] <= 1.0 + 1e-12) and np.any(d[covered] > 1.0 - 1e-12)
    # backward: every source cell is fully covered by the global destination
    back = cr.regrid(cr.transpose(R), torch.ones(R.shape[0], dtype=torch.float64, device=DEV)).cpu().numpy()
    np.testing.assert_allclose(back, 1.0, atol=1e-12)


def test_normalize_in_place_keeps_results(cr):
    _case("normalize_")
    R = cr.Regridder(cr.Grid.from_points(cr.lonlat_points(72, 36)), cr.Grid.from_points(cr.lonlat_points(90, 45)))
    s = torch.tensor(np.random.default_rng(4).random(R.shape[1]), device=DEV)
    d1 = cr.regrid(R, s)
    cr.normalize_(R)
    assert R.intersections.max() == 1.0
    d2 = cr.regrid(R, s)
    torch.testing.assert_close(d1, d2, rtol=1e-14, atol=0)


def test_autotuned_variant_is_accurate(cr, c2):
    _case("tune")
    R, A = c2
    n_dst, n_src = A.shape
    k = 64
    x64 = _fields(n_src, k, 9)
    x = torch.tensor(x64.T.copy(), dtype=torch.float32, device=DEV).T  # cells fastest (n, k)
    y = torch.zeros(k, n_dst, dtype=torch.float32, device=DEV).T
    variant, ms = cr.regrid_tune(y, R, x)
    assert variant >= 0 and ms > 0
    assert variant in (8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 25, 28)  # FP64 products and sums of the FP64 A
    y.zero_()
    cr.regrid_(y, R, x)
    ref = (A @ x.cpu().numpy().astype(np.float64)) / R.dst_areas[:, None]
    assert np.max(np.abs(y.cpu().numpy() - ref)) / np.max(np.abs(ref)) <= 1e-6
    # the transpose direction keeps its own (untuned) choice and stays correct
    back = torch.zeros(k, n_src, dtype=torch.float32, device=DEV).T
    cr.regrid_(back, cr.transpose(R), y)
    trp = A.T.tocsr()
    ref_b = (trp @ y.cpu().numpy().astype(np.float64)) / R.src_areas[:, None]
    assert np.max(np.abs(back.cpu().numpy() - ref_b)) / np.max(np.abs(ref_b)) <= 1e-6


def _ragged_matrix(seed=0):
    """CSR with empty rows, short rows, long rows and rows of more than 1024
    non-zeros (the tile kernel hands those to the long-row kernel)."""
    rng = np.random.default_rng(seed)
    n_dst, n_src = 3000, 5000
    rows, cols = [], []
    for i in range(n_dst):
        u = rng.random()
        n = 0 if u < 0.1 else int(rng.integers(1, 12) if u < 0.9 else rng.integers(12, 300))
        if i in (100, 2000):
            n = 1500 if i == 100 else 3000
        c0 = int(min(max(0, i * n_src // n_dst - n // 2), n_src - n))
        rows += [i] * n
        cols += list(range(c0, c0 + n))
    A = sp.csr_matrix((rng.random(len(rows)) + 0.1, (rows, cols)), shape=(n_dst, n_src))
    a_d = np.asarray(A.sum(1)).ravel() + 0.5
    a_s = np.asarray(A.sum(0)).ravel() + 0.5
    return A, a_d, a_s


@pytest.mark.parametrize("k", [1, 3, 20])
def test_every_kernel_variant_matches_reference(cr, k):
    """Each FP32 and FP64 apply variant, both directions, ragged rows and field
    counts that leave a partial last stage."""
    A, a_d, a_s = _ragged_matrix(1)
    R = cr.Regridder.from_csr(A, a_d, a_s)
    rng = np.random.default_rng(k)
    for trans in (False, True):
        RR = cr.transpose(R) if trans else R
        M = (A.T.tocsr() if trans else A)
        area = a_s if trans else a_d
        n_out, n_in = M.shape
        x64 = rng.random((n_in, k)) - 0.3
        ref = (M @ x64) / area[:, None]
        scale = (abs(M) @ abs(x64)) / area[:, None] + 1e-300
        for dtype, tol in ((torch.float32, 1e-6), (torch.float64, 1e-13)):
            name = "float32" if dtype == torch.float32 else "float64"
            for v in range(cr.regrid_variants(name)):
                _case("variant", k=k, transpose=trans, dtype=name, v=v)
                cr.regrid_set_variant(RR, v, name)
                # cells fastest: (n, k) with cell stride 1
                x = torch.tensor(x64.T.copy(), dtype=dtype, device=DEV).T
                y = torch.full((k, n_out), float("nan"), dtype=dtype, device=DEV).T
                if k > 1:
                    cr.regrid_(y, RR, x)
                else:
                    cr.regrid_(y[:, 0], RR, x[:, 0])
                xin = x.cpu().numpy().astype(np.float64)
                err = np.max(np.abs(y.cpu().numpy() - (M @ xin) / area[:, None]) / scale)
                assert err <= tol, (name, v, trans, err)
            cr.regrid_set_variant(RR, -1, name)
        assert np.all(np.isfinite(ref))


def test_default_variant_on_row_strips_over_nested_sources(cr):
    """Lon-lat destination rows (strips of a row-major grid) over HEALPix nested
    sources: the default FP32 apply (which may pick the block-ordered plan for
    such a matrix) and the block-ordered variants give the FP64 result to the
    FP32 tolerance, both directions, every destination row written."""
    _case("row_strips", src="healpix64 nested", dst="1deg lon-lat")
    src = cr.Grid.healpix(64, "nested", step=1)
    dst = cr.Grid.from_points(cr.lonlat_points(360, 180))
    R = cr.Regridder(dst, src, radius=1.0)
    rng = np.random.default_rng(3)
    k = 12
    for RR in (R, cr.transpose(R)):
        n_out, n_in = RR.shape
        x64 = torch.tensor(rng.random((k, n_in)), dtype=torch.float64, device=DEV)
        ref = torch.zeros(k, n_out, dtype=torch.float64, device=DEV)
        cr.regrid_(ref.T, RR, x64.T)
        x = x64.float()
        for v in (-1, 22, 23, 28):
            cr.regrid_set_variant(RR, v)
            y = torch.full((k, n_out), float("nan"), dtype=torch.float32, device=DEV)
            cr.regrid_(y.T, RR, x.T)
            assert torch.isfinite(y).all(), v
            err = ((y.double() - ref).abs() / ref.abs().clamp_min(1e-30)).max().item()
            assert err <= 1e-6, (v, err)
        cr.regrid_set_variant(RR, -1)


def test_tile_kernel_on_c2_fields(cr, c2):
    _case("tiles_c2")
    R, A = c2
    n_dst, n_src = A.shape
    k = 37
    x64 = _fields(n_src, k, 11)
    x = torch.tensor(x64.T.copy(), dtype=torch.float32, device=DEV).T
    ref = (A @ x.cpu().numpy().astype(np.float64)) / R.dst_areas[:, None]
    outs = []
    for v in range(cr.regrid_variants("float32")):
        cr.regrid_set_variant(R, v)
        y = torch.zeros(k, n_dst, dtype=torch.float32, device=DEV).T
        cr.regrid_(y, R, x)
        outs.append(y.cpu().numpy())
        assert np.max(np.abs(outs[-1] - ref)) / np.max(np.abs(ref)) <= 1e-6, v
    cr.regrid_set_variant(R, -1)
    # the staged kernel is deterministic: same result on a second launch
    cr.regrid_set_variant(R, 6)
    y2 = torch.zeros(k, n_dst, dtype=torch.float32, device=DEV).T
    cr.regrid_(y2, R, x)
    cr.regrid_set_variant(R, -1)
    assert np.array_equal(y2.cpu().numpy(), outs[6])


def test_set_variant_rejects_unknown(cr, c2):
    R, _ = c2
    with pytest.raises(cr.ArgumentError):
        cr.regrid_set_variant(R, 99)


def test_graph_captured_apply_replays_match_eager(cr, c2):
    """regrid_plan launches captured in a HIP graph (torch.cuda.CUDAGraph):
    every replay, with eager work and synchronisation in between, equals the
    eager result bit for bit (FP32 batched tile kernel and FP64 single field)."""
    _case("graph")
    R, A = c2
    n_dst, n_src = A.shape
    k = 24
    x = torch.tensor(_fields(n_src, k, 13).T.copy(), dtype=torch.float32, device=DEV).T
    y = torch.zeros(k, n_dst, dtype=torch.float32, device=DEV).T
    x64 = torch.tensor(np.random.default_rng(14).random(n_src), device=DEV)
    y64 = torch.zeros(n_dst, dtype=torch.float64, device=DEV)
    f32 = cr.regrid_plan(y, R, x)      # first launches (plans built) outside the capture
    f64 = cr.regrid_plan(y64, R, x64)
    torch.cuda.synchronize()
    ref32, ref64 = y.clone(), y64.clone()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        f32()
        f64()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        f32()
        f64()
    for rep in range(3):
        y.zero_()
        y64.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, ref32), rep
        assert torch.equal(y64, ref64), rep
        _ = (x * 2.0).sum().item()  # eager work between replays


def test_capture_needs_prepared_set_up(cr):
    """A cr_regrid whose direction has no plan yet fails inside a capture with
    CR_UNSUPPORTED (its set-up allocates and synchronises); after
    cr_regrid_prepare the same call captures, and every replay equals the
    eager result."""
    _case("capture_prepare")
    L = cr._lib
    src, dst = cr.Grid.from_points(cr.lonlat_points(90, 45)), cr.Grid.from_points(cr.lonlat_points(72, 36))
    R = cr.Regridder(dst, src, radius=1.0)
    n_dst, n_src = R.shape
    k = 8
    x = torch.tensor(_fields(n_src, k, 3).T.copy(), dtype=torch.float32, device=DEV)  # (k, n): cells fastest
    y = torch.zeros(k, n_dst, dtype=torch.float32, device=DEV)
    h = R._core.h

    def call():
        return L.lib.cr_regrid(h, 0, L.CR_F32, k, x.data_ptr(), 1, n_src, y.data_ptr(), 1, n_dst, 1,
                               torch.cuda.current_stream().cuda_stream)

    status = []
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        status.append(call())
    assert status[0] == L.CR_UNSUPPORTED and "cr_regrid_prepare" in L.last_error()
    assert L.lib.cr_regrid_prepare(h, 0, L.CR_F32, k, n_src, n_dst, None) == L.CR_OK
    assert call() == L.CR_OK
    torch.cuda.synchronize()
    ref = y.clone()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2):
        status.append(call())
    assert status[1] == L.CR_OK
    for rep in range(3):
        y.zero_()
        torch.cuda.synchronize()
        g2.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, ref), rep
        _ = (x * 2.0).sum().item()  # eager work between replays
