"""Parity of the headline path at full size: the 0.25 deg lon-lat (1440 x 720)
-> HEALPix nside 256 (nested, 16-gon cells) regridder built on the device,
then

  * the batched FP32 apply of 256 seeded harmonic fields laid out [n, 64, 4]
    (cells fastest, dims=1), forward and through transpose(R), with the
    default kernel (variant 15: products and sums in FP64 from the FP64 A) and
    the pinned variants 17 (the same, 4 field chunks), 21 (A rounded to FP32),
    23 and 28 (21 and 15 on the block-ordered plan),
    against the oracle's FP64 SpMV on the oracle-built CSR for 8 sampled
    fields: relative field error <= 1e-6, conservation drift <= 1e-7;
  * one FP64 field forward and transposed against the oracle to 1e-12;

and the 4 deg -> 5 deg configuration (normalize=true, U[0,1) field, seed 0)
forward and transposed in FP64 to 1e-12.  Reference: regrid!/NDSliceLoop,
src/regridder/regrid.jl:63-118,303-318; transpose, regridder.jl:49-50.
"""
import numpy as np
import pytest
import torch

import bench

pytestmark = pytest.mark.gpu
DEV = "cuda"
SAMPLED = (0, 37, 64, 101, 128, 170, 203, 255)


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def c3(cr, orc):
    _case("c3_setup", src="0.25deg 1440x720", dst="healpix256 nested step4")
    sp_ = cr.lonlat_points(1440, 720)
    dst = cr.Grid.healpix(256, "nested", step=4)
    R = cr.Regridder(dst, cr.Grid.from_points(sp_), radius=1.0)
    ref = orc.build(dst.cells(), cr.points_to_cells(sp_), clip_is_dst=False)
    refT = ref.transpose()
    X2 = bench.synth_lonlat_fields(torch, 1440, 720, 256, 64, seed=0)  # (256, n_src) FP32
    return R, ref, refT, X2


def _fp32_report(got, x64, ref_csr, a_to, a_from):
    """Max relative field error and conservation drift of sampled fields."""
    worst_err = worst_drift = 0.0
    for f in SAMPLED:
        y_ref = ref_csr.apply(x64[f], nthreads=8)
        y = got[f].double().cpu().numpy()
        worst_err = max(worst_err, float(np.max(np.abs(y - y_ref)) / np.max(np.abs(y_ref))))
        drift = abs(float((y * a_to).sum() - (x64[f] * a_from).sum())) / float(np.abs(x64[f] * a_from).sum())
        worst_drift = max(worst_drift, drift)
    return worst_err, worst_drift


@pytest.mark.parametrize("variant", [-1, 17, 21, 23, 28])
def test_c3_batched_fp32_forward_and_transpose(cr, c3, variant):
    _case("c3_batched_fp32", variant=variant, fields=256, layout="[n,64,4]")
    R, ref, refT, X2 = c3
    n_dst, n_src = R.shape
    RT = cr.transpose(R)
    cr.regrid_set_variant(R, variant)
    cr.regrid_set_variant(RT, variant)
    try:
        X = X2.view(4, 64, n_src).permute(2, 1, 0)  # (n_src, 64, 4), strides (1, n, 64 n)
        Y2 = torch.zeros(256, n_dst, dtype=torch.float32, device=DEV)
        Y = Y2.view(4, 64, n_dst).permute(2, 1, 0)
        cr.regrid_(Y, R, X, dims=1)
        torch.cuda.synchronize()
        x64 = {f: X2[f].double().cpu().numpy() for f in SAMPLED}
        err, drift = _fp32_report(Y2, x64, ref, R.dst_areas, R.src_areas)
        print(f"c3 forward variant {variant}: rel err {err:.2e} drift {drift:.2e}", flush=True)
        assert err <= 1e-6 and drift <= 1e-7
        # transpose(R): the destination fields back to the source grid, same layout
        B2 = torch.zeros(256, n_src, dtype=torch.float32, device=DEV)
        B = B2.view(4, 64, n_src).permute(2, 1, 0)
        cr.regrid_(B, RT, Y, dims=1)
        torch.cuda.synchronize()
        y64 = {f: Y2[f].double().cpu().numpy() for f in SAMPLED}
        err, drift = _fp32_report(B2, y64, refT, R.src_areas, R.dst_areas)
        print(f"c3 transpose variant {variant}: rel err {err:.2e} drift {drift:.2e}", flush=True)
        assert err <= 1e-6 and drift <= 1e-7
    finally:
        cr.regrid_set_variant(R, -1)
        cr.regrid_set_variant(RT, -1)


def test_c3_fp64_single_field_forward_and_transpose(cr, c3):
    _case("c3_fp64_single")
    R, ref, refT, X2 = c3
    x = X2[5].double()
    y = torch.zeros(R.shape[0], dtype=torch.float64, device=DEV)
    cr.regrid_(y, R, x)
    y_ref = ref.apply(x.cpu().numpy(), nthreads=8)
    got = y.cpu().numpy()
    assert np.max(np.abs(got - y_ref)) <= 1e-12 * np.max(np.abs(y_ref))
    tot = float((x.cpu().numpy() * R.src_areas).sum())
    assert abs(float((got * R.dst_areas).sum()) - tot) <= 1e-12 * abs(tot)
    b = torch.zeros(R.shape[1], dtype=torch.float64, device=DEV)
    cr.regrid_(b, cr.transpose(R), y)
    b_ref = refT.apply(got, nthreads=8)
    assert np.max(np.abs(b.cpu().numpy() - b_ref)) <= 1e-12 * np.max(np.abs(b_ref))


def test_c1_fp64_normalized_forward_and_transpose(cr, orc):
    _case("c1_fp64", normalize=True, seed=0)
    sp_, dp_ = cr.lonlat_points(90, 45), cr.lonlat_points(72, 36)
    R = cr.Regridder(cr.Grid.from_points(dp_), cr.Grid.from_points(sp_), normalize=True, radius=1.0)
    ref = orc.build(cr.points_to_cells(dp_), cr.points_to_cells(sp_)).normalized()
    s = np.random.default_rng(0).random(90 * 45)
    d = torch.zeros(72 * 36, dtype=torch.float64, device=DEV)
    cr.regrid_(d, R, torch.tensor(s, device=DEV))
    d_ref = ref.apply(s)
    got = d.cpu().numpy()
    assert np.max(np.abs(got - d_ref)) <= 1e-12 * np.max(np.abs(d_ref))
    assert abs((got * R.dst_areas).sum() - (s * R.src_areas).sum()) <= 1e-12 * (s * R.src_areas).sum()
    back = torch.zeros(90 * 45, dtype=torch.float64, device=DEV)
    cr.regrid_(back, cr.transpose(R), d)
    b_ref = ref.transpose().apply(got)
    assert np.max(np.abs(back.cpu().numpy() - b_ref)) <= 1e-12 * np.max(np.abs(b_ref))
