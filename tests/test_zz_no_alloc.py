"""cr_regrid_backward_noalloc (scratch in a static device array, no allocation)
against cr_regrid_backward, through the C ABI.

Two calls in a row on the shared scratch -- a cubed sphere C12 -> 6 degree
regridder with 4 fields, then a smaller C6 -> 12 degree one with 1 field --
each compared with cr_regrid_backward on the same inputs: the field gradient
bit for bit (a gather with a fixed order), the vertex gradients (FP64 atomics,
order-dependent) to 1e-12 of their largest magnitude.  Also: a request over
the array's 64 MiB is rejected before anything runs, a missing adjoint plan is
reported, and a call after set-up leaves the device's free memory unchanged."""
import numpy as np
import pytest
import torch

from gpu_common import jittered_cubed_sphere

pytestmark = pytest.mark.gpu
DEV = "cuda"
F64 = 0


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _st(t):
    return (t.stride(0), t.stride(1) if t.ndim > 1 else 1)


def _problem(cr, ne, nlon, nlat, nf, seed):
    P = jittered_cubed_sphere(cr, ne, seed=seed, frac=0.05)
    d = 360.0 / nlon
    D = cr.lonlat_points(nlon, nlat, lon0=0.5, lat0=-90.0, dlon=d, dlat=180.0 / nlat)
    src, dst = cr.Grid.from_points(P), cr.Grid.from_points(D)
    R = cr.Regridder(dst, src, radius=1.0)
    rng = np.random.default_rng(seed)
    x = torch.tensor(rng.random((src.ncells, nf)), device=DEV) if nf > 1 else torch.tensor(
        rng.random(src.ncells), device=DEV)
    y = cr.regrid(R, x)
    gy = torch.tensor(rng.standard_normal(tuple(y.shape)), device=DEV)
    return R, src, dst, x, y, gy


def _outputs(src, dst, x):
    return (torch.full_like(x, np.nan), torch.zeros(dst.nverts, 3, dtype=torch.float64, device=DEV),
            torch.zeros(src.nverts, 3, dtype=torch.float64, device=DEV))


def _call(L, fn, R, src, dst, x, y, gy, nf, outs=None, verts=True):
    gx, g_dst, g_src = outs if outs is not None else _outputs(src, dst, x)
    (xc, xf), (yc, yf), (gc, gf), (gxc, gxf) = _st(x), _st(y), _st(gy), _st(gx)
    st = fn(R._core.h, dst._h, src._h, 0, 1, nf, x.data_ptr(), xc, xf, y.data_ptr(), yc, yf, gy.data_ptr(), gc, gf,
            gx.data_ptr(), gxc, gxf, g_dst.data_ptr() if verts else None, g_src.data_ptr() if verts else None,
            torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return st, gx, g_dst, g_src


def _prepare(L, R, nf, n_to, gx_field_stride):
    L.check(L.lib.cr_regrid_prepare(R._core.h, 1, F64, nf, n_to, gx_field_stride,
                                    torch.cuda.current_stream().cuda_stream), "prepare")


def _compare(L, R, src, dst, x, y, gy, nf):
    st, gx0, gd0, gs0 = _call(L, L.lib.cr_regrid_backward, R, src, dst, x, y, gy, nf)
    L.check(st, "backward")
    st, gx1, gd1, gs1 = _call(L, L.lib.cr_regrid_backward_noalloc, R, src, dst, x, y, gy, nf)
    L.check(st, "backward_noalloc")
    assert torch.equal(gx1, gx0), "field gradients differ"
    for a, b in ((gd1, gd0), (gs1, gs0)):
        scale = float(b.abs().max())
        assert scale > 0
        assert float((a - b).abs().max()) <= 1e-12 * scale


def test_noalloc_backward_matches_two_calls(cr):
    L = cr._lib
    _case("noalloc big", ne=12, nlon=60, nlat=30, nf=4)
    R, src, dst, x, y, gy = _problem(cr, 12, 60, 30, 4, seed=11)
    # the adjoint reads w (n_dst x nf, cells fastest) and writes gx
    _prepare(L, R, 4, dst.ncells, x.stride(1))
    _compare(L, R, src, dst, x, y, gy, 4)
    _case("noalloc small", ne=6, nlon=30, nlat=15, nf=1)
    R2, src2, dst2, x2, y2, gy2 = _problem(cr, 6, 30, 15, 1, seed=12)
    _prepare(L, R2, 1, dst2.ncells, 1)
    _compare(L, R2, src2, dst2, x2, y2, gy2, 1)
    # the device's free memory is unchanged by a call after set-up
    _case("noalloc free memory")
    outs = _outputs(src2, dst2, x2)
    torch.cuda.synchronize()
    before = torch.cuda.mem_get_info()[0]
    st, *_ = _call(L, L.lib.cr_regrid_backward_noalloc, R2, src2, dst2, x2, y2, gy2, 1, outs=outs)
    L.check(st, "backward_noalloc")
    assert torch.cuda.mem_get_info()[0] == before


def test_noalloc_backward_limits(cr):
    L = cr._lib
    _case("noalloc limits")
    R, src, dst, x, y, gy = _problem(cr, 6, 30, 15, 1, seed=13)
    # over the 64 MiB array: n_dst * nfields doubles; rejected before any launch
    # (x, y and gy are never read)
    nf_big = (64 << 20) // (8 * dst.ncells) + 1
    st = L.lib.cr_regrid_backward_noalloc(R._core.h, dst._h, src._h, 0, 1, nf_big, x.data_ptr(), 1, src.ncells,
                                          y.data_ptr(), 1, dst.ncells, gy.data_ptr(), 1, dst.ncells, x.data_ptr(),
                                          1, src.ncells, None, None, torch.cuda.current_stream().cuda_stream)
    assert st != 0 and "scratch" in L.last_error()
    # the adjoint's plan for 3 fields was never made: reported, nothing launched
    st = L.lib.cr_regrid_backward_noalloc(R._core.h, dst._h, src._h, 0, 1, 3, x.data_ptr(), 1, src.ncells,
                                          y.data_ptr(), 1, dst.ncells, gy.data_ptr(), 1, dst.ncells, x.data_ptr(),
                                          1, src.ncells, None, None, torch.cuda.current_stream().cuda_stream)
    assert st != 0 and "cr_regrid_prepare" in L.last_error()
