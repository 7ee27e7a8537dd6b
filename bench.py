"""Benchmark: Regridder build pairs/s and batched regrid! GB/s on one MI355X.

Workload (the configuration the headline metric is quoted on):
  source      0.25 degree lon-lat, 1440 x 720 cells (1441 x 721 vertices, poles included)
  destination HEALPix nside 256, nested, 786 432 cells as 16-gons (4 points per side)
  A           built in FP64 on the device
  fields      256 FP32 fields laid out [n, 64, 4] cells fastest (what regrid! gets with
              dims=1): seeded spherical-harmonic fields of degree <= 64 + 1 % uniform noise

A step is one batched regrid! of all 256 fields (one kernel launch); `value` is
its algorithmic GB/s with the fields resident in HBM.  Beside it: the plain-FP64
comparison (a loop of 256 single-field FP64 regrid! calls on the same GPU, warm
and with the caches flushed before each), graph replay against eager launches,
the achievable HBM rates of hand-written streaming kernels (libcrhbm.so), the
build timed from the host inputs (the source point matrix, the HEALPix nside),
from host vertex tables and from device-resident grids, the reverse mode, and
the CPU baseline: the FP64 C restatement in oracle/ on this host (build at 1
thread and all threads, single-field applies at 1 and all threads), which also
counts the FP64 operations behind the clip kernels' roofline figures.

  python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--skip-cpu]
  python bench.py --configs     # the other configurations, one JSON line each
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "Regridder build pairs/s; batched regrid! GB/s vs plain FP64, same GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def legendre_table(lmax: int, x: np.ndarray) -> np.ndarray:
    """4pi-normalised associated Legendre functions P[j, l, m] at x = sin(lat)
    (computed with the points as the fastest axis, then transposed)."""
    x = np.asarray(x, np.float64)
    u = np.sqrt(np.clip(1.0 - x * x, 0.0, None))
    P = np.zeros((lmax + 1, lmax + 1, x.shape[0]))
    P[0, 0] = 1.0
    for m in range(1, lmax + 1):
        f = math.sqrt(3.0) if m == 1 else math.sqrt((2 * m + 1) / (2 * m))
        P[m, m] = f * u * P[m - 1, m - 1]
    for m in range(0, lmax):
        P[m + 1, m] = math.sqrt(2 * m + 3) * x * P[m, m]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            a = math.sqrt((4 * l * l - 1) / (l * l - m * m))
            b = math.sqrt(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
            P[l, m] = a * (x * P[l - 1, m] - b * P[l - 2, m])
    return np.ascontiguousarray(np.moveaxis(P, 2, 0))


def synth_lonlat_fields(torch, nlon, nlat, nfields, lmax, seed, noise=0.01, dtype=None):
    """(nfields, nlat*nlon) on the device (FP32 unless dtype is given):
    sum_{l<=lmax, m} c_lm Pbar_lm(sin lat) (cos m lon, sin m lon) at cell centres,
    c ~ N(0,1)/sqrt(#terms), + noise*U[0,1)."""
    lat = np.radians(-90.0 + (np.arange(nlat) + 0.5) * 180.0 / nlat)
    lon = np.radians((np.arange(nlon) + 0.5) * 360.0 / nlon)
    P = torch.tensor(legendre_table(lmax, np.sin(lat)), device="cuda")  # (nlat, L, M)
    g = torch.Generator(device="cpu").manual_seed(seed)
    nterms = (lmax + 1) * (lmax + 2)
    ca = torch.randn(nfields, lmax + 1, lmax + 1, generator=g, dtype=torch.float64).tril().cuda()
    cb = torch.randn(nfields, lmax + 1, lmax + 1, generator=g, dtype=torch.float64).tril().cuda()
    ca, cb = ca / math.sqrt(nterms), cb / math.sqrt(nterms)
    m = torch.arange(lmax + 1, device="cuda", dtype=torch.float64)
    lon_t = torch.tensor(lon, device="cuda")
    cosm, sinm = torch.cos(m[:, None] * lon_t[None]), torch.sin(m[:, None] * lon_t[None])  # (M, nlon)
    out = torch.empty(nfields, nlat * nlon, dtype=dtype or torch.float32, device="cuda")
    u = torch.rand(nfields, nlat * nlon, generator=g, dtype=torch.float32).cuda()
    for f0 in range(0, nfields, 32):
        f1 = min(nfields, f0 + 32)
        A = torch.einsum("jlm,flm->fjm", P, ca[f0:f1])
        B = torch.einsum("jlm,flm->fjm", P, cb[f0:f1])
        F = torch.matmul(A, cosm) + torch.matmul(B, sinm)  # (f, nlat, nlon)
        out[f0:f1] = ((1.0 + F.reshape(f1 - f0, -1)) + noise * u[f0:f1].double()).to(out.dtype)
    return out


def synth_cell_fields(torch, centres, nfields, lmax, seed, noise=0.01, ring_z=None):
    """(nfields, n) FP32 on the device for cells with unit-vector centres
    (n, 3) arranged in rings of one latitude (ring_z: the rings' z = sin(lat),
    each cell taking the nearest; default: the distinct centre z values): the
    same harmonic sums as synth_lonlat_fields at (ring latitude, centre
    longitude), the Legendre functions evaluated once per ring."""
    z = np.clip(centres[:, 2], -1.0, 1.0)
    if ring_z is None:
        rings, ring_of = np.unique(np.round(z, 12), return_inverse=True)
    else:
        rings = np.sort(np.asarray(ring_z, np.float64))
        k = np.clip(np.searchsorted(rings, z), 1, rings.size - 1)
        ring_of = np.where(np.abs(z - rings[k - 1]) <= np.abs(rings[k] - z), k - 1, k)
    if rings.size > 65536:
        raise ValueError(f"{rings.size} distinct ring latitudes: pass ring_z")
    P = torch.tensor(legendre_table(lmax, rings), device="cuda")  # (rings, L, M)
    lon = torch.tensor(np.arctan2(centres[:, 1], centres[:, 0]), device="cuda", dtype=torch.float32)
    ring_t = torch.tensor(ring_of, device="cuda", dtype=torch.int64)
    g = torch.Generator(device="cpu").manual_seed(seed)
    nterms = (lmax + 1) * (lmax + 2)
    m = torch.arange(lmax + 1, device="cuda", dtype=torch.float32)
    out = torch.empty(nfields, centres.shape[0], dtype=torch.float32, device="cuda")
    chunk = 1 << 20
    for f in range(nfields):
        ca = torch.randn(lmax + 1, lmax + 1, generator=g, dtype=torch.float64).tril().cuda() / math.sqrt(nterms)
        cb = torch.randn(lmax + 1, lmax + 1, generator=g, dtype=torch.float64).tril().cuda() / math.sqrt(nterms)
        A = torch.einsum("jlm,lm->jm", P, ca).float()  # (rings, M)
        B = torch.einsum("jlm,lm->jm", P, cb).float()
        for c0 in range(0, centres.shape[0], chunk):
            c1 = min(centres.shape[0], c0 + chunk)
            ml = m[None, :] * lon[c0:c1, None]
            r = ring_t[c0:c1]
            out[f, c0:c1] = 1.0 + (A[r] * torch.cos(ml) + B[r] * torch.sin(ml)).sum(1)
        u = torch.rand(centres.shape[0], generator=g, dtype=torch.float32).cuda()
        out[f] += noise * u
    return out


def apply_bytes(nnz, n_rows, n_cols, k, elem):
    """Algorithmic HBM bytes of one apply: int32 col + FP64 val per non-zero, int64
    row pointers, one FP64 area (reciprocal) per row, k fields read and written."""
    return nnz * 12 + (n_rows + 1) * 8 + n_rows * 8 + k * (n_cols + n_rows) * elem


def read_traffic(name, variant):
    """HBM bytes per launch of the given kernel variant from a committed
    rocprofv3 PMC summary (lines ``traffic_bytes_per_launch[<variant>]: <bytes>``),
    or None when the summary has no figure for that variant."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch[{variant}]:"
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


def _time_ms(torch, fn, reps):
    """Mean device time of fn() over reps calls (HIP events on the current stream)."""
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _build_ms(cr, torch, dst, src, reps=3, **kw):
    """A regridder and the median wall time of `reps` further device builds."""
    R = cr.Regridder(dst, src, **kw)
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        Rb = cr.Regridder(dst, src, **kw)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
        del Rb
    return R, float(np.median(ts)) * 1e3


def _jittered_cubed_sphere(cr, ne, seed, frac=0.05):
    """Cubed-sphere points with every face-interior vertex moved in its tangent
    plane by N(0, frac * cell width), renormalised (cells stay convex)."""
    P = cr.cubed_sphere_points(ne).copy()
    rng = np.random.default_rng(seed)
    inner = P[:, 1:-1, 1:-1, :]
    noise = rng.normal(0.0, frac * (np.pi / 2 / ne), size=inner.shape)
    noise -= np.sum(noise * inner, axis=-1, keepdims=True) * inner
    moved = inner + noise
    P[:, 1:-1, 1:-1, :] = moved / np.linalg.norm(moved, axis=-1, keepdims=True)
    return P


def host_info():
    """nproc, the CPU share this process may use and the CPU model."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count(), "affinity": share, "cpu_model": model}


def cpu_threads():
    """Threads for the all-cores CPU legs: the CPU share of this process, at most 16
    (a one-GPU box's share of its host)."""
    return max(1, min(16, host_info()["affinity"]))


class HbmProbe:
    """The streaming read / copy / write kernels of libcrhbm.so: the achievable
    HBM rates, the ceiling the apply's roofline is read against."""

    def __init__(self, torch):
        import ctypes
        path = os.path.join(ROOT, "conservativeregridding.jl_amd", "libcrhbm.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        self.lib = ctypes.CDLL(path)
        self.lib.crhbm_launch.restype = ctypes.c_int
        self.lib.crhbm_launch.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                          ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        self.torch = torch
        self.sink = torch.zeros(256, dtype=torch.float32, device="cuda")

    def launch(self, kind, a, b, nbytes, blocks):
        st = self.lib.crhbm_launch(kind, a.data_ptr() if a is not None else None,
                                   b.data_ptr() if b is not None else None, nbytes, self.sink.data_ptr(), blocks,
                                   self.torch.cuda.current_stream().cuda_stream)
        if st != 0:
            raise RuntimeError(f"crhbm_launch({kind}) failed: hipError {st}")

    def flush(self, buf):
        """Overwrite buf (larger than every cache level) to evict earlier data."""
        self.launch(2, None, buf, buf.numel() * buf.element_size(), 4096)

    def ceilings(self):
        """Best GB/s over a few grid sizes: read 2 GiB, copy 1 GiB -> 1 GiB (bytes
        read + written), write 1 GiB."""
        torch = self.torch
        a = torch.ones(1 << 29, dtype=torch.float32, device="cuda")  # 2 GiB
        b = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        half = 1 << 30
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        out = {}
        for name, kind, nbytes, moved in (("read_GBps", 0, 2 * half, 2 * half), ("copy_GBps", 1, half, 2 * half),
                                          ("write_GBps", 2, half, half)):
            best = 0.0
            for blocks in (1024, 2048, 4096, 8192):
                self.launch(kind, a, b, nbytes, blocks)
                torch.cuda.synchronize()
                e0.record()
                for _ in range(5):
                    self.launch(kind, a, b, nbytes, blocks)
                e1.record()
                torch.cuda.synchronize()
                best = max(best, moved / (e0.elapsed_time(e1) / 5 * 1e-3) / 1e9)
            out[name] = round(best, 1)
        del a, b
        return out


def _drift(y, a_to, x, a_from):
    """Conservation drift |sum(y a_to) - sum(x a_from)| / sum(|x a_from|) (host FP64)."""
    return abs(float((y * a_to).sum() - (x * a_from).sum())) / float(np.abs(x * a_from).sum())


def report_configs(cr, torch, reps, with_cpu):
    """The reference's other configurations, measured on one GPU (one JSON line
    each): build pairs/s from host vertex arrays, single-field FP64 applies
    forward and transposed (warm and with the caches flushed), the reverse
    mode, and the largest configuration's 64-field FP32 batched apply forward
    and transposed (plain launches and a graph replay); with the configured
    fields.  with_cpu: the CPU leg per configuration -- the oracle's build
    timed at all cores (and 1 thread where it takes seconds) and its apply,
    and the device results' error and drift against it."""
    out = []
    orc = None
    if with_cpu:
        import __graft_entry__ as ge
        orc = ge.load_oracle()
    nth = cpu_threads()
    probe = HbmProbe(torch)
    flush_buf = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB

    def build_from_host(dst_fn, src_fn, reps_b=3, **kw):
        """Median wall time of `reps_b` builds that start from host vertex
        arrays (upload, input checks, build), and of the build alone from the
        device-resident grids (in `dev_ms`)."""
        ts, tb = [], []
        R = None
        for i in range(reps_b + 1):
            torch.cuda.synchronize()
            t = time.perf_counter()
            d_, s_ = dst_fn(), src_fn()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            Rb = cr.Regridder(d_, s_, **kw)
            torch.cuda.synchronize()
            if i:
                ts.append(time.perf_counter() - t)
                tb.append(time.perf_counter() - t1)
            R, dst_g, src_g = Rb, d_, s_
        dev_ms[0] = float(np.median(tb)) * 1e3
        return R, dst_g, src_g, float(np.median(ts)) * 1e3

    dev_ms = [0.0]

    def single_field(name, R, x, ref=None):
        """FP64 forward and transposed applies of field x: warm device time of
        back-to-back launches, cold time (caches flushed before each launch),
        drift, and the error against the oracle's apply of its own CSR."""
        nd, ns = R.shape
        y = torch.zeros(nd, dtype=torch.float64, device="cuda")
        RT = cr.transpose(R)
        xt = torch.zeros(ns, dtype=torch.float64, device="cuda")
        f_fwd, f_bwd = cr.regrid_plan(y, R, x), cr.regrid_plan(xt, RT, y)
        fwd, bwd = _time_ms(torch, f_fwd, reps), _time_ms(torch, f_bwd, reps)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        cold = []
        for fn in (f_fwd, f_bwd):
            ts = []
            for _ in range(5):
                probe.flush(flush_buf)
                e0.record()
                fn()
                e1.record()
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1))
            cold.append(float(np.median(ts)))
        f_fwd()
        f_bwd()
        torch.cuda.synchronize()
        a_d, a_s = np.asarray(R.dst_areas), np.asarray(R.src_areas)
        xh, yh, xth = x.cpu().numpy(), y.cpu().numpy(), xt.cpu().numpy()
        b = apply_bytes(R.nnz, nd, ns, 1, 8), apply_bytes(R.nnz, ns, nd, 1, 8)
        rec = {"fp64_forward_us": round(fwd * 1e3, 2), "fp64_forward_GBps": round(b[0] / (fwd * 1e-3) / 1e9, 1),
               "fp64_forward_cold_us": round(cold[0] * 1e3, 2),
               "fp64_transpose_us": round(bwd * 1e3, 2), "fp64_transpose_GBps": round(b[1] / (bwd * 1e-3) / 1e9, 1),
               "fp64_transpose_cold_us": round(cold[1] * 1e3, 2),
               "fp64_drift": _drift(yh, a_d, xh, a_s)}
        if ref is not None:
            y_ref = ref.apply(xh, nthreads=nth)
            xt_ref = ref.transpose().apply(y_ref, nthreads=nth)
            rec["fp64_forward_rel_err_vs_oracle"] = float(np.max(np.abs(yh - y_ref)) / np.max(np.abs(y_ref)))
            rec["fp64_transpose_rel_err_vs_oracle"] = float(np.max(np.abs(xth - xt_ref)) / np.max(np.abs(xt_ref)))
        log(f"[configs] {name}: FP64 forward {fwd * 1e3:.1f} us warm / {cold[0] * 1e3:.1f} us cold "
            f"({b[0] / (fwd * 1e-3) / 1e9:.0f} GB/s warm), transpose {bwd * 1e3:.1f} / {cold[1] * 1e3:.1f} us; "
            f"drift {rec['fp64_drift']:.1e}"
            + (f"; vs oracle {rec['fp64_forward_rel_err_vs_oracle']:.1e} / "
               f"{rec['fp64_transpose_rel_err_vs_oracle']:.1e}" if ref is not None else ""))
        return rec

    def cpu_leg(name, dst_cells, src_cells, clip_is_dst, one_thread, R):
        """The oracle build at all cores (and 1 thread) and its single-field apply."""
        t = time.perf_counter()
        ref = orc.build(dst_cells, src_cells, clip_is_dst=clip_is_dst, nthreads=nth)
        t_all = time.perf_counter() - t
        rec = {"threads": nth, "build_s": round(t_all, 3), "build_pairs_per_s": round(ref.nnz / t_all, 1),
               "nnz": ref.nnz, "device_nnz": R.nnz}
        if one_thread:
            t = time.perf_counter()
            orc.build(dst_cells, src_cells, clip_is_dst=clip_is_dst, nthreads=1)
            t1 = time.perf_counter() - t
            rec.update({"build_1thread_s": round(t1, 3), "build_1thread_pairs_per_s": round(ref.nnz / t1, 1)})
        x = np.random.default_rng(1).random(ref.n_src)
        t = time.perf_counter()
        for _ in range(5):
            ref.apply(x, nthreads=1)
        ta = (time.perf_counter() - t) / 5
        rec["apply_1thread_GBps"] = round(apply_bytes(ref.nnz, ref.n_dst, ref.n_src, 1, 8) / ta / 1e9, 3)
        log(f"[configs] {name} CPU oracle: build {t_all:.3f} s at {nth} threads "
            f"({ref.nnz / t_all / 1e6:.1f} M pairs/s)"
            + (f", {rec['build_1thread_s']:.3f} s at 1 thread" if one_thread else "")
            + f"; apply {rec['apply_1thread_GBps']:.2f} GB/s (1 thread)")
        return ref, rec

    # ---- 4 deg -> 5 deg lon-lat, normalize=true; one U[0,1) field, seed 0
    sp4, dp5 = cr.lonlat_points(90, 45), cr.lonlat_points(72, 36)
    R, dst, src, ms = build_from_host(lambda: cr.Grid.from_points(dp5), lambda: cr.Grid.from_points(sp4),
                                      normalize=True)
    log(f"[configs] 4deg -> 5deg lon-lat: build {ms:.2f} ms from host arrays, nnz {R.nnz}")
    rec = {"workload": "4deg lon-lat 90x45 -> 5deg lon-lat 72x36, normalize=true; U[0,1) field, seed 0",
           "nnz": R.nnz, "build_ms_from_host": round(ms, 3), "build_pairs_per_s": round(R.nnz / (ms * 1e-3), 1),
           "build_ms_device_grids": round(dev_ms[0], 3)}
    ref = None
    if orc is not None:
        ref, rec["cpu_baseline"] = cpu_leg("4deg -> 5deg", cr.points_to_cells(dp5), cr.points_to_cells(sp4),
                                           True, True, R)
        ref = ref.normalized()
    x = torch.tensor(np.random.default_rng(0).random(90 * 45), device="cuda")
    rec.update(single_field("4deg -> 5deg", R, x, ref))
    out.append(rec)

    # ---- 1 deg lon-lat -> C48; one FP64 field of spherical harmonics up to degree 16
    sp1, cpts = cr.lonlat_points(360, 180), cr.cubed_sphere_points(48)
    R, dst, src, ms = build_from_host(lambda: cr.Grid.from_points(cpts), lambda: cr.Grid.from_points(sp1))
    log(f"[configs] 1deg lon-lat -> C48: build {ms:.2f} ms from host arrays, nnz {R.nnz}")
    rec = {"workload": "1deg lon-lat 360x180 -> cubed sphere C48; FP64 field of degree <= 16 harmonics, "
                       "N(0,1) coefficients, seed 0",
           "nnz": R.nnz, "build_ms_from_host": round(ms, 3), "build_pairs_per_s": round(R.nnz / (ms * 1e-3), 1),
           "build_ms_device_grids": round(dev_ms[0], 3)}
    ref = None
    if orc is not None:
        ref, rec["cpu_baseline"] = cpu_leg("1deg -> C48", cr.points_to_cells(cpts), cr.points_to_cells(sp1),
                                           True, True, R)
    x = synth_lonlat_fields(torch, 360, 180, 1, 16, seed=0, noise=0.0, dtype=torch.float64)[0].contiguous()
    rec.update(single_field("1deg -> C48", R, x, ref))
    out.append(rec)

    # ---- jittered C96 -> 1 deg lon-lat: FP64 field and the reverse mode
    jp = _jittered_cubed_sphere(cr, 96, seed=4)
    R, dst, src, ms = build_from_host(lambda: cr.Grid.from_points(sp1), lambda: cr.Grid.from_points(jp))
    ns, nd = 6 * 96 * 96, 360 * 180
    rec = {"workload": "jittered cubed sphere C96 -> 1deg lon-lat; FP64 U[0,1) field and target, reverse mode",
           "nnz": R.nnz, "build_ms_from_host": round(ms, 3), "build_pairs_per_s": round(R.nnz / (ms * 1e-3), 1),
           "build_ms_device_grids": round(dev_ms[0], 3)}
    ref = None
    if orc is not None:
        ref, rec["cpu_baseline"] = cpu_leg("C96 -> 1deg", cr.points_to_cells(sp1), cr.points_to_cells(jp),
                                           True, True, R)
    rng = np.random.default_rng(4)
    x = torch.tensor(rng.random(ns), device="cuda")
    t_target = torch.tensor(rng.random(nd), device="cuda")
    rec.update(single_field("C96 -> 1deg", R, x, ref))
    y = cr.regrid(R, x)
    gy = 2.0 * (y - t_target)
    ms_bwd = _time_ms(torch, lambda: cr.regrid_backward(R, dst, src, x, y, gy), reps)
    log(f"[configs] jittered C96 -> 1deg: build {ms:.2f} ms, nnz {R.nnz}; backward (field + vertex gradients) "
        f"{ms_bwd:.3f} ms ({R.nnz / (ms_bwd * 1e-3) / 1e6:.0f} M pairs/s)")
    rec.update({"backward_ms": round(ms_bwd, 4), "backward_pairs_per_s": round(R.nnz / (ms_bwd * 1e-3), 1)})
    out.append(rec)
    del R, x, y, gy

    # ---- HEALPix nside 1024 -> 0.1 deg lon-lat: 64 FP32 fields (harmonics of
    # degree <= 256 + 1 % noise, synthesised on the device), forward and transposed
    hsrc = cr.Grid.healpix(1024, "nested", step=1)
    hcells = hsrc.cells()  # host table: the CPU leg's input and the field synthesis' centres
    del hsrc
    dp01 = cr.lonlat_points(3600, 1800)
    # host inputs: the destination point matrix and the HEALPix nside (cells made on the device)
    R, dst, src, ms = build_from_host(lambda: cr.Grid.from_points(dp01),
                                      lambda: cr.Grid.healpix(1024, "nested", step=1), reps_b=1)
    ns, nd, k = src.ncells, dst.ncells, 64
    log(f"[configs] HEALPix 1024 -> 0.1deg: build {ms:.1f} ms from host arrays, nnz {R.nnz}")
    centres = hcells.sum(axis=1)
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    t = time.perf_counter()
    nsd = 1024  # HEALPix ring latitudes: z = 1 - i^2/(3 n^2) in the caps, 4/3 - 2i/(3n) between
    ring = np.arange(1, 4 * nsd)
    ring_z = np.where(ring < nsd, 1.0 - ring ** 2 / (3.0 * nsd ** 2),
                      np.where(ring <= 3 * nsd, 4.0 / 3.0 - 2.0 * ring / (3.0 * nsd),
                               -(1.0 - (4 * nsd - ring) ** 2 / (3.0 * nsd ** 2))))
    X = synth_cell_fields(torch, centres, k, 256, seed=0, ring_z=ring_z)  # (64, ns) FP32
    torch.cuda.synchronize()
    log(f"[configs] HEALPix 1024 -> 0.1deg: 64 fields of degree <= 256 synthesised in {time.perf_counter() - t:.1f} s")
    del centres
    Y = torch.zeros(k, nd, dtype=torch.float32, device="cuda")
    Xc, Yc = X.t(), Y.t()  # cells fastest: (n, 64) with cell stride 1
    RT = cr.transpose(R)
    f_fwd = cr.regrid_plan(Yc, R, Xc)
    Z = torch.zeros(k, ns, dtype=torch.float32, device="cuda")
    f_bwd = cr.regrid_plan(Z.t(), RT, Yc)
    ms_f = _time_ms(torch, f_fwd, reps)  # the default variant
    ms_t = _time_ms(torch, f_bwd, reps)
    # accuracy of the default variants: FP32 against the device FP64 apply of the same fields
    f_fwd()
    f_bwd()
    torch.cuda.synchronize()
    a_d = torch.tensor(R.dst_areas, device="cuda")
    a_s = torch.tensor(R.src_areas, device="cuda")
    err_f = err_t = drift_f = 0.0
    for f in range(0, k, 8):
        x64 = X[f].double()
        y64 = cr.regrid(R, x64)
        z64 = cr.regrid(RT, Y[f].double())
        err_f = max(err_f, float((Y[f].double() - y64).abs().max() / y64.abs().max()))
        err_t = max(err_t, float((Z[f].double() - z64).abs().max() / z64.abs().max()))
        drift_f = max(drift_f, float(((Y[f].double() * a_d).sum() - (x64 * a_s).sum()).abs()
                                     / (x64 * a_s).abs().sum()))
    fp64_vs_oracle = None
    if orc is not None:  # the device FP64 apply against the oracle's SpMV of the same CSR (2 fields)
        rp, col, val = R.csr_arrays()
        fp64_vs_oracle = 0.0
        for f in (0, 63):
            y_ref = orc.spmv(rp, col, val, X[f].double().cpu().numpy(), R.dst_areas, nth)
            y64 = cr.regrid(R, X[f].double()).cpu().numpy()
            fp64_vs_oracle = max(fp64_vs_oracle, float(np.max(np.abs(y64 - y_ref)) / np.max(np.abs(y_ref))))
    # the autotuner's pick per direction
    log(f"[configs] HEALPix 1024 -> 0.1deg: default applies timed, FP32 error {err_f:.1e}; tuning")
    v_f, _ = cr.regrid_tune(Yc, R, Xc)
    v_t, _ = cr.regrid_tune(Z.t(), RT, Yc)
    ms_ft = _time_ms(torch, f_fwd, reps)  # the prepared launches now run the tuned variants
    ms_tt = _time_ms(torch, f_bwd, reps)
    # the tuned forward captured in a HIP graph, 10 launches per graph
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        f_fwd()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        for _ in range(10):
            f_fwd()
    ms_g = _time_ms(torch, graph.replay, max(1, reps // 10)) / 10
    bf, bt = apply_bytes(R.nnz, nd, ns, k, 4), apply_bytes(R.nnz, ns, nd, k, 4)
    x0 = X[0].double().contiguous()
    y0 = cr.regrid(R, x0)
    gy0 = torch.ones_like(y0)
    ms_bwd = _time_ms(torch, lambda: cr.regrid_backward(R, dst, src, x0, y0, gy0), max(1, reps // 2))
    log(f"[configs] HEALPix 1024 -> 0.1deg, 64 FP32 fields, default variant: forward {ms_f:.3f} ms "
        f"({bf / (ms_f * 1e-3) / 1e9:.0f} GB/s), transpose {ms_t:.3f} ms ({bt / (ms_t * 1e-3) / 1e9:.0f} GB/s); "
        f"autotuned: forward variant {v_f} {ms_ft:.3f} ms ({bf / (ms_ft * 1e-3) / 1e9:.0f} GB/s, graph replay "
        f"{ms_g:.3f} ms), transpose variant {v_t} {ms_tt:.3f} ms ({bt / (ms_tt * 1e-3) / 1e9:.0f} GB/s); "
        f"FP32 error {err_f:.1e} / {err_t:.1e}, drift {drift_f:.1e}; "
        f"backward (one FP64 field, vertex gradients) {ms_bwd:.2f} ms")
    rec = {"workload": "HEALPix nside 1024 nested -> 0.1deg lon-lat 3600x1800; 64 FP32 fields of degree <= 256 "
                       "harmonics + 1% noise", "nnz": R.nnz,
           "build_ms_from_host": round(ms, 3), "build_pairs_per_s": round(R.nnz / (ms * 1e-3), 1),
           "build_ms_device_grids": round(dev_ms[0], 3),
           "fp32_forward_ms": round(ms_f, 4), "fp32_forward_GBps": round(bf / (ms_f * 1e-3) / 1e9, 1),
           "fp32_transpose_ms": round(ms_t, 4), "fp32_transpose_GBps": round(bt / (ms_t * 1e-3) / 1e9, 1),
           "fp32_rel_err_vs_fp64": {"forward": err_f, "transpose": err_t}, "fp32_drift": drift_f,
           "fp64_rel_err_vs_oracle_spmv": fp64_vs_oracle,
           "tuned": {"forward_variant": v_f, "forward_ms": round(ms_ft, 4),
                     "forward_GBps": round(bf / (ms_ft * 1e-3) / 1e9, 1), "transpose_variant": v_t,
                     "transpose_ms": round(ms_tt, 4), "transpose_GBps": round(bt / (ms_tt * 1e-3) / 1e9, 1),
                     "forward_graph_ms": round(ms_g, 4)},
           "default_vs_tuned": {"forward": round(ms_f / ms_ft, 4), "transpose": round(ms_t / ms_tt, 4)},
           "backward_ms": round(ms_bwd, 3)}
    if orc is not None:
        t = time.perf_counter()
        refc = orc.build(cr.points_to_cells(dp01), hcells, clip_is_dst=True, nthreads=nth)
        tb = time.perf_counter() - t
        rec["cpu_baseline"] = {"threads": nth, "build_s": round(tb, 3), "build_pairs_per_s": round(refc.nnz / tb, 1),
                               "nnz": refc.nnz, "device_nnz": R.nnz}
        log(f"[configs] HEALPix 1024 -> 0.1deg CPU oracle: build {tb:.2f} s at {nth} threads")
        del refc
    out.append(rec)
    for rec in out:
        print(json.dumps({"configs_report": rec}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--skip-cpu", action="store_true", help="omit the CPU baseline leg")
    ap.add_argument("--builds", type=int, default=3, help="timed device builds")
    ap.add_argument("--configs", action="store_true",
                    help="measure the other reference configurations instead of the headline step")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the apply would shard "
              "by fields, each GPU holding all of A and k/G fields, with no communication)", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as ge
    cr = ge.load_package()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.cuda.get_device_name(0)
    log(f"[bench] device {dev}; library {cr.version()}")
    if args.configs:
        report_configs(cr, torch, max(10, args.steps), not args.skip_cpu)
        return

    NLON, NLAT, NSIDE, STEP, K = 1440, 720, 256, 4, 256
    t0 = time.time()
    src_pts = cr.lonlat_points(NLON, NLAT)
    src = cr.Grid.from_points(src_pts)
    dst = cr.Grid.healpix(NSIDE, "nested", step=STEP)
    torch.cuda.synchronize()
    log(f"[bench] grids ready ({time.time() - t0:.1f} s): src {src.ncells} cells, dst {dst.ncells} cells")

    # ---- build ---------------------------------------------------------------
    R = cr.Regridder(dst, src)  # warm-up build (also the regridder used below)
    times = []
    st = None
    for _ in range(args.builds):
        torch.cuda.synchronize()
        tb = time.perf_counter()
        Rb = cr.Regridder(dst, src)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - tb)
        st = Rb.stats()  # device stage times of a steady-state build
        del Rb
    st = st or R.stats()
    build_s = float(np.median(times))
    nnz, ncand = R.nnz, st["n_candidates"]
    log(f"[bench] build from device-resident grids: nnz={nnz} candidates={ncand} wall={build_s * 1e3:.1f} ms "
        f"(stages ms: bounds {st['ms_bounds']:.2f} tree {st['ms_tree']:.2f} search {st['ms_search']:.2f} "
        f"clip {st['ms_clip']:.2f} assemble {st['ms_assemble']:.2f})")
    # ---- fields --------------------------------------------------------------
    n_dst, n_src = R.shape
    X2 = synth_lonlat_fields(torch, NLON, NLAT, K, 64, seed=0)  # (256, n_src) = Julia [n, 64, 4]
    X = X2.view(4, 64, n_src).permute(2, 1, 0)  # (n_src, 64, 4), strides (1, n, 64n)
    Y2 = torch.zeros(K, n_dst, dtype=torch.float32, device="cuda")
    Y = Y2.view(4, 64, n_dst).permute(2, 1, 0)
    torch.cuda.synchronize()

    # ---- batched FP32 apply (the step) ----------------------------------------
    # every FP32 kernel variant once (5 launches each, median), then the autotuner's pick
    # the first apply of a direction also builds its tile plan (once per regridder)
    torch.cuda.synchronize()
    tf = time.perf_counter()
    cr.regrid_(Y, R, X, dims=1)
    torch.cuda.synchronize()
    first_apply_ms = (time.perf_counter() - tf) * 1e3
    # a second regridder's first apply: its plan without the process's one-time
    # loading of the planner's code object
    R3 = cr.Regridder(dst, src)
    torch.cuda.synchronize()
    tf = time.perf_counter()
    cr.regrid_(Y, R3, X, dims=1)
    torch.cuda.synchronize()
    first_apply2_ms = (time.perf_counter() - tf) * 1e3
    del R3
    log(f"[bench] first batched apply (with its one-time tile plan): {first_apply_ms:.1f} ms in this process, "
        f"{first_apply2_ms:.1f} ms for the next regridder")
    # the same build from the host inputs of the reference's call: the source
    # point matrix (uploaded and checked) and the HEALPix nside (its 16-gons
    # generated on the device, as the reference's HealpixRootNode makes its
    # cells); and from host vertex tables of both grids (the destination's
    # 302 MB of 16-gons uploaded and checked too)
    dst_cells_host = dst.cells()

    def timed_builds(make_grids):
        """Median wall time from the host inputs to the regridder; the grids and
        the regridder are released after the clock stops."""
        ts = []
        for i in range(args.builds + 1):
            torch.cuda.synchronize()
            tb = time.perf_counter()
            g_dst, g_src = make_grids()
            Rh = cr.Regridder(g_dst, g_src)
            torch.cuda.synchronize()
            if i:
                ts.append(time.perf_counter() - tb)
            assert Rh.nnz == nnz, (Rh.nnz, nnz)
            del Rh, g_dst, g_src
        return float(np.median(ts))

    build_host_s = timed_builds(lambda: (cr.Grid.healpix(NSIDE, "nested", step=STEP), cr.Grid.from_points(src_pts)))
    build_tables_s = timed_builds(lambda: (cr.Grid.from_cells(dst_cells_host, convex=False),
                                           cr.Grid.from_points(src_pts)))
    log(f"[bench] build from host inputs (source points, HEALPix nside): {build_host_s * 1e3:.1f} ms "
        f"({nnz / build_host_s / 1e9:.3f} G pairs/s); from host vertex tables of both grids: "
        f"{build_tables_s * 1e3:.1f} ms")

    variant_ms = {}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for v in range(cr.regrid_variants("float32")):
        cr.regrid_set_variant(R, v)
        fn = cr.regrid_plan(Y, R, X, dims=1)
        ts = []
        for _ in range(5):
            ev[0].record()
            fn()
            ev[1].record()
            torch.cuda.synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        variant_ms[v] = round(float(np.median(ts)), 4)
    cr.regrid_set_variant(R, -1)
    log("[bench] FP32 variants (ms): " + ", ".join(f"{k}: {v:.3f}" for k, v in variant_ms.items()))
    variant, tune_ms = cr.regrid_tune(Y, R, X, dims=1)  # picks the kernel variant once
    log(f"[bench] autotuned cells-fastest variant {variant} ({tune_ms:.4f} ms)")
    launch = cr.regrid_plan(Y, R, X, dims=1)
    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    tw = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    wall_steps = time.perf_counter() - tw
    ms_step = e0.elapsed_time(e1) / args.steps
    B32 = apply_bytes(nnz, n_dst, n_src, K, 4)
    gbps = B32 / (ms_step * 1e-3) / 1e9
    log(f"[bench] batched FP32 apply: {ms_step:.4f} ms/step, {gbps:.0f} GB/s algorithmic "
        f"({B32 / 1e9:.3f} GB/launch), wall {wall_steps:.3f} s")

    # the same launch captured in a HIP graph (torch.cuda.CUDAGraph), replayed
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        launch()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    per_graph = 10
    with torch.cuda.graph(graph):
        for _ in range(per_graph):
            launch()
    graph.replay()
    torch.cuda.synchronize()
    # eager launches and graph replays, interleaved over 5 rounds of the same
    # number of steps so that both see the same device state; medians
    n_replays = max(1, args.steps // per_graph)
    eager_r, graph_r = [], []
    for _ in range(5):
        e0.record()
        for _ in range(n_replays * per_graph):
            launch()
        e1.record()
        torch.cuda.synchronize()
        eager_r.append(e0.elapsed_time(e1) / (n_replays * per_graph))
        e0.record()
        for _ in range(n_replays):
            graph.replay()
        e1.record()
        torch.cuda.synchronize()
        graph_r.append(e0.elapsed_time(e1) / (n_replays * per_graph))
    ms_graph, ms_eager_il = float(np.median(graph_r)), float(np.median(eager_r))
    log(f"[bench] batched FP32 apply, graph replay ({per_graph} launches per graph): {ms_graph:.4f} ms/step; "
        f"eager launches interleaved with it: {ms_eager_il:.4f} ms/step")

    # ---- loop of 256 plain FP64 single-field applies --------------------------
    # Timed for each FP64 kernel variant (row kernel, tile kernel); the fastest
    # is the plain-FP64 figure the batched launch is compared with.
    X64 = X2.double()
    Y64 = torch.zeros(K, n_dst, dtype=torch.float64, device="cuda")
    B64 = apply_bytes(nnz, n_dst, n_src, 1, 8)
    loop_ms = {}
    for v, vname in enumerate(("rows", "tiles")[:cr.regrid_variants("float64")]):
        cr.regrid_set_variant(R, v, "float64")
        loop = [cr.regrid_plan(Y64[f], R, X64[f]) for f in range(K)]
        for fn in loop:
            fn()
        torch.cuda.synchronize()
        reps = 3
        e0.record()
        for _ in range(reps):
            for fn in loop:
                fn()
        e1.record()
        torch.cuda.synchronize()
        loop_ms[vname] = e0.elapsed_time(e1) / reps
        log(f"[bench] FP64 loop of {K} single-field applies ({vname}): {loop_ms[vname]:.3f} ms, "
            f"{K * B64 / (loop_ms[vname] * 1e-3) / 1e9:.0f} GB/s")
    best64 = min(loop_ms, key=loop_ms.get)
    cr.regrid_set_variant(R, list(loop_ms).index(best64), "float64")
    for f in range(K):  # Y64 from the fastest variant
        cr.regrid_(Y64[f], R, X64[f])
    ms_loop = loop_ms[best64]
    gbps_loop = K * B64 / (ms_loop * 1e-3) / 1e9
    log(f"[bench] fastest FP64 loop ({best64}) {ms_loop:.3f} ms; batched speed-up {ms_loop / ms_step:.2f}x")
    # the same single-field applies with every cache level flushed before each
    # (a 1 GiB overwrite between launches; A no longer stays in the 256 MB MALL)
    probe = HbmProbe(torch)
    flush_buf = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    cold = []
    for f in range(32):
        fn = cr.regrid_plan(Y64[f], R, X64[f])
        probe.flush(flush_buf)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        cold.append(e0.elapsed_time(e1))
    ms_cold_field = float(np.median(cold))
    log(f"[bench] FP64 single-field apply with flushed caches: {ms_cold_field * 1e3:.1f} us "
        f"({B64 / (ms_cold_field * 1e-3) / 1e9:.0f} GB/s); a loop of {K}: {K * ms_cold_field:.3f} ms, "
        f"batched speed-up {K * ms_cold_field / ms_step:.2f}x")

    # ---- reverse mode: field adjoint + vertex gradients of one FP64 field -------
    x0, y0 = X64[0].contiguous(), Y64[0].contiguous()
    gy0 = torch.ones_like(y0)
    cr.regrid_backward(R, dst, src, x0, y0, gy0)  # warm-up (plans of the transpose direction)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(3):
        cr.regrid_backward(R, dst, src, x0, y0, gy0)
    e1.record()
    torch.cuda.synchronize()
    ms_bwd = e0.elapsed_time(e1) / 3
    e0.record()
    for _ in range(3):
        cr.regrid_backward(R, dst, src, x0, y0, gy0, field_grad=False)
    e1.record()
    torch.cuda.synchronize()
    ms_vert = e0.elapsed_time(e1) / 3
    log(f"[bench] backward (field adjoint + vertex gradients, one FP64 field): {ms_bwd:.3f} ms "
        f"({nnz / (ms_bwd * 1e-3) / 1e6:.0f} M pairs/s); vertex gradients alone {ms_vert:.3f} ms")

    # ---- achievable HBM rates: hand-written streaming kernels ------------------
    del flush_buf
    ceil = probe.ceilings()
    # the apply moves mostly reads (sources, A) and writes each output once: its
    # ceiling is the read and write rates weighted by its algorithmic bytes
    out_bytes = K * n_dst * 4
    in_bytes = B32 - out_bytes
    mix_gbps = B32 / (in_bytes / ceil["read_GBps"] + out_bytes / ceil["write_GBps"])
    copy_gbps = ceil["copy_GBps"]
    log(f"[bench] streaming kernels: read {ceil['read_GBps']:.0f}, copy {ceil['copy_GBps']:.0f}, write "
        f"{ceil['write_GBps']:.0f} GB/s; the apply's read/write mix at those rates {mix_gbps:.0f} GB/s; "
        f"batched apply at {gbps / mix_gbps:.1%} of it, {gbps / HBM_PEAK_GBPS:.1%} of the published "
        f"{HBM_PEAK_GBPS:.0f} GB/s")

    # ---- accuracy vs FP64 -----------------------------------------------------
    Yd = Y2.double()
    scale = Y64.abs().amax(dim=1)
    rel_err = float(((Yd - Y64).abs().amax(dim=1) / scale).max())
    a_d = torch.tensor(R.dst_areas, device="cuda")
    a_s = torch.tensor(R.src_areas, device="cuda")
    tot_s = (X64 * a_s).sum(1)
    drift = float(((Yd * a_d).sum(1) - tot_s).abs().div((X64 * a_s).abs().sum(1)).max())
    drift64 = float(((Y64 * a_d).sum(1) - tot_s).abs().div((X64 * a_s).abs().sum(1)).max())
    log(f"[bench] FP32 vs FP64: max rel field error {rel_err:.2e}; conservation drift FP32 {drift:.2e}, "
        f"FP64 {drift64:.2e}")

    # ---- CPU baseline (oracle) -------------------------------------------------
    # The FP64 restatement in oracle/ on this host's cores: the full build at 1
    # thread and at every core of this process's share, single-field applies at
    # 1 thread (as mul!) and at every core; the oracle also counts the FP64
    # operations of its clips, the numerator of the clip kernels' FP64 rooflines.
    cpu = None
    build_speedup = None
    fp64_kernels = None
    hinfo = host_info()
    if not args.skip_cpu:
        orc = ge.load_oracle()
        nth = cpu_threads()
        scells = cr.points_to_cells(src_pts)
        tc = time.perf_counter()
        ref = orc.build(dst_cells_host, scells, clip_is_dst=False, nthreads=nth)
        cpu_build_s = time.perf_counter() - tc
        tc = time.perf_counter()
        ref1 = orc.build(dst_cells_host, scells, clip_is_dst=False, nthreads=1)
        cpu_build1_s = time.perf_counter() - tc
        assert ref1.nnz == ref.nnz
        del ref1
        cpu_pairs, cpu_pairs1 = ref.nnz / cpu_build_s, ref.nnz / cpu_build1_s
        build_speedup = (nnz / build_host_s) / cpu_pairs
        refc = orc.build(dst_cells_host, scells, clip_is_dst=False, nthreads=nth, count_flops=True)
        flops_per_cand = refc.clip_flops / max(1, refc.n_candidates)
        flops_per_kept = refc.clip_flops_kept / max(1, refc.nnz)
        del refc
        # FP64 single-field applies: 16 of the 256 fields, 1 thread and all threads
        rp, col, val = R.csr_arrays()
        xs = X64[:16].cpu().numpy()
        tc = time.perf_counter()
        for f in range(16):
            orc.spmv(rp, col, val, xs[f], R.dst_areas, 1)
        cpu_apply1_s = (time.perf_counter() - tc) / 16
        tc = time.perf_counter()
        for f in range(16):
            orc.spmv(rp, col, val, xs[f], R.dst_areas, nth)
        cpu_apply_s = (time.perf_counter() - tc) / 16
        cpu_gbps, cpu_gbps1 = B64 / cpu_apply_s / 1e9, B64 / cpu_apply1_s / 1e9
        # parity of the device FP64 apply against the oracle apply of the same CSR
        y0 = orc.spmv(rp, col, val, xs[0], R.dst_areas, nth)
        fp64_err = float(np.max(np.abs(Y64[0].cpu().numpy() - y0)) / np.max(np.abs(y0)))
        cpu = {"value": round(cpu_gbps, 3), "unit": "GB/s", "cores": nth, "kind": "port",
               "sample": f"16 of the 256 FP64 single-field applies on oracle/ at {nth} threads",
               "apply_1thread_GBps": round(cpu_gbps1, 3),
               "build_pairs_per_s": round(cpu_pairs, 1), "build_threads": nth,
               "build_1thread_pairs_per_s": round(cpu_pairs1, 1),
               "build_sample": "full build of this workload on oracle/ (from host cell tables)",
               "build_seconds": round(cpu_build_s, 3), "build_1thread_seconds": round(cpu_build1_s, 3),
               "cpu_nnz": ref.nnz, "fp64_field_rel_err_vs_oracle": fp64_err, **hinfo}
        del ref
        log(f"[bench] CPU oracle ({hinfo['cpu_model']}, nproc {hinfo['nproc']}, {nth} threads used): build "
            f"{cpu_build_s:.2f} s ({cpu_pairs / 1e6:.2f} M pairs/s), 1 thread {cpu_build1_s:.2f} s; apply "
            f"{cpu_apply_s * 1e3:.2f} ms/field at {nth} threads ({cpu_gbps:.2f} GB/s), "
            f"{cpu_apply1_s * 1e3:.1f} ms at 1 ({cpu_gbps1:.2f} GB/s); GPU build (from host arrays) "
            f"{build_speedup:.1f}x the {nth}-thread oracle; FP64 GPU vs oracle {fp64_err:.1e}")
        # FP64 rooflines of the two FP64-bound kernels.  k_clip_compact: the
        # oracle's operations per candidate x the device's candidates over the
        # clip stage's time.  The vertex-gradient pass (k_clip_grad_edges, the
        # declined pairs' k_clip_grad_list and the area term) differentiates
        # the clip of every kept pair: 2x the clip's operations is counted, a
        # lower bound, over the whole pass, so the fraction is a lower bound too.
        peak64 = 78.6e12
        clip_flops = flops_per_cand * ncand
        grad_flops = 2.0 * flops_per_kept * nnz
        fp64_kernels = {
            "peak_TFLOPs": 78.6,
            "flops_source": "FP64 adds/subs/muls/divs/sqrts of the oracle's clip and area (atan2 as one), "
                            "per candidate pair of its own candidate search",
            "k_clip_compact": {"flops": clip_flops, "ms": round(st["ms_clip"], 4),
                               "achieved_TFLOPs": round(clip_flops / (st["ms_clip"] * 1e-3) / 1e12, 3),
                               "frac": round(clip_flops / (st["ms_clip"] * 1e-3) / peak64, 4),
                               "flops_per_candidate": round(flops_per_cand, 1)},
            "vertex_pass": {"kernels": "k_clip_grad_edges + k_clip_grad_list + area term",
                            "flops_lower_bound": grad_flops, "ms_vertex_pass": round(ms_vert, 4),
                            "achieved_TFLOPs_lower_bound": round(grad_flops / (ms_vert * 1e-3) / 1e12, 3),
                            "frac_lower_bound": round(grad_flops / (ms_vert * 1e-3) / peak64, 4),
                            "clip_flops_per_kept_pair": round(flops_per_kept, 1)}}
        log(f"[bench] FP64 rooflines: k_clip_compact {fp64_kernels['k_clip_compact']['achieved_TFLOPs']:.2f} "
            f"TFLOP/s ({fp64_kernels['k_clip_compact']['frac']:.1%} of 78.6), vertex pass >= "
            f"{fp64_kernels['vertex_pass']['achieved_TFLOPs_lower_bound']:.2f} TFLOP/s")

    out = {
        "metric": METRIC,
        "value": round(gbps, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # variants 5-7, 18-19 sum in FP32 (Kahan); 20-27 use A rounded to FP32 (but 25)
        "dtype": ("f32 fields, f32 A, f32 Kahan sums" if variant in (5, 6, 7, 18, 19) else
                  "f32 fields, f32-rounded A, f64 sums" if variant in (20, 21, 22, 23, 24, 26, 27) else
                  "f32 fields, f64 A, f64 sums"),
        "data": "synthetic",
        "config": {
            "workload": "0.25deg lon-lat 1440x720 -> HEALPix nside 256 nested (16-gon cells); "
                        "256 FP32 fields [n,64,4] cells fastest, dims=1",
            "fields": K, "n_src": n_src, "n_dst": n_dst, "nnz": nnz, "apply_variant": variant,
            "variant_ms": variant_ms,
        },
        # pairs/s from the host inputs (source point matrix uploaded and
        # checked, HEALPix cells generated from nside on the device) to device
        # CSR(A), CSR(A^T) and areas; the build from host vertex tables of both
        # grids and from device-resident grids beside it
        "build": {"pairs_per_s": round(nnz / build_host_s, 1), "ms": round(build_host_s * 1e3, 3), "nnz": nnz,
                  "from_host_vertex_tables": {"ms": round(build_tables_s * 1e3, 3),
                                              "pairs_per_s": round(nnz / build_tables_s, 1)},
                  "from_device_grids": {"ms": round(build_s * 1e3, 3), "pairs_per_s": round(nnz / build_s, 1)},
                  "candidates": ncand, "candidate_pairs_per_s": round(ncand / build_s, 1),
                  "stage_ms": {k: round(st[k], 3) for k in
                               ("ms_bounds", "ms_tree", "ms_search", "ms_clip", "ms_assemble", "ms_total",
                                "ms_host")},
                  "speedup_vs_cpu": None if build_speedup is None else round(build_speedup, 2)},
        "fp64_loop": {"ms": round(ms_loop, 4), "GBps": round(gbps_loop, 2),
                      "batched_speedup": round(ms_loop / ms_step, 3), "kernel": best64,
                      "ms_by_kernel": {k: round(v, 4) for k, v in loop_ms.items()},
                      "cold_ms_per_field": round(ms_cold_field, 5), "cold_loop_ms": round(K * ms_cold_field, 4),
                      "cold_batched_speedup": round(K * ms_cold_field / ms_step, 3)},
        # hand-written streaming kernels (libcrhbm.so): the achievable HBM rates
        "hbm_ceilings": {**ceil, "apply_mix_GBps": round(mix_gbps, 1)},
        "copy_GBps": round(copy_gbps, 1),
        # the batched apply's algorithmic GB/s over the achievable rate of its
        # read/write mix; roofline.frac below is over the published 8 TB/s
        "frac_of_measured_copy": round(gbps / copy_gbps, 4),
        "frac_of_measured_mix": round(gbps / mix_gbps, 4),
        "backward": {"ms": round(ms_bwd, 4), "pairs_per_s": round(nnz / (ms_bwd * 1e-3), 1),
                     "vertex_gradients_ms": round(ms_vert, 4)},
        "graph_ms_per_step": round(ms_graph, 5),
        "graph_vs_eager": {"graph_ms": round(ms_graph, 5), "eager_ms": round(ms_eager_il, 5),
                           "launches_per_graph": per_graph, "rounds": 5},
        "first_apply_ms": round(first_apply_ms, 2),
        "first_apply_next_regridder_ms": round(first_apply2_ms, 2),
        "accuracy": {"fp32_max_rel_field_error": rel_err, "fp32_conservation_drift": drift,
                     "fp64_conservation_drift": drift64},
        "roofline": {"bound": "hbm", "achieved": round(gbps, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps / HBM_PEAK_GBPS, 4),
                     "traffic": read_traffic("spmm_f32_cells_c3_pmc.txt", variant)},
        "fp64_kernels": fp64_kernels,
        "cpu_baseline": cpu,
        "host": hinfo,
        "device": dev,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
