"""Helpers shared by the GPU parity tests (imported by tests/test_gpu_*.py)."""
import numpy as np


def pattern_and_values(R, ref, tol=1e-12, thr=1e-14):
    """Compare a device Regridder (built with radius=1) with the oracle CSR.

    Areas must agree to ``tol`` relative to the smaller of the two cells, and a
    pair missing on one side counts as area 0 there: so the sparsity patterns
    must be equal except for pairs whose area is at most (thr + tol) * min cell
    area, i.e. pairs the two sides may round to either side of the drop
    threshold (slivers between nearly parallel edges do).
    Returns (max relative error, number of threshold flips)."""
    rp, col, val = R.csr_arrays()
    n_dst = rp.shape[0] - 1
    rows = np.repeat(np.arange(n_dst), np.diff(rp))
    rrows = ref.row_index()
    key = rows.astype(np.int64) * ref.n_src + col
    rkey = rrows.astype(np.int64) * ref.n_src + ref.col
    common, ia, ib = np.intersect1d(key, rkey, assume_unique=True, return_indices=True)
    only_dev = np.setdiff1d(np.arange(key.shape[0]), ia)
    only_ref = np.setdiff1d(np.arange(rkey.shape[0]), ib)
    a_dst, a_src = ref.a_dst, ref.a_src
    flips = 0
    for idx, r, c, v in ((only_dev, rows, col, val), (only_ref, rrows, ref.col, ref.val)):
        if idx.size:
            scale = np.minimum(a_dst[r[idx]], a_src[c[idx]])
            worst = float(np.max(v[idx] / scale))
            assert worst <= thr + tol, f"sparsity pattern differs beyond the threshold ({worst:.3e})"
            flips += idx.size
    scale = np.minimum(a_dst[rows[ia]], a_src[col[ia]])
    err = np.max(np.abs(val[ia] - ref.val[ib]) / scale) if ia.size else 0.0
    assert err <= tol, f"intersection areas differ: {err:.3e}"
    # each pair's error relative to its own area: over every common pair, and
    # over the pairs of at least 1e-4 of the smaller cell (slivers below that
    # carry their cells' absolute rounding)
    if ia.size:
        rel = np.abs(val[ia] - ref.val[ib]) / ref.val[ib]
        big = ref.val[ib] >= 1e-4 * scale
        print(f"[areas] {ia.size} pairs: max error {err:.2e} of the smaller cell; per pair {rel.max():.2e} "
              f"(all), {rel[big].max() if big.any() else 0.0:.2e} (pairs >= 1e-4 of the smaller cell)", flush=True)
        assert not big.any() or rel[big].max() <= 1e-8
    np.testing.assert_allclose(R.dst_areas, a_dst, rtol=1e-12)
    np.testing.assert_allclose(R.src_areas, a_src, rtol=1e-12)
    return err, flips


def jittered_cubed_sphere(cr, ne, seed, frac=0.05):
    """Cubed sphere with every face-interior vertex moved in its tangent plane
    by N(0, frac * cell width), renormalised (cells stay convex)."""
    P = cr.cubed_sphere_points(ne).copy()
    rng = np.random.default_rng(seed)
    width = np.pi / 2 / ne
    inner = P[:, 1:-1, 1:-1, :]
    noise = rng.normal(0.0, frac * width, size=inner.shape)
    noise -= np.sum(noise * inner, axis=-1, keepdims=True) * inner  # tangent plane
    moved = inner + noise
    P[:, 1:-1, 1:-1, :] = moved / np.linalg.norm(moved, axis=-1, keepdims=True)
    return P
