"""Tile-plan helpers shared by the plan tests: synthetic CSR matrices, a host
emulation of the staged tile kernel on a plan, and a device-plan fetch.

The emulation follows the kernel's contract, independent of how a plan is
made: per tile, thread t walks its slots e = 0..ent-1 (slot ent_off + e*T + t),
adds value * x[cols[col_off + (key & 0xffff)]] to one accumulator and, where a
segment ends (kRounds: bit e of the wave's flush mask; kEven: the next slot's
segment differs, or e is the last slot), stores the sum to segment key >> 16
and restarts.  A row is the sum of its segments rseg[row0 + tile + lr] ..
rseg[row0 + tile + lr + 1] - 1; rows listed as huge are plain CSR rows."""
import ctypes

import numpy as np

ROUNDS, EVEN = 0, 1


def synthetic_csr(kind, seed=0):
    """(nrows, ncols, rowptr, col, val) of a synthetic matrix.

    banded   -- regridding-like: 6000 rows, 2..12 neighbouring columns each
    ragged   -- row lengths from 0 to 3000 (some rows huge for one or both
                geometries), random columns with repeats across rows, a few
                columns repeated inside a row, unsorted
    wide     -- 400 rows of 300 distinct columns each: tiles of a few rows,
                so tiles start off the 32-row grid
    empty    -- 100 rows, no non-zeros
    """
    rng = np.random.default_rng(seed)
    if kind == "banded":
        nrows, ncols = 6000, 4000
        n = rng.integers(2, 13, nrows)
        cols = []
        for r in range(nrows):
            c0 = int(r * ncols / nrows) - int(n[r]) // 2
            cols.append(np.clip(np.arange(c0, c0 + n[r]), 0, ncols - 1))
        cols = [np.unique(c) for c in cols]
    elif kind == "ragged":
        nrows, ncols = 3000, 5000
        u = rng.random(nrows)
        n = np.where(u < 0.80, rng.integers(0, 11, nrows),
                     np.where(u < 0.97, rng.integers(11, 400, nrows), rng.integers(1000, 3000, nrows)))
        cols = []
        for r in range(nrows):
            c = rng.integers(0, ncols, n[r])
            if n[r] > 3 and r % 7 == 0:
                c[-1] = c[0]  # a column twice in one row
            cols.append(c)
    elif kind == "wide":
        nrows, ncols = 400, 6000
        cols = [rng.choice(ncols, 300, replace=False) for _ in range(nrows)]
    elif kind == "empty":
        nrows, ncols = 100, 50
        cols = [np.zeros(0, np.int64) for _ in range(nrows)]
    else:
        raise ValueError(kind)
    lens = np.array([len(c) for c in cols], np.int64)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = (np.concatenate(cols) if rp[-1] else np.zeros(0)).astype(np.int32)
    val = rng.random(col.shape[0]) + 0.5
    return nrows, ncols, rp, col, val


def _pairs(c, v):
    """(column, value) pairs sorted lexicographically."""
    p = np.stack([np.asarray(c, np.float64), np.asarray(v, np.float64)], axis=1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def emulate(P, style, rp, col, val, x, threads):
    """y = A x through plan P (host arrays), FP64, with structural checks:
    tiles cover the rows in order, every segment below nseg is stored exactly
    once, and the live slots of a tile hold exactly the (column, value) pairs
    of its non-huge rows."""
    nrows = rp.shape[0] - 1
    y = np.zeros(nrows)
    T = threads
    n = np.diff(rp)
    is_huge = np.zeros(nrows, bool)
    is_huge[P["huge"].astype(np.int64)] = True
    prev_end = 0
    for t, d in enumerate(P["desc"]):
        row0, nr, ent, nseg = int(d["row0"]), int(d["nrows"]), int(d["ent"]), int(d["nseg"])
        assert row0 == prev_end, "tiles are contiguous"
        prev_end = row0 + nr
        sl = slice(int(d["ent_off"]), int(d["ent_off"]) + ent * T)
        keys = P["key"][sl].reshape(ent, T).astype(np.int64)
        vals = P["v64"][sl].reshape(ent, T)
        assert np.array_equal(P["v32"][sl].reshape(ent, T), vals.astype(np.float32))
        tc = P["cols"][int(d["col_off"]):int(d["col_off"]) + int(d["ncols"])].astype(np.int64)
        assert np.all(np.diff(tc) > 0), "a tile's columns are distinct and ascending"
        lc, sg = keys & 0xFFFF, keys >> 16
        assert np.all(sg <= nseg)
        live = vals != 0.0
        assert np.all(lc[live] < tc.shape[0])
        xv = x[tc[np.minimum(lc, tc.shape[0] - 1)]] if tc.shape[0] else np.zeros(lc.shape)
        prod = np.where(live, vals * xv, 0.0)
        seg = np.zeros(nseg + 1)
        writes = np.zeros(nseg + 1, np.int64)
        acc = np.zeros(T)
        wave = np.arange(T) // 64
        for e in range(ent):
            acc += prod[e]
            if style == ROUNDS:
                fl = ((d["flush"][wave] >> np.uint32(e)) & 1).astype(bool)
            else:
                fl = np.ones(T, bool) if e == ent - 1 else sg[e + 1] != sg[e]
            seg[sg[e, fl]] = acc[fl]
            np.add.at(writes, sg[e, fl], 1)
            acc[fl] = 0.0
        assert np.all(writes[:nseg] == 1), "every segment is stored once"
        rs = P["rseg"][row0 + t:row0 + t + nr + 1].astype(np.int64)
        assert rs[0] == 0 and rs[-1] == nseg and np.all(np.diff(rs) >= 0)
        np.add.at(y, row0 + np.repeat(np.arange(nr), np.diff(rs)), seg[:nseg])
        # the live slots are the tile's non-zeros
        rows = np.arange(row0, row0 + nr)
        rows = rows[~is_huge[rows] & (n[rows] > 0)]
        ks = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in rows]) if rows.size else np.zeros(0, np.int64)
        want = _pairs(col[ks], val[ks])
        got = _pairs(tc[lc[live]], vals[live])
        assert np.array_equal(got, want), "slots hold the tile's non-zeros once each"
    assert prev_end == nrows, "tiles cover every row"
    for r in np.flatnonzero(is_huge):
        assert n[r] > 0
        y[r] = np.dot(val[rp[r]:rp[r + 1]], x[col[rp[r]:rp[r + 1]]])
    return y


def device_plan(L, handle, transpose, style, geom, empty_plan, blocks=False):
    """The device plan of one direction of a regridder handle, as host arrays
    (blocks: the plan over 32-row blocks sorted by their smallest column, with
    its row map)."""
    sizes = np.zeros(6, np.int64)
    st = int(style) | (2 if blocks else 0)
    L.check(L.lib.cr_regrid_tile_plan(handle, int(transpose), st, int(geom), sizes.ctypes.data, None), "tile_plan")
    P = empty_plan(sizes[:5])
    P["rowmap"] = np.zeros(int(sizes[5]), np.int32)
    L.check(L.lib.cr_regrid_copy_tile_plan(handle, int(transpose), st, int(geom),
                                           *(P[k].ctypes.data for k in ("desc", "key", "v32", "v64", "cols",
                                                                        "rseg", "huge", "rowmap"))),
            "copy_tile_plan")
    return P


def block_order(rp, col):
    """The 32-row block order of a block-ordered plan, restated: blocks sorted
    by their smallest column (stable; a block with no non-zeros sorts last),
    plan row v = matrix row 32 * order[v // 32] + v % 32, or -1 past the last
    row.  Returns (rowmap, the block-ordered CSR (rp, col, val index))."""
    nrows = rp.shape[0] - 1
    nb = (nrows + 31) // 32
    key = np.full(nb, 0xFFFFFFFF, np.uint64)
    for b in range(nb):
        k0, k1 = rp[32 * b], rp[min(nrows, 32 * b + 32)]
        if k1 > k0:
            key[b] = col[k0:k1].min()
    order = np.argsort(key, kind="stable")
    v = np.arange(nb * 32)
    rowmap = order[v // 32] * 32 + v % 32
    rowmap = np.where(rowmap < nrows, rowmap, -1).astype(np.int32)
    n = np.where(rowmap >= 0, np.diff(rp)[np.maximum(rowmap, 0)], 0)
    rp2 = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    idx = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in rowmap if r >= 0] or [np.zeros(0, np.int64)])
    return rowmap, rp2, idx.astype(np.int64)


def raw_regridder(L, nrows, ncols, rp, col, val):
    """A regridder handle made from raw CSR arrays (kept as given: unsorted
    columns, repeats inside a row), unit areas."""
    h = ctypes.c_void_p()
    a_dst, a_src = np.ones(nrows), np.ones(ncols)
    L.check(L.lib.cr_regridder_from_csr(ctypes.byref(h), nrows, ncols, col.shape[0], L.i64ptr(rp),
                                        L.i32ptr(col), L.dptr(val), L.dptr(a_dst), L.dptr(a_src), None),
            "from_csr")
    return h


def assert_same_plan(dev, ref):
    """Bit-for-bit equal plans (column lists compared where tiles use them)."""
    for k in ("desc", "key", "v32", "v64", "rseg", "huge"):
        assert dev[k].shape == ref[k].shape, f"{k}: {dev[k].shape} vs {ref[k].shape}"
        if k == "desc":
            for f in ref[k].dtype.names:
                bad = np.flatnonzero(np.any((dev[k][f] != ref[k][f]).reshape(ref[k].shape[0], -1), axis=1))
                assert bad.size == 0, f"desc.{f} differs at tiles {bad[:5]}: {dev[k][f][bad[:3]]} vs {ref[k][f][bad[:3]]}"
        else:
            a, b = dev[k].view(np.uint8), ref[k].view(np.uint8)
            bad = np.flatnonzero(a != b)
            assert bad.size == 0, f"{k} differs at byte {bad[:5]}"
    assert dev["cols"].shape == ref["cols"].shape
    for d in ref["desc"]:
        s = slice(int(d["col_off"]), int(d["col_off"]) + int(d["ncols"]))
        assert np.array_equal(dev["cols"][s], ref["cols"][s])
