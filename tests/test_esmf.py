"""ESMF offline-weights files (ext/ConservativeRegriddingNCDatasetsExt.jl:15-59;
test/extensions/ncdatasets.jl:32-122 semantics).  The library writes netCDF
classic itself.  CDF-2 files are read back with scipy's independent netCDF
reader (scipy.io.netcdf_file, which predates CDF-5); the library's reader is
checked on files scipy writes and on a CDF-5 file assembled here byte by byte
from the format specification; CDF-5 files (64-bit row/col, the default)
round-trip through the library's reader.  The CPU tests write from host
arrays; the GPU tests from device regridders (the README example and a global
build), and rebuild a device regridder from a file."""
import struct
import json
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _read(path):
    f = scipy.io.netcdf_file(path, "r", mmap=False)
    out = {"dims": dict(f.dimensions), "attrs": {k: v for k, v in f._attributes.items()},
           "vars": {k: np.array(v[:]) for k, v in f.variables.items()},
           "var_attrs": {k: dict(v._attributes) for k, v in f.variables.items()},
           "version": f.version_byte}
    f.close()
    return out


def _readme():
    d = json.load(open(os.path.join(GOLDEN, "readme_example.json")))
    return sp.csc_matrix(np.array(d["A_dense"])), np.array(d["dst_areas"]), np.array(d["src_areas"])


def _esmf_arrays(A, a_dst, a_src):
    A = sp.csc_matrix(A)
    A.sort_indices()
    col = np.repeat(np.arange(A.shape[1]), np.diff(A.indptr))
    row = A.indices
    S = A.data / a_dst[row]
    frac_a = np.asarray(A.sum(0)).ravel() / a_src
    frac_b = np.asarray(A.sum(1)).ravel() / a_dst
    return S, row + 1, col + 1, frac_a, frac_b


def test_schema_and_values_from_arrays(cr, tmp_path):
    A, a_dst, a_src = _readme()
    S, row, col, fa, fb = _esmf_arrays(A, a_dst, a_src)
    path = str(tmp_path / "w.nc")
    assert cr.write_esmf_arrays(path, S, row, col, fa, a_src, fb, a_dst, cdf_version=2) == path
    f = _read(path)
    assert f["version"] == 2
    assert f["dims"] == {"n_s": A.nnz, "n_a": 5, "n_b": 4}
    assert set(f["vars"]) == {"S", "row", "col", "frac_a", "frac_b", "area_a", "area_b"}
    np.testing.assert_array_equal(f["vars"]["S"], S)
    np.testing.assert_array_equal(f["vars"]["row"], row)
    np.testing.assert_array_equal(f["vars"]["col"], col)
    np.testing.assert_array_equal(f["vars"]["area_a"], a_src)
    np.testing.assert_array_equal(f["vars"]["area_b"], a_dst)
    assert f["attrs"]["normalization"] == b"destarea"
    assert f["attrs"]["source_grid"] == b"source" and f["attrs"]["destination_grid"] == b"destination"
    assert f["attrs"]["map_method"] == b"Conservative remapping"
    for k in ("created_at", "source_grid_shape", "destination_grid_shape"):
        assert k not in f["attrs"]
    assert f["var_attrs"]["S"]["long_name"] == b"weight value"
    # rebuild the matrix from (S, row, col)
    R = sp.csr_matrix((f["vars"]["S"] * a_dst[f["vars"]["row"] - 1], (f["vars"]["row"] - 1, f["vars"]["col"] - 1)),
                      shape=A.shape)
    np.testing.assert_allclose(R.toarray(), A.toarray(), rtol=1e-15)


def test_optional_attributes(cr, tmp_path):
    A, a_dst, a_src = _readme()
    S, row, col, fa, fb = _esmf_arrays(A, a_dst, a_src)
    path = str(tmp_path / "sub" / "w.nc")
    os.makedirs(os.path.dirname(path))
    cr.write_esmf_arrays(path, S, row, col, fa, a_src, fb, a_dst, src_grid_name="era5_0.25deg",
                         dst_grid_name="c90", src_shape=(720, 361), dst_shape=(90, 90, 6),
                         created_at="2026-04-21T00:00:00", cdf_version=2)
    f = _read(path)
    assert f["attrs"]["source_grid"] == b"era5_0.25deg"
    assert f["attrs"]["destination_grid"] == b"c90"
    assert f["attrs"]["created_at"] == b"2026-04-21T00:00:00"
    assert list(f["attrs"]["source_grid_shape"]) == [720, 361]
    assert list(f["attrs"]["destination_grid_shape"]) == [90, 90, 6]


def test_odd_lengths_are_padded(cr, tmp_path):
    """32-bit variables of odd length and odd attribute strings exercise the
    4-byte padding of the classic format."""
    rng = np.random.default_rng(0)
    n_s, n_a, n_b = 7, 3, 5
    S = rng.random(n_s)
    row = rng.integers(1, n_b + 1, n_s).astype(np.int32)
    col = np.sort(rng.integers(1, n_a + 1, n_s)).astype(np.int32)
    path = str(tmp_path / "o.nc")
    cr.write_esmf_arrays(path, S, row, col, rng.random(n_a), rng.random(n_a), rng.random(n_b), rng.random(n_b),
                         src_grid_name="abcde", created_at="x", cdf_version=2)
    f = _read(path)
    np.testing.assert_array_equal(f["vars"]["row"], row)
    np.testing.assert_array_equal(f["vars"]["col"], col)
    np.testing.assert_array_equal(f["vars"]["S"], S)
    assert f["attrs"]["source_grid"] == b"abcde" and f["attrs"]["created_at"] == b"x"


def test_cdf5_is_the_default_and_round_trips(cr, tmp_path):
    """CDF-5 keeps row/col as 64-bit integers (the reference's Int64): values
    beyond 2^31 survive, and every variable reads back bit for bit."""
    rng = np.random.default_rng(1)
    n_s, n_a, n_b = 9, 4, 6
    S = rng.random(n_s)
    row = rng.integers(1, n_b + 1, n_s)
    col = np.sort(rng.integers(1, n_a + 1, n_s))
    row[3] = 3_000_000_007  # no 32-bit integer holds it
    arrs = (rng.random(n_a), rng.random(n_a), rng.random(n_b), rng.random(n_b))
    path = str(tmp_path / "w5.nc")
    cr.write_esmf_arrays(path, S, row, col, *arrs)
    assert open(path, "rb").read(4) == b"CDF\x05"
    got = cr.read_esmf_weights(path)
    assert got["row"].dtype == np.int64
    np.testing.assert_array_equal(got["row"], row)
    np.testing.assert_array_equal(got["col"], col)
    np.testing.assert_array_equal(got["S"], S)
    for k, v in zip(("frac_a", "area_a", "frac_b", "area_b"), arrs):
        np.testing.assert_array_equal(got[k], v)
    # CDF-2 cannot hold that index
    with pytest.raises(cr.MethodError):
        cr.write_esmf_arrays(str(tmp_path / "w2.nc"), S, row, col, *arrs, cdf_version=2)


def test_reader_takes_files_scipy_writes(cr, tmp_path):
    """An independent writer (scipy.io.netcdf_file, CDF-1 and CDF-2, NC_INT and
    NC_FLOAT variables) read by the library."""
    A, a_dst, a_src = _readme()
    S, row, col, fa, fb = _esmf_arrays(A, a_dst, a_src)
    for version in (1, 2):
        path = str(tmp_path / f"scipy{version}.nc")
        f = scipy.io.netcdf_file(path, "w", version=version)
        f.createDimension("n_s", S.size)
        f.createDimension("n_a", 5)
        f.createDimension("n_b", 4)
        f.title = "weights"
        for name, data, dim, typ in (("S", S, "n_s", "d"), ("row", row, "n_s", "i"), ("col", col, "n_s", "i"),
                                     ("frac_a", fa, "n_a", "f"), ("area_a", a_src, "n_a", "d"),
                                     ("frac_b", fb, "n_b", "d"), ("area_b", a_dst, "n_b", "d")):
            v = f.createVariable(name, typ, (dim,))
            v[:] = data
            v.units = "1"
        f.close()
        got = cr.read_esmf_weights(path)
        np.testing.assert_array_equal(got["S"], S)
        np.testing.assert_array_equal(got["row"], row)
        np.testing.assert_array_equal(got["col"], col)
        np.testing.assert_array_equal(got["frac_a"], fa.astype(np.float32))
        np.testing.assert_array_equal(got["area_b"], a_dst)


def _cdf5_bytes(n_s, S, row, col, a):
    """A CDF-5 ESMF file assembled from "The NetCDF Classic Format
    Specification" (64-bit counts, NC_INT64 = 10, NC_DOUBLE = 6), with one
    global and one variable attribute and the data of each variable at its
    recorded 64-bit offset."""
    be64, be32 = (lambda v: struct.pack(">q", v)), (lambda v: struct.pack(">i", v))

    def name(s):
        b = s.encode()
        return be64(len(b)) + b + b"\0" * ((4 - len(b) % 4) % 4)

    dims = [("n_s", n_s), ("n_a", len(a)), ("n_b", len(a))]
    variables = [("S", 0, 6, S.astype(">f8").tobytes()), ("row", 0, 10, row.astype(">i8").tobytes()),
                 ("col", 0, 10, col.astype(">i8").tobytes()), ("frac_a", 1, 6, a.astype(">f8").tobytes()),
                 ("area_a", 1, 6, a.astype(">f8").tobytes()), ("frac_b", 2, 6, a.astype(">f8").tobytes()),
                 ("area_b", 2, 6, a.astype(">f8").tobytes())]

    def header(offsets):
        h = b"CDF\x05" + be64(0)
        h += be32(0x0A) + be64(len(dims)) + b"".join(name(n) + be64(v) for n, v in dims)
        h += be32(0x0C) + be64(1) + name("title") + be32(2) + be64(2) + b"ab\0\0"
        h += be32(0x0B) + be64(len(variables))
        for (vn, dim, typ, data), off in zip(variables, offsets):
            h += name(vn) + be64(1) + be64(dim)
            h += be32(0x0C) + be64(1) + name("units") + be32(2) + be64(1) + b"m\0\0\0"
            h += be32(typ) + be64(len(data)) + be64(off)
        return h

    size = len(header([0] * len(variables)))
    offsets, pos = [], size
    for v in variables:
        offsets.append(pos)
        pos += len(v[3])
    return header(offsets) + b"".join(v[3] for v in variables)


def test_reader_takes_a_hand_assembled_cdf5_file(cr, tmp_path):
    S = np.array([0.5, 0.25, 1.0])
    row = np.array([1, 2, 2], dtype=np.int64)
    col = np.array([1, 1, 2], dtype=np.int64)
    a = np.array([2.0, 4.0])
    path = tmp_path / "hand5.nc"
    path.write_bytes(_cdf5_bytes(3, S, row, col, a))
    got = cr.read_esmf_weights(str(path))
    np.testing.assert_array_equal(got["S"], S)
    np.testing.assert_array_equal(got["row"], row)
    np.testing.assert_array_equal(got["col"], col)
    np.testing.assert_array_equal(got["area_b"], a)


def test_reader_rejects_what_it_cannot_read(cr, tmp_path):
    bad = tmp_path / "not.nc"
    bad.write_bytes(b"HDF5 is not netCDF classic")
    with pytest.raises(cr.ArgumentError):
        cr.read_esmf_weights(str(bad))
    hdf = tmp_path / "nc4.nc"
    hdf.write_bytes(b"CDF\x03" + b"\0" * 32)
    with pytest.raises(cr.MethodError):
        cr.read_esmf_weights(str(hdf))
    path = str(tmp_path / "w.nc")
    f = scipy.io.netcdf_file(path, "w", version=2)
    f.createDimension("n_s", 1)
    f.createDimension("n_a", 1)
    f.createDimension("n_b", 1)
    v = f.createVariable("S", "d", ("n_s",))
    v[:] = [1.0]
    f.close()
    with pytest.raises(cr.ArgumentError, match="no variable row"):
        cr.read_esmf_weights(path)
    with pytest.raises(cr.ArgumentError):
        cr.read_esmf_weights(str(tmp_path / "missing.nc"))


@pytest.mark.gpu
def test_regridder_from_esmf_file(cr, tmp_path):
    """A device regridder rebuilt from a CDF-5 file applies like the original
    (A = S .* area_b to rounding)."""
    R = cr.Regridder(cr.Grid.from_points(cr.lonlat_points(72, 36)), cr.Grid.from_points(cr.lonlat_points(90, 45)))
    path = str(tmp_path / "g5.nc")
    cr.save_esmf_weights(path, R)
    R2 = cr.regridder_from_esmf(path)
    assert R2.shape == R.shape and R2.nnz == R.nnz
    np.testing.assert_array_equal(R2.dst_areas, R.dst_areas)
    np.testing.assert_array_equal(R2.src_areas, R.src_areas)
    A, A2 = R.intersections, R2.intersections
    assert abs(A2 - A).max() <= 4e-16 * A.max()
    import torch
    x = torch.rand(90 * 45, dtype=torch.float64, device="cuda")
    y, y2 = cr.regrid(R, x), cr.regrid(R2, x)
    assert float((y - y2).abs().max()) <= 1e-15 * float(y.abs().max())


@pytest.mark.gpu
def test_regridder_file_round_trip(cr, tmp_path):
    """ncdatasets.jl "round-trip values": the README regridder (normalize=false)."""
    A, a_dst, a_src = _readme()
    R = cr.Regridder.from_csr(A, a_dst, a_src)
    path = str(tmp_path / "r.nc")
    assert cr.save_esmf_weights(path, R, cdf_version=2) == path
    f = _read(path)
    assert f["dims"] == {"n_s": A.nnz, "n_a": 5, "n_b": 4}
    S, row, col, fa, fb = _esmf_arrays(A, a_dst, a_src)
    np.testing.assert_array_equal(f["vars"]["row"], row)
    np.testing.assert_array_equal(f["vars"]["col"], col)
    np.testing.assert_allclose(f["vars"]["S"], S, rtol=1e-15)
    np.testing.assert_allclose(f["vars"]["frac_a"], fa, rtol=1e-15)
    np.testing.assert_allclose(f["vars"]["frac_b"], fb, rtol=1e-15)
    # transpose(R): roles of the grids swap
    path_t = str(tmp_path / "t.nc")
    cr.save_esmf_weights(path_t, cr.transpose(R), cdf_version=2)
    ft = _read(path_t)
    assert ft["dims"] == {"n_s": A.nnz, "n_a": 4, "n_b": 5}
    np.testing.assert_array_equal(ft["vars"]["area_a"], a_dst)


@pytest.mark.gpu
def test_global_build_fractions_are_one(cr, tmp_path):
    """Both grids cover the sphere: frac_a = frac_b = 1 (ncdatasets.jl:107-110)."""
    R = cr.Regridder(cr.Grid.from_points(cr.lonlat_points(72, 36)), cr.Grid.from_points(cr.lonlat_points(90, 45)),
                     radius=1.0)
    path = str(tmp_path / "g.nc")
    cr.save_esmf_weights(path, R, src_shape=(90, 45), dst_shape=(72, 36), cdf_version=2)
    f = _read(path)
    np.testing.assert_allclose(f["vars"]["frac_a"], 1.0, atol=1e-12)
    np.testing.assert_allclose(f["vars"]["frac_b"], 1.0, atol=1e-12)
    A = R.intersections
    rebuilt = sp.csr_matrix((f["vars"]["S"] * R.dst_areas[f["vars"]["row"] - 1],
                             (f["vars"]["row"] - 1, f["vars"]["col"] - 1)), shape=A.shape)
    assert abs(rebuilt - A).max() <= 1e-15 * A.max()
    # column-major order with ascending rows inside each column
    key = (f["vars"]["col"].astype(np.int64) << 32) | f["vars"]["row"]
    assert np.all(np.diff(key) > 0)
