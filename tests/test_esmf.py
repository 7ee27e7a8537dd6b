"""ESMF offline-weights files (ext/ConservativeRegriddingNCDatasetsExt.jl:15-59;
test/extensions/ncdatasets.jl:32-122 semantics).  The library writes netCDF
classic itself; the files are read back with scipy's independent netCDF
reader (scipy.io.netcdf_file).  The CPU tests write from host arrays; the GPU
tests from device regridders (the README example and a global build)."""
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
    assert cr.write_esmf_arrays(path, S, row, col, fa, a_src, fb, a_dst) == path
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
                         created_at="2026-04-21T00:00:00")
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
                         src_grid_name="abcde", created_at="x")
    f = _read(path)
    np.testing.assert_array_equal(f["vars"]["row"], row)
    np.testing.assert_array_equal(f["vars"]["col"], col)
    np.testing.assert_array_equal(f["vars"]["S"], S)
    assert f["attrs"]["source_grid"] == b"abcde" and f["attrs"]["created_at"] == b"x"


@pytest.mark.gpu
def test_regridder_file_round_trip(cr, tmp_path):
    """ncdatasets.jl "round-trip values": the README regridder (normalize=false)."""
    A, a_dst, a_src = _readme()
    R = cr.Regridder.from_csr(A, a_dst, a_src)
    path = str(tmp_path / "r.nc")
    assert cr.save_esmf_weights(path, R) == path
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
    cr.save_esmf_weights(path_t, cr.transpose(R))
    ft = _read(path_t)
    assert ft["dims"] == {"n_s": A.nnz, "n_a": 4, "n_b": 5}
    np.testing.assert_array_equal(ft["vars"]["area_a"], a_dst)


@pytest.mark.gpu
def test_global_build_fractions_are_one(cr, tmp_path):
    """Both grids cover the sphere: frac_a = frac_b = 1 (ncdatasets.jl:107-110)."""
    R = cr.Regridder(cr.Grid.from_points(cr.lonlat_points(72, 36)), cr.Grid.from_points(cr.lonlat_points(90, 45)),
                     radius=1.0)
    path = str(tmp_path / "g.nc")
    cr.save_esmf_weights(path, R, src_shape=(90, 45), dst_shape=(72, 36))
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
