"""MatrixMarket interchange of assembled systems (shesolve.interchange).

Pinned against scipy's independent MatrixMarket reader/writer: what we write,
scipy reads back entry for entry, and what scipy writes, we read; FP64 values
round-trip bit for bit; block ranges and the unknown mapping round-trip."""
import numpy as np
import scipy.io
import scipy.sparse as sp

from conftest import oracle_system


def test_system_round_trip_and_scipy_agreement(tmp_path):
    from shesolve import interchange as X
    s = oracle_system("diode_L3")
    A, b, _, _ = s.reduced()
    ranges = s.block_ranges()
    ei, eo, _, _ = s.mapping()
    prefix = str(tmp_path / "diode")
    X.write_system(prefix, A.row_ptr, A.col, A.val, b, block_ranges=ranges, mapping=(ei, eo))
    rp, ci, v, b2, blk, (ei2, eo2) = X.read_system(prefix, blocks=True, mapping=True)
    assert np.array_equal(rp, A.row_ptr) and np.array_equal(ci, A.col)
    assert np.array_equal(v, A.val) and np.array_equal(b2, b)
    assert np.array_equal(blk, np.asarray(ranges).reshape(-1, 2))
    assert np.array_equal(ei2, np.asarray(ei).ravel()) and np.array_equal(eo2, np.asarray(eo).ravel())
    # scipy reads our file
    M = scipy.io.mmread(prefix + ".mtx").tocsr()
    M.sort_indices()
    assert np.array_equal(M.indptr, A.row_ptr) and np.array_equal(M.indices, A.col)
    assert np.array_equal(M.data, A.val)
    assert np.array_equal(np.asarray(scipy.io.mmread(prefix + "_b.mtx")).ravel(), b)


def test_reads_scipy_written_files_and_rejects_bad_ones(tmp_path):
    import pytest
    from shesolve import interchange as X
    rng = np.random.default_rng(2)
    M = sp.random(40, 40, density=0.1, random_state=2, format="csr") + sp.eye(40)
    path = str(tmp_path / "m.mtx")
    scipy.io.mmwrite(path, M, precision=17)
    rp, ci, v = X.read_matrix(path)
    Ms = M.tocsr()
    Ms.sort_indices()
    assert np.array_equal(rp, Ms.indptr) and np.array_equal(ci, Ms.indices) and np.array_equal(v, Ms.data)
    x = rng.standard_normal(7)
    scipy.io.mmwrite(str(tmp_path / "x.mtx"), x.reshape(-1, 1), precision=17)
    assert np.array_equal(X.read_array(str(tmp_path / "x.mtx"))[:, 0], x)
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n")
    with pytest.raises(ValueError):
        X.read_matrix(str(bad))
