"""Matrix Market reader / writer (gDMRC.h) against scipy.io's independent implementation.

The reader must give the matrix scipy.io.mmread gives, as CSR with sorted columns and summed
duplicates, for every supported format / field / symmetry; the writer must round-trip bit for
bit (17 significant digits)."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import gdmr


def _dense(A):
    return A.to_scipy().toarray()


def _check_csr(A):
    assert A.rowptr[0] == 0 and A.rowptr[-1] == len(A.col) == len(A.val)
    for r in range(A.n):
        cols = A.col[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(cols) > 0), "columns sorted and unique in each row"


@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric"])
@pytest.mark.parametrize("field", ["real", "integer", "pattern"])
def test_coordinate_matches_scipy(tmp_path, symmetry, field):
    rng = np.random.default_rng(5)
    n = 37
    M = sp.random(n, n, density=0.15, random_state=6, format="coo").toarray()
    if field == "integer":
        M = np.round(M * 20)
    if field == "pattern":
        M = (M != 0).astype(float)
    if symmetry == "symmetric":
        M = np.tril(M) + np.tril(M, -1).T
    elif symmetry == "skew-symmetric":
        if field == "pattern":
            pytest.skip("the format has no skew-symmetric pattern matrices")
        M = np.tril(M, -1) - np.tril(M, -1).T
    if symmetry != "skew-symmetric":
        d = rng.random(n)
        M[np.diag_indices(n)] = 1.0 if field == "pattern" else np.round(20 * d) + 1
    path = str(tmp_path / "a.mtx")
    scipy.io.mmwrite(path, sp.coo_matrix(M), field=field, symmetry=symmetry)
    ref = scipy.io.mmread(path).toarray()
    A = gdmr.read_matrix_market(path)
    _check_csr(A)
    assert A.n == n
    np.testing.assert_array_equal(_dense(A), ref)


def test_array_format_and_duplicates(tmp_path):
    M = np.array([[1.5, 0.0, -2.0], [0.0, 0.0, 3.25], [4.0, 5.0, 0.0]])
    p1 = str(tmp_path / "d.mtx")
    scipy.io.mmwrite(p1, M)
    A = gdmr.read_matrix_market(p1)
    _check_csr(A)
    np.testing.assert_array_equal(_dense(A), M)
    assert A.nnz == 5  # explicit zeros of a dense file are dropped
    # Duplicate coordinates are summed (file order), as scipy's coo -> csr conversion does.
    p2 = str(tmp_path / "dup.mtx")
    with open(p2, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% comment\n\n3 3 5\n"
                "1 1 1.0\n3 2 2.0\n1 1 0.5\n2 3 -1.0\n3 2 0.25\n")
    B = gdmr.read_matrix_market(p2)
    _check_csr(B)
    np.testing.assert_array_equal(_dense(B), scipy.io.mmread(p2).tocsr().toarray())
    assert B.nnz == 3


def test_round_trip_and_rejections(tmp_path):
    A = gdmr.conv_diff_3d(6, jitter=0.01, seed=3, stream=1)
    p = str(tmp_path / "cd.mtx")
    gdmr.write_matrix_market(p, A)
    B = gdmr.read_matrix_market(p)
    np.testing.assert_array_equal(A.rowptr, B.rowptr)
    np.testing.assert_array_equal(A.col, B.col)
    np.testing.assert_array_equal(A.val, B.val)
    np.testing.assert_array_equal(scipy.io.mmread(p).toarray(), _dense(A))
    bad = tmp_path / "c.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1.0 0.0\n")
    with pytest.raises(ValueError):
        gdmr.read_matrix_market(str(bad))
    oob = tmp_path / "oob.mtx"
    oob.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(ValueError):
        gdmr.read_matrix_market(str(oob))
    with pytest.raises(ValueError):
        gdmr.read_matrix_market(str(tmp_path / "missing.mtx"))
    assert gdmr.lib.dmrReadMatrixMarket(str(bad).encode(), None, None, None, None, None,
                                        None) == gdmr.DMR_ERR_NOTBUILT
