"""Matrix Market I/O against MTJ's reading rules and scipy.io (an
independent implementation of the published format)."""
import io

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from ressim_amd import mmio


def _dense(n, m, rp, ci, d):
    return sp.csr_matrix((d, ci, rp), shape=(n, m)).toarray()


def test_round_trip_and_scipy_agree(tmp_path):
    rng = np.random.default_rng(0)
    A = sp.random(37, 29, density=0.2, random_state=1, format="csr")
    A.data = rng.standard_normal(A.nnz)
    A.sort_indices()
    txt = mmio.write_matrix(A.indptr, A.indices, A.data, 37, 29, digits=17)
    n, m, rp, ci, d = mmio.read_matrix(io.StringIO(txt))
    assert (n, m) == (37, 29)
    assert np.array_equal(rp, A.indptr) and np.array_equal(ci, A.indices)
    assert np.array_equal(d, A.data)
    p = tmp_path / "a.mtx"
    p.write_text(txt)
    B = scipy.io.mmread(str(p)).tocsr()
    assert np.array_equal(B.toarray(), A.toarray())
    # MTJ's 13 significant digits
    txt12 = mmio.write_matrix(A.indptr, A.indices, A.data, 37, 29)
    _, _, _, _, d12 = mmio.read_matrix(io.StringIO(txt12))
    assert np.allclose(d12, A.data, rtol=1e-12, atol=0) and not np.array_equal(d12, A.data)
    r0 = int(np.searchsorted(A.indptr, 0, side="right") - 1)
    assert txt12.splitlines()[1] == "%10d %10d %19d" % (37, 29, A.nnz)
    assert txt12.splitlines()[2] == "%10d %10d % .12e" % (r0 + 1, A.indices[0] + 1, A.data[0])


def test_scipy_written_symmetric_file_is_expanded(tmp_path):
    S = sp.random(20, 20, density=0.15, random_state=3)
    S = (S + S.T).tocoo()
    p = tmp_path / "s.mtx"
    scipy.io.mmwrite(str(p), S, symmetry="symmetric")
    n, m, rp, ci, d = mmio.read_matrix(str(p))
    assert np.allclose(_dense(n, m, rp, ci, d), S.toarray(), rtol=1e-15, atol=0)


def test_skew_symmetric_header_comments_and_duplicates():
    txt = ("%%MatrixMarket MATRIX coordinate Real Skew-Symmetric\n"
           "% a comment line\n"
           "3 3 4   % trailing comment\n"
           "2 1 5.0\n"
           "3 1 -2.5\n"
           "3 2 1.0\n"
           "3 2 7.0\n")
    n, m, rp, ci, d = mmio.read_matrix(io.StringIO(txt))
    D = _dense(n, m, rp, ci, d)
    E = np.array([[0, -5.0, 2.5], [5.0, 0, -7.0], [-2.5, 7.0, 0]])
    assert np.array_equal(D, E)   # the repeated (3,2) keeps its last value
    assert np.array_equal(rp, [0, 2, 4, 6])


def test_no_header_is_coordinate_real_general():
    n, m, rp, ci, d = mmio.read_matrix(io.StringIO("2 3 2\n1 3 4.0\n2 1 -1\n"))
    assert (n, m) == (2, 3) and np.array_equal(_dense(n, m, rp, ci, d), [[0, 0, 4], [-1, 0, 0]])


@pytest.mark.parametrize("head,msg", [
    ("%%MatrixMarket matrix coordinate pattern general", "Pattern"),
    ("%%MatrixMarket matrix array real general", "Dense"),
    ("%%MatrixMarket matrix coordinate complex general", "Complex")])
def test_unsupported_kinds(head, msg):
    with pytest.raises(mmio.UnsupportedOperationException, match=msg):
        mmio.read_matrix(io.StringIO(head + "\n2 2 1\n1 1 1.0\n"))


def test_bad_header_and_truncated_file():
    with pytest.raises(IOError):
        mmio.read_matrix(io.StringIO("%%MatrixMarket vector coordinate real general\n1 1 0\n"))
    with pytest.raises(EOFError):
        mmio.read_matrix(io.StringIO("3 3 2\n1 1 1.0\n"))


def test_vectors():
    x = np.array([1.5, -2.0, 0.0, 3.25])
    assert np.array_equal(mmio.read_vector(io.StringIO(mmio.write_vector(x))), x)
    v = mmio.read_vector(io.StringIO("%%MatrixMarket vector coordinate real\n5 2\n2 1.0\n5 -3\n"))
    assert np.array_equal(v, [0, 1.0, 0, 0, -3.0])
