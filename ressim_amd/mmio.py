"""Matrix Market exchange format, as MTJ reads and writes it.

Reading follows ``CompRowMatrix(MatrixVectorReader)``
(mtj/src/no/uib/cipr/matrix/sparse/CompRowMatrix.java:68-148) and
``DenseVector(MatrixVectorReader)`` (mtj/src/no/uib/cipr/matrix/
DenseVector.java:52-95) over ``MatrixVectorReader``
(mtj/src/no/uib/cipr/matrix/io/MatrixVectorReader.java): the
``%%MatrixMarket`` header is optional (without it a coordinate real general
matrix / coordinate real vector is assumed), header words are compared
case-insensitively, ``%`` starts a comment anywhere, indices are 1-based;
pattern, dense (array) and complex matrices are rejected as MTJ rejects
them; symmetric and skew-symmetric files are expanded (the mirrored entries
are set after every stored one, negated for skew-symmetric); each row's
columns are sorted (``construct``, :188) and entries are stored with
``set`` in file order, so a repeated coordinate keeps its last value.

Writing follows ``MatrixVectorWriter`` / ``AbstractMatrix.toString``
(mtj/src/no/uib/cipr/matrix/io/MatrixVectorWriter.java:87-112,
AbstractMatrix.java:796-809): the header, then ``%10d %10d %19d`` (rows,
columns, nonzero entries) and one ``%10d %10d % .12e`` line per nonzero in
row-major order.  ``digits`` = 12 is MTJ's output; 17 makes a round trip
exact.

The files are host data; the matrix goes to the device through
``CompRowMatrix.from_csr`` (``read_comprow`` does both).
"""
from __future__ import annotations

import io
from typing import Optional

import numpy as np


class UnsupportedOperationException(NotImplementedError):
    """java.lang.UnsupportedOperationException."""


def _tokens(text: str):
    """Whitespace-separated tokens, with '%' starting a comment that runs to
    the end of the line (the StreamTokenizer setup of MatrixVectorReader)."""
    for line in text.splitlines():
        k = line.find("%")
        if k >= 0:
            line = line[:k]
        yield from line.split()


def _header(text: str):
    """(has_header, words) from the first line (readMatrixInfo /
    readVectorInfo)."""
    first = text.lstrip("\n").split("\n", 1)[0].strip()
    words = first.split()
    if words and words[0].lower() == "%%matrixmarket":
        if len(words) < 4:
            raise IOError("Invalid Matrix Market header")
        return True, [w.lower() for w in words]
    return False, []


def _read_text(src) -> str:
    if hasattr(src, "read"):
        t = src.read()
        return t.decode() if isinstance(t, bytes) else t
    with open(src) as f:
        return f.read()


def read_matrix(src):
    """CompRowMatrix(MatrixVectorReader): returns (nrows, ncols, rowptr,
    colind, data) with sorted rows."""
    text = _read_text(src)
    has, w = _header(text)
    if has:
        if len(w) < 5:
            raise IOError("Invalid Matrix Market header")
        if w[1] != "matrix":
            raise IOError('Expected "matrix", got ' + w[1])
        if w[2] not in ("coordinate", "array"):
            raise IOError("Unknown layout " + w[2])
        if w[3] not in ("real", "integer", "complex", "pattern"):
            raise IOError("Unknown field specification " + w[3])
        if w[4] not in ("general", "symmetric", "skew-symmetric", "hermitian"):
            raise IOError("Unknown symmetry specification " + w[4])
        sparse, field, sym = w[2] == "coordinate", w[3], w[4]
    else:
        sparse, field, sym = True, "real", "general"
    if field == "pattern":
        raise UnsupportedOperationException("Pattern matrices are not supported")
    if not sparse:
        raise UnsupportedOperationException("Dense matrices are not supported")
    if field == "complex":
        raise UnsupportedOperationException("Complex matrices are not supported")
    tok = _tokens(text.split("\n", 1)[1] if has else text)
    try:
        nrows, ncols, nent = int(next(tok)), int(next(tok)), int(next(tok))
        row = np.empty(nent, np.int64)
        col = np.empty(nent, np.int64)
        val = np.empty(nent)
        for i in range(nent):
            row[i] = int(next(tok)) - 1
            col[i] = int(next(tok)) - 1
            val[i] = float(next(tok))
    except StopIteration:
        raise EOFError("End-of-File encountered during parsing") from None
    if nent and (row.min() < 0 or row.max() >= nrows or col.min() < 0 or col.max() >= ncols):
        raise IndexError("entry outside the matrix")
    mirror = sym in ("symmetric", "skew-symmetric")
    off = row != col
    rr = np.concatenate([row, col[off]]) if mirror else row
    cc = np.concatenate([col, row[off]]) if mirror else col
    # the pattern: each row's distinct columns, sorted (construct, :160-190)
    key = np.unique(rr * max(ncols, 1) + cc)
    prow, pcol = key // max(ncols, 1), key % max(ncols, 1)
    rowptr = np.zeros(nrows + 1, np.int64)
    np.add.at(rowptr, prow + 1, 1)
    rowptr = np.cumsum(rowptr)
    data = np.zeros(len(key))

    def set_in_order(k, v):
        # set() one entry after another: a repeated coordinate keeps the
        # value of its last occurrence
        u, first_rev = np.unique(k[::-1], return_index=True)
        data[np.searchsorted(key, u)] = v[len(k) - 1 - first_rev]

    set_in_order(row * max(ncols, 1) + col, val)
    if mirror:
        # the mirrored entries are set after every stored one (:130-147)
        set_in_order(col[off] * max(ncols, 1) + row[off],
                     val[off] if sym == "symmetric" else -val[off])
    return nrows, ncols, rowptr.astype(np.int32), pcol.astype(np.int32), data


def read_vector(src) -> np.ndarray:
    """DenseVector(MatrixVectorReader): array or coordinate vectors."""
    text = _read_text(src)
    has, w = _header(text)
    if has:
        if w[1] != "vector":
            raise IOError('Expected "vector", got ' + w[1])
        if w[2] not in ("coordinate", "array"):
            raise IOError("Unknown layout " + w[2])
        if w[3] not in ("real", "integer", "complex", "pattern"):
            raise IOError("Unknown field specification " + w[3])
        sparse, field = w[2] == "coordinate", w[3]
    else:
        sparse, field = True, "real"
    tok = _tokens(text.split("\n", 1)[1] if has else text)
    try:
        size = int(next(tok))
        nent = int(next(tok)) if sparse else size
        if field == "pattern":
            raise UnsupportedOperationException("Pattern vectors are not supported")
        if field == "complex":
            raise UnsupportedOperationException("Complex vectors are not supported")
        x = np.zeros(size)
        if sparse:
            for _ in range(nent):
                i = int(next(tok)) - 1
                x[i] = float(next(tok))
        else:
            for i in range(size):
                x[i] = float(next(tok))
    except StopIteration:
        raise EOFError("End-of-File encountered during parsing") from None
    return x


def write_matrix(rowptr, colind, data, nrows: Optional[int] = None,
                 ncols: Optional[int] = None, digits: int = 12, header: bool = True) -> str:
    """Matrix Market text of a CSR matrix: nonzero entries only, row-major."""
    rowptr = np.asarray(rowptr)
    nrows = len(rowptr) - 1 if nrows is None else nrows
    ncols = nrows if ncols is None else ncols
    colind = np.asarray(colind)
    data = np.asarray(data, np.float64)
    rows = np.repeat(np.arange(nrows), np.diff(rowptr))
    nz = data != 0
    out = io.StringIO()
    if header:
        out.write("%%MatrixMarket matrix coordinate real general\n")
    out.write("%10d %10d %19d\n" % (nrows, ncols, int(nz.sum())))
    fmt = "%%10d %%10d %% .%de\n" % digits
    for r, c, v in zip(rows[nz], colind[nz], data[nz]):
        out.write(fmt % (r + 1, c + 1, v))
    return out.getvalue()


def write_vector(x, digits: int = 12, header: bool = True) -> str:
    """Matrix Market array vector (MatrixVectorWriter.printArray)."""
    x = np.asarray(x, np.float64)
    out = io.StringIO()
    if header:
        out.write("%%MatrixMarket vector array real\n")
    out.write("%10d\n" % x.size)
    fmt = "%% .%de\n" % digits
    for v in x:
        out.write(fmt % v)
    return out.getvalue()


def read_comprow(src, ctx=None):
    """A device CompRowMatrix read from a Matrix Market file or stream."""
    from .mtj import CompRowMatrix
    n, m, rp, ci, d = read_matrix(src)
    return CompRowMatrix.from_csr(n, m, rp, ci, d, ctx=ctx)
