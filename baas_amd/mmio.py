"""Matrix Market ingestion with the reference loaders' semantics.

* :func:`load_matrix_market` -- full matrix; symmetric/hermitian files are
  expanded to both triangles, indices shifted 1 -> 0, rows sorted
  (spm/mmio_highlevel.cpp:748-921 ``mmio_allinone_coo`` then
  spm/COO.cpp:284-320 ``dcoo2csr``).
* :func:`load_lower_matrix_market` -- the "lower" loader used by the SpTRSV
  experiments: keeps the strictly lower entries and puts 1.0 on the diagonal
  (spm/COO.cpp:554-622 ``loadLowerMatrixMarket`` /
  ``splitLowerAndAddMajorDiagonal``).

Files may be gzip-compressed (``.gz``).  Pattern matrices get value 1.0.
"""
from __future__ import annotations

import gzip

import numpy as np

from .matgen import Csr


def _open(path):
    return gzip.open(path, "rt") if str(path).endswith(".gz") else open(path, "r")


def read_coo(path):
    """Return (m, n, rows, cols, vals, symmetric) with 0-based indices, expanded."""
    with _open(path) as f:
        banner = f.readline().lower().split()
        if len(banner) < 5 or banner[0] != "%%matrixmarket" or banner[2] != "coordinate":
            raise ValueError(f"{path}: only coordinate Matrix Market files are supported")
        field, sym = banner[3], banner[4]
        if field == "complex":
            raise ValueError(f"{path}: complex matrices are not supported")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        m, n, nnz = (int(t) for t in line.split())
        ncol = 2 if field == "pattern" else 3
        data = np.loadtxt(f, dtype=np.float64, ndmin=2, max_rows=nnz)
    if data.shape[0] != nnz or data.shape[1] < ncol:
        raise ValueError(f"{path}: expected {nnz} entries")
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    vals = np.ones(nnz) if field == "pattern" else data[:, 2].copy()
    symmetric = sym in ("symmetric", "hermitian", "skew-symmetric")
    if symmetric:
        off = rows != cols
        sign = -1.0 if sym == "skew-symmetric" else 1.0
        rows, cols, vals = (np.concatenate([rows, cols[off]]), np.concatenate([cols, rows[off]]),
                            np.concatenate([vals, sign * vals[off]]))
    return m, n, rows, cols, vals, symmetric


def _to_csr(m, rows, cols, vals):
    import scipy.sparse as sp
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m, m))
    return Csr.from_scipy(A)


def load_matrix_market(path) -> Csr:
    m, n, r, c, v, _ = read_coo(path)
    if m != n:
        raise ValueError(f"{path}: matrix is not square ({m}x{n})")
    return _to_csr(m, r, c, v)


def load_lower_matrix_market(path, diag: float = 1.0) -> Csr:
    m, n, r, c, v, _ = read_coo(path)
    if m != n:
        raise ValueError(f"{path}: non-square matrix can not split lower matrix")
    keep = c < r
    d = np.arange(m)
    return _to_csr(m, np.concatenate([r[keep], d]), np.concatenate([c[keep], d]),
                   np.concatenate([v[keep], np.full(m, diag)]))
