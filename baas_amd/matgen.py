"""Deterministic generators for the benchmark and parity workloads.

Every matrix comes back as a :class:`Csr` (0-based int32 ``rowptr``/``colidx``,
float64 ``values``, rows sorted ascending by column, square), the layout of the
reference's ``SPM::CSR`` (spm/include/CSR.hpp:16-27).  Random draws use
``numpy.random.default_rng(seed)`` (PCG64) in the order documented at each
function, so the same inputs are rebuilt on any machine without the reference.

Workloads (sizes are the ones the parity tests and ``bench.py`` use):

* ``laplace2d_lower(128)``   -- lower triangle (diag 4, off -1) of the 2D
  5-point Laplacian, natural ordering, n = 16384; b ~ U(-1,1), seed 0.
* ``laplace3d(64)``          -- 3D 7-point Laplacian (diag 6, off -1),
  n = 262144, natural ordering; ``random_symmetric_permutation(A, 1)``.
* ``convdiff3d(64)``         -- nonsymmetric 7-point convection-diffusion.
* ``knn_stiffness()``        -- 12-NN 3x3-block SPD matrix, n = 36000.
* ``batch_laplace2d()``      -- 256 2D Laplacians, sides U[16,256], seed 3.
* ``diffusion3d_varcoef(48)`` -- lognormal edge conductivities, seed 4.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Csr:
    n: int
    rowptr: np.ndarray  # int32[n+1]
    colidx: np.ndarray  # int32[nnz]
    values: np.ndarray  # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values, self.colidx, self.rowptr), shape=(self.n, self.n))

    @staticmethod
    def from_scipy(m) -> "Csr":
        m = m.tocsr()
        m.sort_indices()
        m.sum_duplicates()
        return Csr(m.shape[0], m.indptr.astype(np.int32), m.indices.astype(np.int32),
                   m.data.astype(np.float64))

    def diagptr(self) -> np.ndarray:
        """Position of the first entry with column >= row (spm/CSR.cpp:853-871)."""
        rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.rowptr))
        below = np.bincount(rows[self.colidx < rows], minlength=self.n)
        return (self.rowptr[:-1] + below).astype(np.int32)

    def lower(self) -> "Csr":
        """Lower triangle including the diagonal (test/ICRuntime.hpp:416-435)."""
        dp = self.diagptr()
        lens = dp - self.rowptr[:-1] + 1
        rp = np.zeros(self.n + 1, dtype=np.int32)
        np.cumsum(lens, out=rp[1:])
        keep = np.zeros(self.nnz, dtype=bool)
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        keep[:] = self.colidx <= rows
        return Csr(self.n, rp, self.colidx[keep].copy(), self.values[keep].copy())

    def upper(self) -> "Csr":
        """Upper triangle including the diagonal."""
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        keep = self.colidx >= rows
        lens = np.bincount(rows[keep], minlength=self.n)
        rp = np.zeros(self.n + 1, dtype=np.int32)
        np.cumsum(lens, out=rp[1:])
        return Csr(self.n, rp, self.colidx[keep].copy(), self.values[keep].copy())


def _grid_operator(shape, diag, offs):
    """Assemble a structured-grid operator.  ``offs`` lists (axis, step, coef)."""
    import scipy.sparse as sp
    n = int(np.prod(shape))
    idx = np.arange(n).reshape(shape[::-1])  # index = x + nx*(y + ny*z)
    rows = [np.arange(n)]
    cols = [np.arange(n)]
    vals = [np.broadcast_to(np.asarray(diag, dtype=np.float64), (n,)).ravel()]
    for axis, step, coef in offs:
        ax = len(shape) - 1 - axis  # axis 0 = x is the fastest
        sl_src = [slice(None)] * len(shape)
        sl_dst = [slice(None)] * len(shape)
        if step > 0:
            sl_src[ax] = slice(0, shape[axis] - step)
            sl_dst[ax] = slice(step, None)
        else:
            sl_src[ax] = slice(-step, None)
            sl_dst[ax] = slice(0, shape[axis] + step)
        r = idx[tuple(sl_src)].ravel()
        c = idx[tuple(sl_dst)].ravel()
        rows.append(r)
        cols.append(c)
        vals.append(np.full(r.size, coef, dtype=np.float64))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    return Csr.from_scipy(A)


def laplace2d(side: int) -> Csr:
    """2D 5-point Laplacian, diag 4, off -1, natural ordering (x fastest)."""
    offs = [(a, s, -1.0) for a in (0, 1) for s in (-1, 1)]
    return _grid_operator((side, side), 4.0, offs)


def laplace2d_lower(side: int = 128) -> Csr:
    return laplace2d(side).lower()


def laplace_grid(shape) -> Csr:
    """Laplacian on a grid of any shape (x fastest, natural ordering): 2 x
    dimensions on the diagonal, -1 to each grid neighbour."""
    offs = [(a, s, -1.0) for a in range(len(shape)) for s in (-1, 1)]
    return _grid_operator(tuple(shape), 2.0 * len(shape), offs)


def laplace3d(side: int = 64) -> Csr:
    """3D 7-point Laplacian, diag 6, off -1, natural ordering (x fastest)."""
    offs = [(a, s, -1.0) for a in (0, 1, 2) for s in (-1, 1)]
    return _grid_operator((side, side, side), 6.0, offs)


def convdiff3d(side: int = 64, vel=(0.5, 0.3, 0.2), dominance: float = 1.05) -> Csr:
    """Nonsymmetric 7-point convection-diffusion, scaled by h^2.

    Diffusion 1 gives 6 / -1; first-order upwind convection with velocity
    vel*h^-1 (all components positive) adds vel[d] to the diagonal and
    -vel[d] to the upwind (minus-side) neighbour.  The diagonal is then the
    stencil diagonal (6 + sum(vel)) times ``dominance``, so every row is
    strictly diagonally dominant (interior rows have sum |off| = 6 + sum(vel)).
    """
    offs = []
    for a in (0, 1, 2):
        offs.append((a, -1, -1.0 - float(vel[a])))
        offs.append((a, +1, -1.0))
    diag = (6.0 + float(sum(vel))) * dominance
    return _grid_operator((side, side, side), diag, offs)


def random_symmetric_permutation(A: Csr, seed: int = 1):
    """P A P^T with perm = default_rng(seed).permutation(n); new row i is old row perm[i]."""
    perm = np.random.default_rng(seed).permutation(A.n)
    M = A.to_scipy()[perm][:, perm]
    return Csr.from_scipy(M), perm


def symmetric_permute(A: Csr, perm_to_orig) -> Csr:
    """P A P^T for a given order (new row i is old row perm_to_orig[i]),
    rows sorted by column: the renumbering by a schedule's execution order
    (Schedule.unit_order()) that makes each work unit's rows consecutive."""
    perm = np.asarray(perm_to_orig, dtype=np.int64)
    return Csr.from_scipy(A.to_scipy()[perm][:, perm])


def rhs_uniform(n: int, seed: int = 0, k: int = 1) -> np.ndarray:
    r = np.random.default_rng(seed).uniform(-1.0, 1.0, size=(n, k) if k > 1 else n)
    return np.ascontiguousarray(r)


def knn_stiffness(npoints: int = 12000, k: int = 12, seed: int = 2):
    """SPD stiffness-like matrix from a k-NN graph expanded to 3x3 blocks.

    Draw order from default_rng(seed): points U[0,1)^3 (npoints x 3), then one
    weight w ~ U(0.1, 1) per undirected edge in (i<j) row-major order, then
    b ~ N(0,1) of length 3*npoints.  The graph is the union-symmetrised k-NN
    graph (k neighbours besides the point itself, scipy cKDTree).  Off-diagonal
    blocks are -w I, diagonal blocks (1.01 * sum_j w_ij + 1e-3) I.
    Returns (A, b).
    """
    import scipy.sparse as sp
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    pts = rng.random((npoints, 3))
    _, nb = cKDTree(pts).query(pts, k=k + 1)
    r = np.repeat(np.arange(npoints), k)
    c = nb[:, 1:].ravel()
    G = sp.csr_matrix((np.ones(r.size), (r, c)), shape=(npoints, npoints))
    G = ((G + G.T) > 0).astype(np.float64).tocsr()
    Gu = sp.triu(G, k=1).tocoo()
    order = np.lexsort((Gu.col, Gu.row))
    ei, ej = Gu.row[order], Gu.col[order]
    w = rng.uniform(0.1, 1.0, size=ei.size)
    W = sp.csr_matrix((np.concatenate([w, w]), (np.concatenate([ei, ej]), np.concatenate([ej, ei]))),
                      shape=(npoints, npoints))
    rowsum = np.asarray(W.sum(axis=1)).ravel()
    Dn = sp.diags(1.01 * rowsum + 1e-3)
    node = (Dn - W).tocsr()
    A = sp.kron(node, sp.identity(3), format="csr")
    b = rng.standard_normal(3 * npoints)
    return Csr.from_scipy(A), b


def batch_laplace2d(nsys: int = 256, lo: int = 16, hi: int = 256, nrhs: int = 32, seed: int = 3):
    """``nsys`` 2D 5-point Laplacians with sides ~ U{lo..hi} and ``nrhs`` RHS each.

    Draw order from default_rng(seed): the sides (integers, inclusive), then for
    each system in order its RHS block (n_i x nrhs, row-major) ~ U(-1,1).
    Returns (list of Csr, list of rhs arrays).
    """
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, size=nsys)
    mats, rhs = [], []
    for s in sides:
        A = laplace2d(int(s))
        mats.append(A)
        rhs.append(np.ascontiguousarray(rng.uniform(-1.0, 1.0, size=(A.n, nrhs))))
    return mats, rhs


def diffusion3d_varcoef(side: int = 48, seed: int = 4, shift: float = 1e-2):
    """3D 7-point variable-coefficient diffusion with lognormal(0,1) edge conductivity.

    Draw order from default_rng(seed): x-edges, y-edges, z-edges (each in
    natural order of their lower endpoint, x fastest), then x* ~ N(0,1) (n).
    Off-diagonals are -k_e, the diagonal is sum of incident k_e plus ``shift``.
    Returns (A, x_star).
    """
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    n = side ** 3
    idx = np.arange(n).reshape(side, side, side)  # [z, y, x]
    rows, cols, ks = [], [], []
    for ax in (2, 1, 0):  # x, y, z
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        src[ax] = slice(0, side - 1)
        dst[ax] = slice(1, None)
        a = idx[tuple(src)].ravel()
        bb = idx[tuple(dst)].ravel()
        order = np.argsort(a, kind="stable")
        a, bb = a[order], bb[order]
        k = rng.lognormal(0.0, 1.0, size=a.size)
        rows += [a, bb]
        cols += [bb, a]
        ks += [k, k]
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    ks = np.concatenate(ks)
    W = sp.csr_matrix((ks, (rows, cols)), shape=(n, n))
    d = np.asarray(W.sum(axis=1)).ravel() + shift
    A = (sp.diags(d) - W).tocsr()
    xs = rng.standard_normal(n)
    return Csr.from_scipy(A), xs
