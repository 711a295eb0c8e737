"""Finite-difference reference for the ICCG adjoint at 48^3 (CPU, oracle)."""
import concurrent.futures as cf
import os

import numpy as np

import oracle as O
from baas_amd import matgen as g

H = 1e-6


def iccg_fd_48(nentries=64, workers=None):
    """Central finite differences of loss = 0.5 ||A^-1 b - x*||^2 on the 48^3
    variable-coefficient diffusion matrix, for `nentries` entries of b and
    as many stored entries of A (one in each row drawn), step 1e-6.

    All on the CPU oracle's FP64 PCG at tol 1e-14.  With x = A^-1 b and
    u = A^-1 e_i:
    * b + h e_i has the solution x + h u exactly;
    * A + h e_i e_j^T has the solution x - h x_j u / (1 + h u_j)
      (Sherman-Morrison),
    so every perturbed problem is solved exactly up to the accuracy of u and
    x, and no perturbed matrix is factorised.  Returns
    (A, x*, b, rows, entries, fd_b, fd_A)."""
    A, xs = g.diffusion3d_varcoef(48)
    rng = np.random.default_rng(42)
    b = rng.standard_normal(A.n)
    rows_i = np.sort(rng.choice(A.n, nentries, replace=False))
    ents = np.array([rng.integers(A.rowptr[i], A.rowptr[i + 1]) for i in rows_i])

    def tight(rhs):
        x, k, _ = O.pcg(A, rhs, maxiter=20000, tol=1e-14)
        assert k > 0
        return x

    def unit(i):
        e = np.zeros(A.n)
        e[i] = 1.0
        return tight(e)
    # the oracle's PCG releases the GIL (ctypes), so the solves run in parallel
    w = workers or min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=w) as ex:
        fx = ex.submit(tight, b)
        us = list(ex.map(unit, rows_i))
        x = fx.result()

    def fd(xp, xm):
        return 0.5 * float(np.dot(xp - xm, xp + xm - 2 * xs)) / (2 * H)
    fd_b = np.array([fd(x + H * u, x - H * u) for u in us])
    fd_a = []
    for u, e in zip(us, ents):
        j = A.colidx[e]
        fd_a.append(fd(x - H * x[j] * u / (1 + H * u[j]), x + H * x[j] * u / (1 - H * u[j])))
    return A, xs, b, rows_i, ents, fd_b, np.array(fd_a)
