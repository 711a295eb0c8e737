"""Named parity workloads shared by the tests and tests/golden/make_golden.py."""
import functools
import os

from baas_amd import matgen as g
from baas_amd import mmio

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@functools.lru_cache(maxsize=None)
def workload(name):
    """name -> (lower Csr, rowptr for the thread rule, full Csr or None)."""
    if name == "C1_lap2d_128":
        L = g.laplace2d_lower(128)
        return L, L.rowptr, None
    if name == "C2_lap3d_64":
        A = g.laplace3d(64)
    elif name == "C2_lap3d_64_perm1":
        A, _ = g.random_symmetric_permutation(g.laplace3d(64), 1)
    elif name == "C2_convdiff_64":
        A = g.convdiff3d(64)
    elif name == "C2_convdiff_64_perm1":
        A, _ = g.random_symmetric_permutation(g.convdiff3d(64), 1)
    elif name == "C3_knn":
        A, _ = g.knn_stiffness()
    elif name == "C5_diff3d_48":
        A, _ = g.diffusion3d_varcoef(48)
    elif name in ("bcsstk18", "bcsstk38"):
        A = mmio.load_matrix_market(os.path.join(GOLDEN, name + ".mtx.gz"))
    else:
        raise KeyError(name)
    return A.lower(), A.rowptr, A


NAMES = ("C1_lap2d_128", "C2_lap3d_64", "C2_lap3d_64_perm1", "C2_convdiff_64", "C2_convdiff_64_perm1",
         "C3_knn", "C5_diff3d_48", "bcsstk18", "bcsstk38")
# nonsymmetric matrices: ILU0 only (IC0 needs SPD)
NONSYMMETRIC = ("C2_convdiff_64", "C2_convdiff_64_perm1")
