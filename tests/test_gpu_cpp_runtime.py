"""The C++ runtime (baas_amd/host) driven like the reference's mains.

* baas_demo: x = 1 manufactured SpTRSV, IC0 and ILU0 against serial factors
  at 1e-10, and the ICCG scheduler's IC0 + forward/backward solves.
* tests/cpp/*: drivers with the reference mains' shape (same types, class
  names, constructor arguments and calls; test/SpTRSV_level_merge.cpp,
  test/ic0_level_merge.cpp, test/ilu0_level_merge.cpp, the Scheduler use of
  test/iccg_level_merge.cpp) on Matrix Market files.  Each exits non-zero if
  a runtime's own check (x = alpha, IC0/ILU0 against the serial factor at
  1e-10, U = L^T, ICCG convergence) fails.
"""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "baas_amd", "bin", "baas_demo")
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def _run(exe, *args, cwd=None):
    announce(os.path.basename(exe), *args)
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300, cwd=cwd)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("args", [("trsv", 32, 8, 20), ("trsv", 64, 2, 10), ("ic0", 24, 8, 5), ("ilu0", 24, 8, 5),
                                  ("sched", 20, 8, 1)])
def test_runtime_generated(args):
    out = _run(DEMO, *args)
    assert out.count(",") >= 8


@pytest.fixture(scope="module")
def mtx(tmp_path_factory):
    """bcsstk18 (the reference's data file), a 20^3 Laplacian and a 16^3
    convection-diffusion operator as Matrix Market files."""
    import scipy.io
    import scipy.sparse as sp
    from baas_amd import matgen as g
    d = tmp_path_factory.mktemp("mtx")
    with gzip.open(os.path.join(ROOT, "tests", "golden", "bcsstk18.mtx.gz"), "rb") as fi, \
            open(d / "bcsstk18.mtx", "wb") as fo:
        shutil.copyfileobj(fi, fo)
    scipy.io.mmwrite(str(d / "lap3d20.mtx"), sp.tril(g.laplace3d(20).to_scipy()).tocoo(), symmetry="symmetric")
    scipy.io.mmwrite(str(d / "convdiff16.mtx"), g.convdiff3d(16).to_scipy().tocoo())
    return d


def test_runtime_bcsstk18(mtx):
    _run(DEMO, "trsv", mtx / "bcsstk18.mtx", 8, 10)
    _run(DEMO, "ic0", mtx / "bcsstk18.mtx", 8, 3)
    _run(DEMO, "ilu0", mtx / "bcsstk18.mtx", 8, 3)


@pytest.mark.parametrize("name", ["bcsstk18", "lap3d20"])
def test_reference_shaped_ic0_main(mtx, name):
    out = _run(os.path.join(BIN, "ic0_level_merge"), mtx / f"{name}.mtx", 8, mtx / f"ic0_{name}", 5, cwd=mtx)
    assert out.count("code IC0 check pass.") == 3 and "the transposed U matrix is equal to L matrix!" in out
    rows = open(mtx / f"ic0_{name}.csv").read().strip().splitlines()
    assert rows[0].startswith("Matrix_Name,row,nnz") and len(rows) >= 4


def test_reference_shaped_sptrsv_main(mtx):
    out = _run(os.path.join(BIN, "sptrsv_level_merge"), mtx / "lap3d20.mtx", 8, mtx / "trsv", 5, cwd=mtx)
    assert out.count("code solutions check pass.") == 8
    rows = open(mtx / "trsv.csv").read().strip().splitlines()
    assert len(rows) == 1 + 7


@pytest.mark.parametrize("name", ["convdiff16", "bcsstk18"])
def test_reference_shaped_ilu0_main(mtx, name):
    out = _run(os.path.join(BIN, "ilu0_level_merge"), mtx / f"{name}.mtx", 8, mtx / f"ilu0_{name}", 5, cwd=mtx)
    assert out.count("code ILU0 check pass.") == 3


@pytest.mark.parametrize("name", ["lap3d20", "bcsstk18"])
def test_scheduler_iccg(mtx, name):
    out = _run(os.path.join(BIN, "iccg_scheduler"), mtx / f"{name}.mtx", 8, 1e-8, 5000)
    assert "ICCG:" in out and "ICCG perm:" in out
