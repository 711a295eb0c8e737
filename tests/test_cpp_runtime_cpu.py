"""The C++ runtime's Matrix Market loaders (SPM::CSR(file) and the lower
loader CSR(file, true, 0)) against the reference loader's digests for
bcsstk18 (tests/golden, written from the compiled reference).  CPU only: the
loader makes no GPU call."""
import gzip
import hashlib
import os
import shutil
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def _tool(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        from baas_amd import build as B
        B.build_demo()
    return path


def test_cpp_loader_matches_reference_loader(golden, tmp_path):
    man, _ = golden
    src = os.path.join(ROOT, "tests", "golden", "bcsstk18.mtx.gz")
    mtx = tmp_path / "bcsstk18.mtx"
    with gzip.open(src, "rb") as fi, open(mtx, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    r = subprocess.run([_tool("mtx_dump"), str(mtx), str(tmp_path / "b18")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    for lower, tag in ((0, "full"), (1, "lower")):
        want = man["inputs"][f"bcsstk18_loader_lower{lower}"]
        arrs = {k: np.fromfile(tmp_path / f"b18.{tag}.{k}", dtype=dt)
                for k, dt in (("rowptr", np.int32), ("colidx", np.int32), ("values", np.float64))}
        assert len(arrs["rowptr"]) - 1 == want["n"] and len(arrs["colidx"]) == want["nnz"]
        for k, a in arrs.items():
            assert hashlib.sha256(a.tobytes()).hexdigest() == want[k], (tag, k)


def test_reference_shaped_mains_are_built():
    for name in ("ic0_level_merge", "ilu0_level_merge", "sptrsv_level_merge", "iccg_scheduler"):
        assert os.access(_tool(name), os.X_OK), name
