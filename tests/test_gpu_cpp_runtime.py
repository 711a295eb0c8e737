"""The C++ runtime classes (baas_amd/host), driven like the reference's mains:
x = 1 manufactured SpTRSV, IC0 and ILU0 against serial factors at 1e-10, and
the ICCG scheduler's IC0 + forward/backward solves."""
import gzip
import os
import shutil
import subprocess

import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "baas_amd", "bin", "baas_demo")


def _run(*args):
    announce("baas_demo", *args)
    r = subprocess.run([DEMO, *map(str, args)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("args", [("trsv", 32, 8, 20), ("trsv", 64, 2, 10), ("ic0", 24, 8, 5), ("ilu0", 24, 8, 5),
                                  ("sched", 20, 8, 1)])
def test_runtime_generated(args):
    out = _run(*args)
    assert out.count(",") >= 8


def test_runtime_bcsstk18(tmp_path):
    src = os.path.join(ROOT, "tests", "golden", "bcsstk18.mtx.gz")
    dst = tmp_path / "bcsstk18.mtx"
    with gzip.open(src, "rb") as fi, open(dst, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    _run("trsv", dst, 8, 10)
    _run("ic0", dst, 8, 3)
    _run("ilu0", dst, 8, 3)
