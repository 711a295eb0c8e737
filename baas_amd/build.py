"""Build libbaas_hip.so (every HIP source of baas_amd/csrc) for gfx950.

    python -m baas_amd.build        # or __graft_entry__.build()

Objects go to baas_amd/_build/, the library to baas_amd/libbaas_hip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbaas_hip.so")
ARCH = os.environ.get("BAAS_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every fused multiply-add in the kernels is written out
# (the reference's contraction pattern), nothing else may be fused.
# -packed-fp32-ops: keep hipcc from emitting packed FP32 instructions.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [_hipcc()] + FLAGS + lang + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_demo()
    if verbose:
        print("built", LIB, "and", DEMO)
    return LIB


DEMO = os.path.join(HERE, "bin", "baas_demo")
ROOT = os.path.dirname(HERE)
CPP_MAINS = os.path.join(ROOT, "tests", "cpp")  # drivers shaped like the reference's mains


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_demo():
    """The C++ runtime classes (host/), the driver baas_demo and the
    reference-shaped mains under tests/cpp, linked to the library."""
    os.makedirs(os.path.dirname(DEMO), exist_ok=True)
    host = os.path.join(HERE, "host")
    os.makedirs(BUILD, exist_ok=True)
    cflags = ["-O2", "-std=c++17", "-fopenmp", "-Wno-unused-result", "-I" + host]
    rt_obj = os.path.join(BUILD, "baas_runtime.o")
    _run([_hipcc()] + cflags + ["-c", os.path.join(host, "baas_runtime.cpp"), "-o", rt_obj], "runtime build")
    targets = [(os.path.join(host, "baas_demo.cpp"), DEMO)]
    if os.path.isdir(CPP_MAINS):
        os.makedirs(os.path.join(CPP_MAINS, "bin"), exist_ok=True)
        for f in sorted(os.listdir(CPP_MAINS)):
            if f.endswith(".cpp"):
                targets.append((os.path.join(CPP_MAINS, f), os.path.join(CPP_MAINS, "bin", f[:-4])))

    def link(t):
        src, out = t
        obj = os.path.join(BUILD, "host_" + os.path.basename(src) + ".o")
        _run([_hipcc()] + cflags + ["-c", src, "-o", obj], "host build")
        rpath = "$ORIGIN/" + os.path.relpath(HERE, os.path.dirname(out))
        _run([_hipcc(), "-fopenmp", obj, rt_obj, "-L" + HERE, "-lbaas_hip", "-Wl,-rpath," + rpath, "-o", out + ".tmp"],
             "host link")
        os.replace(out + ".tmp", out)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(targets))) as ex:
        list(ex.map(link, targets))


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
