"""Builds liblinqt_amd.so (gfx950) in place with hipcc.

Usage: python -m linqt_amd.build   (or linqt_amd.build.build())
Objects go to linqt_amd/_build/, the library to linqt_amd/liblinqt_amd.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblinqt_amd.so")
ARCH = os.environ.get("LQ_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


# Packed FP32 ops are disabled for every device compile.
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [hipcc(), *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


CXX_LIB = os.path.join(HERE, "liblinqt_cxx.so")
BIN = os.path.join(HERE, "bin")
ROOT = os.path.dirname(HERE)
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include", "linqt"),
            "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed:\n{r.stdout}\n{r.stderr}")


def build_cxx(verbose: bool = False) -> None:
    """The C++ facade (reference class names) and the reference's CLI drivers,
    host-only code over the C ABI."""
    cxx = os.environ.get("CXX", "g++")
    tmp = CXX_LIB + ".tmp"
    _run([cxx, *CXXFLAGS, "-shared", *sorted(glob.glob(os.path.join(HERE, "cpp", "*.cpp"))), "-o", tmp,
          "-L", HERE, "-llinqt_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CXX_LIB)
    os.makedirs(BIN, exist_ok=True)
    drivers = sorted(glob.glob(os.path.join(HERE, "cpp", "drivers", "*.cpp")))

    def one(src):
        exe = os.path.join(BIN, os.path.basename(src)[:-4])
        _run([cxx, *CXXFLAGS, src, "-o", exe + ".tmp", "-L", HERE, "-llinqt_cxx", "-llinqt_amd",
              "-Wl,-rpath,$ORIGIN/.."])
        os.replace(exe + ".tmp", exe)

    with cf.ThreadPoolExecutor(max_workers=min(8, len(drivers))) as ex:
        list(ex.map(one, drivers))
    if verbose:
        print("built", CXX_LIB, "and", len(drivers), "drivers in", BIN)


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_cxx(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
