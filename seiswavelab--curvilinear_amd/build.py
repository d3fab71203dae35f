"""Builds csrc/libswl.so for gfx950 with hipcc (in-tree, no JIT cache).

Every HIP translation unit is compiled with -ffp-contract=off (the reference's CUDA build
uses --fmad=false and its CPU path has no contraction, so FP64 results stay bitwise
comparable) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
INCLUDE = ROOT / "include"
LIB = CSRC / "libswl.so"

HIP_SOURCES = ["engine2d.hip", "engine3d.hip", "geometry2d.hip", "geometry3d.hip"]
GCC_OPENMP_SOURCES = {"vts2d.cpp", "vts3d.cpp"}
CXX_SOURCES = ["model2d.cpp", "sim2d.cpp", "model3d.cpp", "sim3d.cpp", "batch2d.cpp", "segy.cpp", "async2d.cpp", "async3d.cpp", "vts2d.cpp", "vts3d.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


COMMON = ["-O3", "-fPIC", "-std=c++20", "-ffp-contract=off", f"-I{INCLUDE}", f"-I{CSRC}"]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def _compile(src: pathlib.Path, obj: pathlib.Path, verbose: bool):
    if src.name.startswith("model"):
        # FP64 session construction is plain host C++: build it with g++ so libm / complex
        # arithmetic round exactly as in the reference's CPU build.
        cmd = ["g++"] + COMMON + ["-fopenmp", "-c", str(src), "-o", str(obj)]
    elif src.name in GCC_OPENMP_SOURCES:
        # host code with OpenMP loops: one OpenMP runtime (libgomp, linked by g++) for the whole library
        cmd = ["g++"] + COMMON + ["-fopenmp", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", str(src), "-o", str(obj)]
    else:
        flags = COMMON + (HIP_FLAGS if src.suffix == ".hip" else
                          ["-x", "c++", "-fopenmp", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"])
        cmd = [hipcc()] + flags + ["-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> pathlib.Path:
    out = CSRC / "_obj"
    out.mkdir(exist_ok=True)
    sources = [CSRC / s for s in HIP_SOURCES + CXX_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(lambda s: _compile(s, out / (s.name + ".o"), verbose), sources))
    tmp = LIB.with_suffix(".so.tmp")
    # Link with g++ so complex-arithmetic helpers (__divdc3) come from libgcc, as in the
    # reference's CPU build; the device code objects are embedded in each HIP object file.
    rocm = pathlib.Path(hipcc()).resolve().parents[1]
    cmd = (["g++", "-shared", "-fopenmp", "-o", str(tmp)] + [str(o) for o in objs] +
           [f"-L{rocm / 'lib'}", f"-Wl,-rpath,{rocm / 'lib'}", "-lamdhip64", "-Wl,--no-undefined"])
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
