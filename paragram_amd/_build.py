"""Builds libparagram_hip.so in-tree with hipcc for gfx950.

Mirrors the role of the reference's lazy JIT build (paragram/_backend.py:147-190)
with an explicit, in-tree hipcc step.  Environment knobs with the reference's
names: ``MAX_JOBS`` (parallel compiles), ``FAST_COMPILE=1`` (-O0),
``DEBUG_CUDA=1`` (-O0 -g), ``PROFILE_CUDA=1`` (-gline-tables-only),
``VERBOSE=1`` (print commands).  ``NO_FAST_MATH`` is accepted and has no
effect: the predicates never use fast math.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libparagram_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
CPU_LIB_PATH = os.path.join(HERE, "libparagram_cpu.so")
SOURCES = ["pg_abi.hip", "pg_index.hip", "pg_cell.hip", "pg_csr.hip", "pg_knn.hip", "pg_mlp.hip"]
ARCH = "gfx950"


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; a ROCm toolchain is required to build paragram_amd")


def _flags() -> list[str]:
    opt = ["-O3"]
    if os.environ.get("FAST_COMPILE") == "1":
        opt = ["-O0"]
    if os.environ.get("DEBUG_CUDA") == "1":
        opt = ["-O0", "-g"]
    if os.environ.get("PROFILE_CUDA") == "1":
        opt += ["-gline-tables-only"]
    return [f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", *opt,
            # packed FP32 instructions are kept out of the device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            f"-I{INCLUDE}", "-Wno-unused-command-line-argument"]


def build(verbose: bool | None = None, lib_path: str = LIB_PATH) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    if verbose is None:
        verbose = os.environ.get("VERBOSE") == "1"
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = _flags()
    jobs = int(os.environ.get("MAX_JOBS", "4"))

    def compile_one(src: str) -> str:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = lib_path + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, lib_path)
    return lib_path


def build_cpu(verbose: bool | None = None, lib_path: str = CPU_LIB_PATH) -> str:
    """The host-core build (csrc/pg_cpu.cpp, C++/OpenMP) with the system C++
    compiler; returns its path."""
    if verbose is None:
        verbose = os.environ.get("VERBOSE") == "1"
    cxx = os.environ.get("CXX") or shutil.which("g++") or shutil.which("c++")
    if not cxx:
        raise RuntimeError("no C++ compiler found for the host-core build")
    opt = ["-O0", "-g"] if os.environ.get("DEBUG_CUDA") == "1" else ["-O2"]
    tmp = lib_path + ".tmp"
    cmd = [cxx, *opt, "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off", "-Wno-unknown-pragmas",
           f"-I{INCLUDE}", os.path.join(CSRC, "pg_cpu.cpp"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host-core build failed:\n{r.stderr}")
    os.replace(tmp, lib_path)
    return lib_path


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cpu(verbose=True))
