"""Build recipes for the HIP library (libcloudvox.so) and the CPU oracle.

Both are compiled in-tree so the shared objects travel with the source tree.
Floating-point contraction is disabled on host and device
(``-ffp-contract=off``) and device division / square root stay IEEE
correctly rounded, so FP32 results of the kernels and of the oracle agree
bit for bit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
LIB_PATH = PKG_DIR / "libcloudvox.so"
ORACLE_DIR = REPO_DIR / "oracle"
ORACLE_LIB = ORACLE_DIR / "libcv_oracle.so"

HIP_SOURCES = ["cv_kernels.hip", "cv_mlp.hip", "cv_runtime.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (expected /opt/rocm/bin/hipcc)")


def _run(cmd: list[str], cwd: Path) -> None:
    proc = subprocess.run(cmd, cwd=str(cwd), capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        sys.stderr.write(proc.stderr)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def build_library(verbose: bool = False) -> Path:
    """Compile the gfx950 kernels and the host runtime into libcloudvox.so."""
    hipcc = _hipcc()
    build_dir = PKG_DIR / "_build"
    build_dir.mkdir(exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = build_dir / (Path(src).stem + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, "-I", str(REPO_DIR / "include"), "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd, CSRC)
        objs.append(str(obj))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd))
    _run(cmd, CSRC)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


CPP_FRAME_SRC = REPO_DIR / "tests" / "cpp" / "reference_frame.cpp"
CPP_FRAME_BIN = PKG_DIR / "_build" / "reference_frame"


def cpp_frame_command(out: Path = CPP_FRAME_BIN) -> list:
    """g++ line of the C++ facade test program (links libcloudvox.so and the HIP runtime)."""
    rocm = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
    return ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I", str(REPO_DIR / "include"),
            "-I", str(rocm / "include"), str(CPP_FRAME_SRC), "-o", str(out), "-L", str(PKG_DIR), "-lcloudvox",
            "-L", str(rocm / "lib"), "-lamdhip64", f"-Wl,-rpath,$ORIGIN/..:{rocm / 'lib'}"]


def build_cpp_frame(verbose: bool = False) -> Path:
    """Compile the reference-frame program that drives the C++ facade (a test of the boundary)."""
    CPP_FRAME_BIN.parent.mkdir(exist_ok=True)
    cmd = cpp_frame_command()
    if verbose:
        print(" ".join(cmd))
    _run(cmd, REPO_DIR)
    return CPP_FRAME_BIN


def build_oracle(verbose: bool = False) -> Path:
    """Compile the CPU oracle (test infrastructure) with the recipe in oracle/Makefile."""
    cmd = ["make", "-s", "-C", str(ORACLE_DIR)]
    if verbose:
        print(" ".join(cmd))
    _run(cmd, ORACLE_DIR)
    return ORACLE_LIB


def build_all(verbose: bool = False) -> None:
    build_library(verbose)
    build_oracle(verbose)
    build_cpp_frame(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print(LIB_PATH)
    print(ORACLE_LIB)
