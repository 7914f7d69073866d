"""Build steps for the gfx950 library and the CPU checker.

``build_library()`` compiles csrc/*.hip with hipcc into
``flexrt_imaging_reion_amd/libflexrt_gpu.so`` (in-tree, so it travels with the
repository snapshot).  ``build_oracle()`` compiles the CPU restatement under
``oracle/`` with g++; it is test infrastructure and only tests, smoke() and the
CPU-baseline leg of bench.py load it.
"""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libflexrt_gpu.so")
ORACLE_SRC = os.path.join(REPO_DIR, "oracle", "frt_oracle.cc")
ORACLE_LIB = os.path.join(REPO_DIR, "oracle", "liboracle.so")
ORACLE_OMP_LIB = os.path.join(REPO_DIR, "oracle", "liboracle_omp.so")

HIP_SOURCES = ["frt_kernels.hip", "frt_capi.hip", "frt_adjoint.hip", "frt_scan_sort.hip"]
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # no implicit a*b+c fusion: the path must reproduce the CPU arithmetic
    "-ffp-contract=off",
    # keep the compiler from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))


def hipcc_path():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    p = os.path.join(rocm, "bin", "hipcc")
    return p if os.path.exists(p) else "hipcc"


def build_library(out=LIB_PATH):
    tmp = out + ".tmp"
    cmd = [hipcc_path()] + HIPCC_FLAGS + ["-o", tmp] + HIP_SOURCES
    _run(cmd, CSRC)
    os.replace(tmp, out)
    return out


def build_oracle(out=ORACLE_LIB, openmp=False):
    """The serial build is the checker; the OpenMP build (same source, the
    reference's parallel loops over all threads) is the timed CPU baseline."""
    tmp = out + ".tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-mfma", "-ffp-contract=off",
           "-fno-fast-math", "-Wno-unknown-pragmas"] + (["-fopenmp"] if openmp else []) + ["-o", tmp, ORACLE_SRC]
    _run(cmd, os.path.dirname(ORACLE_SRC))
    os.replace(tmp, out)
    return out


def build_oracle_omp(out=ORACLE_OMP_LIB):
    return build_oracle(out, openmp=True)
