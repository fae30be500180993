"""Build recipe of the soapfilm HIP library (libsoapfilm.so) for gfx950.

Every ``csrc/*.hip`` file is compiled by hipcc into one shared library next
to this file; the C ABI is declared in ``include/soapfilm.h`` at the
repository root.  The build runs on a machine without a GPU (hipcc
cross-compiles for gfx950).

    python -m soap_bubble_simulation_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO_DIR, "include")
OBJ_DIR = os.path.join(PKG_DIR, "_build")
LIB_PATH = os.path.join(PKG_DIR, "libsoapfilm.so")
ARCH = os.environ.get("SOAPFILM_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# Packed FP32 VALU instructions are switched off (-packed-fp32-ops): the
# kernels are written in scalar float and must not be turned into
# v_pk_{add,mul,fma}_f32.
COMMON_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-I", INCLUDE, "-I", CSRC,
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed on %s" % os.path.basename(src))
    return obj


PROBE_SRC = os.path.join(PKG_DIR, "probe", "bandwidth.hip")
PROBE_LIB = os.path.join(PKG_DIR, "probe", "libsfprobe.so")


def build_probe() -> str:
    """Compile the streaming-bandwidth probes (probe/bandwidth.hip) into
    libsfprobe.so: measurement tooling for the bench, not part of the
    soapfilm library."""
    tmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC, *COMMON_FLAGS, "-shared", PROBE_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed on probe/bandwidth.hip")
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


CPP_TEST_SRC = os.path.join(REPO_DIR, "tests", "cpp", "test_api.cpp")
CPP_TEST_BIN = os.path.join(REPO_DIR, "tests", "cpp", "test_api")


def build_cpp_test() -> str:
    """Compile the C++ API test (tests/cpp/test_api.cpp) against the library."""
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I", INCLUDE, CPP_TEST_SRC, "-o", CPP_TEST_BIN,
           "-L", PKG_DIR, "-lsoapfilm", "-Wl,-rpath," + PKG_DIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("C++ API test failed to compile")
    return CPP_TEST_BIN


def build(verbose: bool = False) -> str:
    """Compile all HIP sources and link libsoapfilm.so; returns its path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libsoapfilm.so failed")
    os.replace(tmp, LIB_PATH)
    if os.path.exists(CPP_TEST_SRC):
        build_cpp_test()
    build_probe()
    if verbose:
        print("built", LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
