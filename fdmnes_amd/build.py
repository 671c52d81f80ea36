"""Builds libfdmgpu.so (gfx950) in-tree with hipcc.

    python -m fdmnes_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfdmgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["solver.hip", "refine.hip", "batch.hip", "assemble.hip", "tau.hip", "sphere.hip", "tensor.hip", "api.hip"]
HOST_LIB = os.path.join(HERE, "libfdmgpu_host.so")
HOST_SOURCES = ["cluster.cpp"]  # host-only: no HIP runtime dependency

# Packed FP32 instructions are disabled for every device pass; the host pass
# warns that it does not know the feature, which is harmless.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wl,-soname,libfdmgpu.so",
]
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
SHIM_LIB = os.path.join(HERE, "libfdmnes_shim.so")
# test-only native code: the MFMA lane-layout and XCC probe of tests/test_gpu_mfma_layout.py
TESTS = os.path.join(HERE, "..", "tests")
PROBE_LIB = os.path.join(TESTS, "libfdmgpu_probe.so")
SHIM_SRC = os.path.join(HERE, "fortran", "mat_solve_hip.f90")
# the shim is compiled inside FDMNES against its module `declarations`;
# here against the module file of the reference build (oracle/_ref)
REF_MOD_DIR = os.path.join(HERE, "..", "oracle", "_ref")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(HERE, "..", "include"), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libfdmgpu.so")
    os.replace(tmp, LIB)
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-o", tmp,
           *[os.path.join(CSRC, s) for s in HOST_SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building libfdmgpu_host.so")
    os.replace(tmp, HOST_LIB)
    tmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC, *FLAGS[:-1], "-o", tmp, os.path.join(TESTS, "probe.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libfdmgpu_probe.so")
    os.replace(tmp, PROBE_LIB)
    if os.path.exists(os.path.join(REF_MOD_DIR, "declarations.mod")) and os.path.exists(FLANG):
        tmp = SHIM_LIB + ".tmp"
        cmd = [FLANG, "-O2", "-fPIC", "-shared", "-I", REF_MOD_DIR, "-module-dir", "/tmp", "-o", tmp, SHIM_SRC,
               "-L", HERE, "-lfdmgpu", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed building libfdmnes_shim.so")
        os.replace(tmp, SHIM_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
