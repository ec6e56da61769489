"""Build step of the package: compiles the gfx950 HIP library libgi_hip.so
(and the read-bandwidth benchmark libgi_membench.so) in-tree with hipcc (no
JIT cache), so the built files travel with the tree.

    python -m globillum_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libgi_hip.so")
SOURCES = ["gi_kernels.hip", "bvh_gpu.hip", "bvh_build.cpp", "scene_io.cpp", "image_io.cpp"]
# the memory-hierarchy read benchmark used by bench.py (not part of the path)
MEMBENCH = os.path.join(PKG, "libgi_membench.so")
# exhaustive check of the traversal's reciprocal against the IEEE division
RCP_CHECK = os.path.join(PKG, "gi_rcp_check")
# device evaluation of the sampler for given (index, dimension) pairs
HALTON_CHECK = os.path.join(PKG, "gi_halton_check")

# -ffp-contract=off: hit ids, sample directions and image values are
# reproduced bit-for-bit on the host, so no a*b+c may be fused.
# -packed-fp32-ops: no v_pk_*_f32 instructions in any kernel.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libgi_hip.so)")


HOST_DIR = os.path.join(PKG, "host")
HOST_BINS = {"gi_driver": ["gi_driver.cpp", "mi355x_pt.cpp"], "registry_test": ["registry_test.cpp", "mi355x_pt.cpp"]}


def build_host(verbose: bool = False) -> None:
    """C++ host side: the thin driver and the registry test, linked against
    libgi_hip.so (rpath $ORIGIN)."""
    cxx = shutil.which("g++") or "g++"
    for name, srcs in HOST_BINS.items():
        out = os.path.join(PKG, name)
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-o", out + ".tmp"] + [os.path.join(HOST_DIR, s) for s in srcs]
        cmd += ["-L" + PKG, "-lgi_hip", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"g++ failed building {name}")
        os.replace(out + ".tmp", out)


def build(verbose: bool = False) -> str:
    """Compile every source into libgi_hip.so, the reciprocal and sampler
    checks and the host binaries (always recompiles)."""
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-shared", "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES] + ["-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgi_hip.so")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    tmp = MEMBENCH + ".tmp"
    r = subprocess.run([hipcc(), *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "membench.hip")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgi_membench.so")
    os.replace(tmp, MEMBENCH)
    tmp = RCP_CHECK + ".tmp"
    r = subprocess.run([hipcc(), *[f for f in HIP_FLAGS if f != "-fPIC"], "-o", tmp,
                        os.path.join(CSRC, "rcp_check.hip")], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building gi_rcp_check")
    os.replace(tmp, RCP_CHECK)
    tmp = HALTON_CHECK + ".tmp"
    r = subprocess.run([hipcc(), *[f for f in HIP_FLAGS if f != "-fPIC"], "-o", tmp,
                        os.path.join(CSRC, "halton_check.hip"), os.path.join(CSRC, "bvh_build.cpp"), "-pthread"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building gi_halton_check")
    os.replace(tmp, HALTON_CHECK)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
