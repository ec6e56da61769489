"""The host builder's binary -> 4-wide collapse (surface-area cost, dynamic
program over slot counts), checked on the CPU by tests/csrc/collapse_check.cpp
compiled against the library's own bvh_build.cpp: every primitive in exactly
one leaf, every child box containing the triangles below it, consistent
references and depth, merged leaves within the limit -- on SAH trees and on
median-split trees with one triangle per leaf, where the collapse merges."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_collapse_invariants(tmp_path):
    hipcc = next((c for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")) if c and os.path.exists(c)), None)
    if hipcc is None:
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "collapse_check")
    csrc = os.path.join(ROOT, "globillum_amd", "csrc")
    cmd = [hipcc, "-O2", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I" + csrc,
           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "csrc", "collapse_check.cpp"),
           os.path.join(csrc, "bvh_build.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    lines = r.stdout.splitlines()
    assert len(lines) == 42 and all(l.startswith("ok ") for l in lines)
    # the median trees (one triangle per leaf) do get merged leaves at limit 4 / 8
    med = {(l.split()[2], l.split()[3]): int(l.split()[5].split("=")[1]) for l in lines if l.split()[1] == "median"}
    assert med[("n=20000", "merge=4")] < med[("n=20000", "merge=1")]
