"""The host builder's binary -> 4-wide collapse (surface-area cost, dynamic
program over slot counts), checked on the CPU by tests/csrc/collapse_check.cpp
compiled against the library's own bvh_build.cpp: every primitive in exactly
one leaf, every child box containing the triangles below it, consistent
references and depth, merged leaves within the limit, at most `width` child
slots filled (widths 2, 3, 4) -- on SAH trees and on median-split trees
with one triangle per leaf, where the collapse merges.  Also: the threaded
SAH build produces the sequential build's nodes and permutation exactly."""
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
           os.path.join(csrc, "bvh_build.cpp"), "-pthread", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    lines = r.stdout.splitlines()
    assert len(lines) == 71 and all(l.startswith("ok ") for l in lines)
    assert lines[0].startswith("ok parallel n=300000")  # threaded SAH build == sequential, bit for bit
    lines = lines[1:]
    f = {tuple(l.split()[1:5]): {kv.split("=")[0]: int(kv.split("=")[1]) for kv in l.split()[5:]} for l in lines}
    # the median trees (one triangle per leaf) do get merged leaves at limit 4
    assert f[("median", "n=20000", "merge=4", "width=4")]["leaves"] < f[("median", "n=20000", "merge=1", "width=4")]["leaves"]
    # narrower nodes: more of them, deeper trees
    for tree in ("sah", "median"):
        w2, w3, w4 = (f[(tree, "n=20000", "merge=4", f"width={w}")] for w in (2, 3, 4))
        assert w2["wide"] > w3["wide"] > w4["wide"] and w2["depth"] > w4["depth"]
