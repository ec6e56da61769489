"""bench.py's command-line contract, on the CPU: more than one GPU is
refused with a message and exit status 2 before anything touches a device,
and the help text lists the legs that can be skipped."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_help_lists_the_legs():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--no-cpu", "--no-traffic", "--no-plain", "--no-batching", "--no-adjoint",
                 "--no-bvh", "--no-instancing", "--no-small", "--no-membench", "--copy-mib", "--tuning-cache"):
        assert flag in r.stdout, flag
