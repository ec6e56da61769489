"""The kernels' sampler (gi_device.h: chunk tables, a fixed chunk count per
dimension, FP64 sums) evaluated on the GPU for every dimension a depth-32
render uses, against the oracle's digit loop, bit for bit
(globillum_amd/gi_halton_check runs the device function on given pairs)."""
import os
import subprocess

import numpy as np
import pytest

import globillum_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [0, 12345])
def test_device_sampler_matches_oracle(seed, oracle, tmp_path):
    exe = os.path.join(os.path.dirname(globillum_amd.__file__), "gi_halton_check")
    assert os.path.exists(exe), "gi_halton_check not built (python -m globillum_amd.build)"
    rng = np.random.default_rng(seed + 9)
    pairs = []
    for dim in range(4 + 7 * 32):
        ns = [0, 1, 2, 2 ** 31, 2 ** 32 - 1] + [int(x) for x in rng.integers(0, 2 ** 32, 150)]
        ns += [int(x) for x in rng.integers(0, 2 ** 20, 50)]
        pairs += [(n, dim) for n in ns]
    arr = np.array(pairs, np.uint32)
    fin, fout = str(tmp_path / "pairs.bin"), str(tmp_path / "vals.bin")
    arr.tofile(fin)
    r = subprocess.run([exe, str(seed), fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(fout, np.float32)
    assert len(got) == len(arr)
    want = np.array([oracle.halton_sample(int(n), int(d), seed) for n, d in arr], np.float32)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert len(bad) == 0, [(int(arr[i, 0]), int(arr[i, 1]), float(got[i]), float(want[i])) for i in bad[:5]]
