"""The read-bandwidth benchmark bench.py reports beside the rooflines:
a buffer held in L2 reads faster than one held in the Infinity Cache, which
reads faster than one in HBM, and the HBM figure is below the 8 TB/s
peak."""
import ctypes as C
import os

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_read_bandwidth_levels():
    lib = C.CDLL(os.path.join(ROOT, "globillum_amd", "libgi_membench.so"))
    f = lib.gi_membench_read
    f.restype = C.c_int
    f.argtypes = [C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    bw = {}
    for name, mib, passes in (("l2", 16, 3), ("mall", 128, 3), ("hbm", 2048, 2)):
        g = C.c_double(0.0)
        assert f(0, mib << 20, passes, C.byref(g)) == 0
        bw[name] = g.value
    print(bw)
    assert 1000.0 < bw["hbm"] < 8000.0
    assert bw["l2"] > bw["mall"] > bw["hbm"]
