import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    """New Context per call on device 0 (fails loudly without a GPU)."""
    from globillum_amd.context import Context

    def make(scene=None, **kw):
        ctx = Context(0)
        if scene is not None:
            ctx.upload(scene, **kw)
        return ctx

    return make
