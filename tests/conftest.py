import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "metal-vibe-tracer_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


def _ensure_built():
    lib = os.path.join(PKG, "libvt.so")
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not (os.path.exists(lib) and os.path.exists(orc)):
        import __graft_entry__
        __graft_entry__.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    import vt
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    torch.cuda.init()
    ctx = vt.Context(0)
    yield ctx
    ctx.close()
