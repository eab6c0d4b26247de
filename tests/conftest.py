import os
import sys

import pytest

try:  # import torch at collection time: the first import on a fresh machine can take a minute
    import torch  # noqa: F401
except Exception:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def vctx():
    import torch
    import vren_amd
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    ctx = vren_amd.Context(0)
    yield ctx
    torch.cuda.synchronize()
    ctx.close()
