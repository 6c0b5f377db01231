import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import tapp_cases as T  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    return T.Oracle()


@pytest.fixture(scope="session")
def ref():
    if not os.path.exists(T.REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    return T.tc.TappLib(T.REF_LIB)


@pytest.fixture(scope="session")
def lib():
    """The GPU library.  Missing library is an error, not a skip."""
    return T.tc.TappLib()


@pytest.fixture(scope="session")
def gpu_exec(lib):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.init()
    rc, ex = lib.create_executor()
    assert rc == 0
    yield ex
    lib.destroy_executor(ex)
