import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def sfh():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def emu():
    path = os.path.join(ROOT, "tests", "emu", "libpsf_emu.so")
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    import oracle
    dp = ctypes.POINTER(ctypes.c_double)
    L.emu_bin_cmd_smooth.argtypes = [dp, dp, dp, dp, dp, dp, ctypes.c_int64, ctypes.c_double,
                                     oracle.Range, oracle.Range, dp]
    L.emu_bin_cmd_smooth.restype = ctypes.c_int64
    return L


@pytest.fixture(scope="session")
def gpu_ctx(sfh):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU (run with -m 'not gpu' on CPU)"
    return sfh.default_context()
