import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library's C ABI)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build(force=False)
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def zh():
    """libzmlhip.so through its C ABI (fails loudly when it is not built)."""
    from zml_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    torch.cuda.init()
    return torch


@pytest.fixture(params=["by_height", "one_launch"])
def lu_panel(request, zh):
    """Runs an FP64 LU test under both panel strategies: the default choice by
    block-column height (leaf by leaf up to 16384 rows, one launch above), and
    every block column in one launch. Both must give zml's pivots."""
    zh.zmlhip_set_panel_one_launch_rows(0 if request.param == "one_launch" else -1)
    yield request.param
    zh.zmlhip_set_panel_one_launch_rows(-1)
