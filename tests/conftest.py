import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "sai-qmc_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    if not os.path.exists(pyoracle.ORACLE_SO):
        pytest.fail("oracle/liboracle.so missing: run __graft_entry__.build()")
    return pyoracle


@pytest.fixture(scope="session")
def eng_mod():
    import dqmc_gpu
    dqmc_gpu.lib()  # raises if the HIP engine is missing
    return dqmc_gpu
