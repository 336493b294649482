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


@pytest.fixture(scope="session")
def exact():
    """exact(kind, Lx, Ly, U, dtau, L, field, tau=0, model="engine"): the
    Green's function of a field in 80-bit extended precision
    (oracle/exact.c), built on the FP64 propagators of the engine's host
    layer (model="engine", dqmc_gpu.model) or of the oracle (model="oracle",
    pyoracle.model): each FP64 path is measured against the exact G of its
    own inputs."""
    import dqmc_gpu
    import pyoracle

    def f(kind, Lx, Ly, U, dtau, L, field, tau=0, model="engine", blk=8):
        th, _, _ = dqmc_gpu.lattice(kind, Lx, Ly)
        if model == "engine":
            eK, _, _, lam = dqmc_gpu.model(th, U, dtau)
        else:
            eK, _, lam = pyoracle.model(th, U, dtau)
        return pyoracle.green_exact(eK, lam, U, dtau, field, L, tau, blk)

    return f
