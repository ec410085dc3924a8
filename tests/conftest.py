import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "tblite-cuda_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def orc(oracle_mod):
    return oracle_mod.Oracle()


@pytest.fixture(scope="session")
def engine():
    import tblite_amd
    # fixed routing by size (no cost-model routing of small batches to the
    # multi-workgroup path), so a molecule takes the same path alone and in a
    # batch; tests/test_gpu_tuning.py covers the routing
    eng = tblite_amd.Engine(tuning={"auto_grid": 0})
    yield eng
    eng.close()
