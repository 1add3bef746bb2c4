import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def product():
    """The HIP library; fails loudly if it is not built or no GPU is present."""
    from simplebubblyflowsolver_amd import solver
    be = solver.product_backend()
    be.set_dt_coefs(0.2, 0.5, 0.95)
    return be
