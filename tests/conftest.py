import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def mech():
    from nrg_amd.mechanism import load_mechanism
    return load_mechanism()


@pytest.fixture(scope="session")
def oracle(mech):
    from oracle import Oracle
    return Oracle(mech)


@pytest.fixture(scope="session")
def engine(mech):
    """The HIP chemistry step (fails loudly if libnrg_chem.so is missing or no GPU)."""
    import torch
    assert torch.cuda.is_available(), "GPU test collected on a host without a GPU"
    from nrg_amd.solver import ChemistryStep
    e = ChemistryStep(mech)
    yield e
    e.close()
