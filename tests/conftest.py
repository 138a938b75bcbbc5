import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def tiny_system():
    from publicrelease_2020_amd import system as S
    return S.make_synthetic(3, 3, 7, nrad=24, nang=50, box=4.5)


@pytest.fixture(scope="session")
def tiny_oracle(tiny_system):
    from oracle import oracle as O
    return O.evaluate(tiny_system)


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"))
