import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)                                  # oracle (test infrastructure)
sys.path.insert(0, os.path.join(ROOT, "pandora-2.0_amd"))  # pandora_amd (product binding)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def rel(a, b):
    """Norm-wise relative difference ||a-b||_2 / ||b||_2 (the parity criterion)."""
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle
