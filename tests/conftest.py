import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
for _p in (ROOT, ROOT / "tests"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


@pytest.fixture
def rng():
    return np.random.default_rng(2024)


# (norb, nelec) edge cases of the reference's test matrix
# (python/ffsim/testing/testing.py:24-46, generate_norb_nelec(exhaustive=False))
NORB_NELEC_EDGE = [
    (0, (0, 0)),
    (1, (0, 0)),
    (1, (1, 0)),
    (1, (0, 1)),
    (1, (1, 1)),
    (2, (1, 1)),
    (2, (2, 2)),
    (3, (2, 1)),
    (3, (1, 2)),
    (4, (3, 2)),
]
