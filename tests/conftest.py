import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle32():
    from oracle.oracle import Oracle
    import numpy as np
    return Oracle(np.float32)


@pytest.fixture(scope="session")
def oracle64():
    from oracle.oracle import Oracle
    import numpy as np
    return Oracle(np.float64)


@pytest.fixture(scope="session")
def oracle32_fma():
    from oracle.oracle import Oracle
    import numpy as np
    return Oracle(np.float32, fma=True)
