import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
for sub in ("oracle", "hal-cpp_amd", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libhalhip.so")


# Worker threads for the oracle's line loops: this host's CPU share, at most
# 16 (results do not depend on the count).
ORACLE_THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    o = Oracle()
    o.set_threads(ORACLE_THREADS)
    return o


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "hal_reference_vectors.npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def ctx():
    """A device context.  GPU tests run on a GPU host: no device is a failure."""
    import halhip
    c = halhip.Context(0)
    yield c
    c.close()
