import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "traversal_small.npz")))


@pytest.fixture(scope="session")
def small_scene():
    from alpine_amd import scenes
    return scenes.small_scene()


@pytest.fixture(scope="session")
def microbench_scene():
    from alpine_amd import scenes
    return scenes.microbench_scene()
