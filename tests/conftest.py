import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def scene_c1():
    from rpba_amd import scenes
    return scenes.generate("c1_50")


@pytest.fixture(scope="session")
def tiny_scene():
    """6 cameras, 60 points: small enough for dense normal equations."""
    from rpba_amd import scenes
    return scenes.generate(None, seed=11, n_cams=6, n_pts=60, track="uniform2_8", outlier_frac=0.0)
