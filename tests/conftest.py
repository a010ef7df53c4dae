import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU check")


import numpy as _np
import pytest as _pytest


@_pytest.fixture(scope="session")
def golden():
    d = dict(_np.load(os.path.join(ROOT, "tests", "golden", "golden_fft.npz")))
    d["sizes"] = {int(k[2:]) for k in d if k.startswith("x_")}
    return d
