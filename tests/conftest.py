import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def ir_npz():
    return np.load(os.path.join(GOLDEN, "ir_1e3_72.npz"))


@pytest.fixture(scope="session")
def lih_npz():
    return np.load(os.path.join(GOLDEN, "lih_nk2.npz"))


@pytest.fixture(scope="session")
def lih(lih_npz, ir_npz):
    from mbsymmetry_amd.system import lih_fixture
    from mbsymmetry_amd.transforms import ir_from_fixture
    # LiH test parameters: beta = 5, ni = 72, ns = 2 (test_cugw.cpp:19-34)
    return lih_fixture(lih_npz, ir_from_fixture(ir_npz, 5.0))


@pytest.fixture(scope="session")
def config1():
    from mbsymmetry_amd.system import config1
    return config1()
