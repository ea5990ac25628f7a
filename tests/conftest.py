import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def born_scene():
    """Config-1 scene (2,562 patches) with its oracle."""
    import oracle
    from beep3_amd import configs

    from beep3_amd.scene import Scene

    b = configs.born_ion()
    sc = Scene(b.Dsolvent, b.kappa, b.quad, b.qual, b.nbsize)
    sc.add_instance(b.library[0], 2.0)
    sc.build()
    a = sc.arrays()
    return sc, a, oracle.Oracle(a, b.kappa, b.Dsolvent)
