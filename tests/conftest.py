import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on device)")
    config.addinivalue_line("markers", "slow: CPU test taking more than a few seconds")


@pytest.fixture(scope="session")
def cube():
    import oracle_api
    return oracle_api.refined_cube()


@pytest.fixture(scope="session")
def cube_charges(cube):
    import oracle_api
    return oracle_api.glibc_rand_charges(len(cube))


@pytest.fixture(scope="session")
def cube_exact(cube, cube_charges):
    """Direct sum at every point of the refined cube, kappa = 1."""
    import oracle_api
    return oracle_api.direct(cube, cube_charges, cube, 1.0)
