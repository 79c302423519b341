import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "damascus-earth-evap_amd"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def oracle_solar():
    import oracle_lib
    return oracle_lib.Solar()


@pytest.fixture(scope="session")
def gpu_ctx():
    import dmsun
    return dmsun.Context(device=0)
