import pytest

from tests._util import ROOT  # noqa: F401  (puts the repo root on sys.path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def gpu_ctx():
    import kerrtracecpp2_amd as kt
    ctx = kt.Context(0)
    yield ctx
    ctx.close()
