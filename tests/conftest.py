import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (str(ROOT), str(ROOT / "tests")):
  if p not in sys.path:
    sys.path.insert(0, p)


def pytest_configure(config):
  config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libraries")


def gpu_available() -> bool:
  try:
    from qtqp_amd import _native
    lib = _native.kkt_lib()
    return lib.qtqp_device_count() > 0
  except OSError:
    return False


@pytest.fixture(scope="session")
def require_gpu():
  if not gpu_available():
    pytest.fail("GPU test run without a usable HIP device / built library")
