import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


def _ensure_built():
    lib = os.path.join(ROOT, "sph2s_final_amd", "libsph2s.so")
    if not os.path.exists(lib):
        from sph2s_final_amd import build
        build.build_library()
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def has_gpu():
    import sph2s_final_amd
    return sph2s_final_amd.device_count() > 0


@pytest.fixture(autouse=True)
def _require_gpu(request):
    if request.node.get_closest_marker("gpu") is not None:
        import sph2s_final_amd
        if sph2s_final_amd.device_count() == 0:
            pytest.fail("GPU test run without a visible HIP device")
