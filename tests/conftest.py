import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the native kernels")


def _ensure_built():
    """Load what is built; build only if nothing is (build() always rebuilds)."""
    lib = os.path.join(ROOT, "newman_amd", "libnewman_amd.so")
    if not os.path.exists(lib):
        from newman_amd import build as b
        b.build_library()
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc):
        import oracle_lib
        oracle_lib.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU guard: a gpu-marked test must run on a HIP device and fails
    loudly (no fallback) when the engine cannot start."""
    import newman_amd as nm
    m = nm.Mandelbrot(2, 2)
    m.N = 4
    m.computeAll()  # raises NewmanError(NM_ERR_NODEVICE) without a GPU
    return nm
