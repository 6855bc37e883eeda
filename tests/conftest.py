import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP kernels")
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "pynear_amd", "libpynear_amd.so")):  # fresh checkout
        subprocess.run([sys.executable, "-m", "pynear_amd.build"], cwd=ROOT, check=True)


def _have_gpu():
    from pynear_amd import _native
    try:
        _native.lib()
    except _native.NativeLibraryError:
        if os.path.exists("/dev/kfd"):  # a GPU machine without the HIP library: fail, never skip
            raise
        return False
    try:
        return _native.device_count() > 0
    except RuntimeError:
        return False


HAVE_GPU = _have_gpu()


def pytest_collection_modifyitems(config, items):
    if HAVE_GPU:
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def ref_module():
    import oracle
    mod = oracle.load_ref()
    if mod is None:
        pytest.skip("reference module not built (make -C oracle ref)")
    return mod
