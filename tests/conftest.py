import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "cgmlst-hamming-gpu_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    """Import cgmlst-hamming-gpu_amd/ (a directory name that is not a Python
    identifier) as the module cgmlst_hamming_gpu_amd."""
    name = "cgmlst_hamming_gpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: CPU test of a minute or more")


@pytest.fixture(scope="session")
def cgh():
    return load_package()


@pytest.fixture(scope="session")
def workloads(cgh):
    import cgmlst_hamming_gpu_amd.workloads as w
    return w


@pytest.fixture(scope="session")
def gpu(cgh):
    """Initialise torch's HIP runtime first (the library then shares it) and
    require a device."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test needs a visible MI355X")
    torch.zeros(1, device="cuda")
    assert cgh.device_count() >= 1
    return torch
