"""Test configuration: registers the `gpu` marker, loads the package and the oracle."""
import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(REPO, "real-time-depth-fusion-directional-tsdf_amd")
ORACLE_DIR = os.path.join(REPO, "oracle")


def _load_pkg():
    if "vsg_amd" in sys.modules:
        return sys.modules["vsg_amd"]
    spec = importlib.util.spec_from_file_location("vsg_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vsg_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_pkg()
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch
