"""Shared fixtures: the package loader, the oracle and the HIP library."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def load_pkg():
    """Import the `msbg-rs_amd/` package directory as `msbg_rs_amd`."""
    if "msbg_rs_amd" in sys.modules:
        return sys.modules["msbg_rs_amd"]
    path = os.path.join(ROOT, "msbg-rs_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("msbg_rs_amd", path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["msbg_rs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_pkg()  # test modules import `msbg_rs_amd` submodules at collection time


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def orc():
    import oracle_ffi
    return oracle_ffi.load()


@pytest.fixture(scope="session")
def hip(pkg):
    """The HIP product library; fails loudly if it is not built or no GPU."""
    lib = pkg.lib.load()
    n = lib.device_count()
    if n < 1:
        pytest.fail("no HIP device visible: the gpu tests need an MI355X")
    return lib
