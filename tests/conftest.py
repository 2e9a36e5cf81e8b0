"""Shared fixtures: the package (directory name has a hyphen, so it is loaded by path) and the oracle."""
import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_package():
    name = "blackframe_renderer_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, ROOT / "blackframe-renderer_amd" / "__init__.py",
        submodule_search_locations=[str(ROOT / "blackframe-renderer_amd")])
    module = importlib.util.module_from_spec(spec)
    sys.modules[name] = module
    spec.loader.exec_module(module)
    return module


def load_oracle_module():
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # noqa: E402

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def bf():
    return load_package()


@pytest.fixture(scope="session")
def abi(bf):
    return bf.abi


@pytest.fixture(scope="session")
def oracle(bf):
    return load_oracle_module().Oracle(bf.abi)


@pytest.fixture(scope="session")
def oracle_module():
    return load_oracle_module()


@pytest.fixture(scope="session")
def hip(bf):
    import importlib

    return importlib.import_module("blackframe_renderer_amd.hip")
