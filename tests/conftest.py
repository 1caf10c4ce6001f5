import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_fixtures.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(ROOT, "tests", "golden")


def load_pkg():
    """Import the package directory (its name is not a Python identifier)."""
    import importlib.util
    if "hsps_amd" in sys.modules:
        return sys.modules["hsps_amd"]
    path = os.path.join(ROOT, "hybrid-sparse-pde-solver_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("hsps_amd", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hsps_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def H():
    return load_pkg()


@pytest.fixture(scope="session")
def ctx(H):
    c = H.Context(0)
    yield c
    c.close()
