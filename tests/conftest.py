"""Shared test setup: markers, the product binding (parts-sim_amd) and the oracle bindings.

GPU tests are marked ``gpu``; everything else runs on the CPU. The oracle (oracle/) is the checker
only: tests compare the product's results against it.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
# chunked reductions of the oracle and of the compiled reference follow this thread count
os.environ.setdefault("PARTS_SIM_THREADS", "8")


def _load_package():
    if "parts_sim_amd" in sys.modules:
        return sys.modules["parts_sim_amd"]
    pkg = os.path.join(ROOT, "parts-sim_amd")
    spec = importlib.util.spec_from_file_location("parts_sim_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["parts_sim_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


psx = _load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def P():
    return psx


@pytest.fixture(scope="session")
def W():
    import parts_sim_amd.workloads as w
    return w


@pytest.fixture(scope="session")
def cpu():
    import cpu as c
    return c


def ref_available():
    import cpu as c
    return c.available("ref")
