"""Shared fixtures: the product package (loaded from vespa.jl_amd/), the CPU
oracle (the checker) and the `gpu` marker."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def load_vespa():
    """Import vespa.jl_amd/ (a directory name with a dot) as `vespa_jl_amd`."""
    if "vespa_jl_amd" in sys.modules:
        return sys.modules["vespa_jl_amd"]
    d = os.path.join(ROOT, "vespa.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "vespa_jl_amd", os.path.join(d, "__init__.py"), submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vespa_jl_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); parity tests proper")


def pytest_collection_modifyitems(config, items):
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    V = load_vespa()
    if V.device_count() == 0:
        skip = pytest.mark.skip(reason="no HIP device visible")
        for it in gpu_items:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def V():
    return load_vespa()


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle
