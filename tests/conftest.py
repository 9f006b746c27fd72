"""Shared fixtures.  GPU tests carry the ``gpu`` marker; everything else runs
on a machine without a GPU."""

import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    module = importlib.util.module_from_spec(spec)
    sys.modules[name] = module
    spec.loader.exec_module(module)
    return module


def load_package():
    return _load("ratpoints_gpu_amd", os.path.join(ROOT, "ratpoints-gpu_amd", "__init__.py"))


def compiled_variants():
    """(index, name) of every row-tile variant compiled into the library, read
    from its C ABI (no device needed); empty when it has not been built."""
    try:
        return list(enumerate(load_package().variants()))
    except Exception:
        return []


def load_workloads():
    return _load("ratpoints_gpu_amd_workloads",
                 os.path.join(ROOT, "ratpoints-gpu_amd", "workloads.py"))


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def workloads():
    return load_workloads()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def oracle_lib(oracle_mod):
    return oracle_mod.OracleLib()


@pytest.fixture(scope="session")
def gpu(pkg):
    """One device context shared by the GPU tests; options are reset to their
    defaults before each use by ``fresh_gpu``."""
    ctx = pkg.SieveContext(0)
    yield ctx
    ctx.close()


@pytest.fixture
def fresh_gpu(gpu):
    gpu.set_option("variant", -1)
    gpu.set_option("run_log2", -1)
    gpu.set_option("survivor_capacity", 1 << 24)
    gpu.set_option("batch_size", 65536)
    gpu.set_option("min_tiles", 2048)
    gpu.set_option("direct_words", 2048)
    gpu.set_option("primitive_only", 0)
    return gpu
