import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def tables():
    from uclchem_amd import makerates_hip as mh
    return mh.load_network(mh.network_path("default"))


@pytest.fixture(scope="session")
def oracle(tables):
    from oracle import Oracle
    return Oracle(tables)


@pytest.fixture(scope="session")
def reflib():
    from oracle import REF_SO, RefLib
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libucref.so not built (needs the reference sources)")
    return RefLib()


@pytest.fixture(scope="session")
def gpu_net(tables):
    from uclchem_amd import gpu
    return gpu.Network(tables)


@pytest.fixture(scope="session")
def oracle_of():
    """Oracle (C restatement) bound to a given network's tables, one per network."""
    from oracle import Oracle
    cache = {}

    def get(t):
        if t.name not in cache:
            cache[t.name] = Oracle(t)
        return cache[t.name]
    return get
