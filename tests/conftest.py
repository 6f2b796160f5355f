import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels through libthcpp_amd.so on an MI355X")


@pytest.fixture(scope="session")
def ctx():
    import torch
    from thcpp_amd.api import Context
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_graph():
    from thcpp_amd.api import Context
    c = Context(0, graphs=True)
    yield c
    c.close()


@pytest.fixture(scope="session")
def cfg0():
    from thcpp_amd import synthetic
    return synthetic.make_config("diamond8_n32")


@pytest.fixture(scope="session")
def cfg0_kmeans(cfg0):
    from oracle import kmeans_oracle as KO
    s = cfg0
    return KO.kmeans(s.grid, s.rho, s.K, s.max_it, -1.0, 0, keep_maps=True)
