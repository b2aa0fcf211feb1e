"""pytest configuration: the `gpu` marker and shared fixtures.

CPU tests (-m "not gpu") cover the oracle against the golden fixtures and
closed forms, the host-side logic, and that libnekrs_hip.so loads and exports
every symbol of include/nrs_hip.h.  GPU tests (-m gpu) call the HIP library
through its C ABI and compare with the oracle.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def load(name):
        if name not in cache:
            with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
                cache[name] = {k: f[k] for k in f.files}
        return cache[name]
    return load


@pytest.fixture(scope="session")
def gpu():
    """The product package on cuda:0 (fails loudly if the library is missing)."""
    import torch
    import nekrs_amd
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    nekrs_amd.init(0)
    return nekrs_amd


@pytest.fixture(scope="session")
def torch_mod():
    import torch
    return torch
