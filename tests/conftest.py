import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP kernels)")


def pytest_sessionstart(session):
    # torch ships its own HIP runtime, which finds the GPU only if it
    # initialises before libptamd's: GPU sessions bring it up first
    expr = session.config.getoption("markexpr") or ""
    if "gpu" in expr and "not gpu" not in expr:
        try:
            import torch
            torch.cuda.init()
        except Exception:
            pass


def pytest_collection_modifyitems(config, items):
    # the session-scoped GPU fixtures are shared; keep file order stable
    pass


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)


def golden_npz(name):
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", name), allow_pickle=False)
