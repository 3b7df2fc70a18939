import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def _ensure_built():
    from oracle import build as ob
    if not os.path.exists(ob.OUT):
        ob.build()
    lib = os.path.join(ROOT, "rtnn_amd", "librtnn_hip.so")
    if not os.path.exists(lib) and os.path.exists("/opt/rocm/bin/hipcc"):
        from rtnn_amd import build as rb
        rb.build()


_ensure_built()


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "samplepc_expected.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def samplepc():
    return np.loadtxt(os.path.join(GOLDEN, "samplepc.txt"), delimiter=",").astype(np.float32)


@pytest.fixture(scope="session")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rtnn_amd
    c = rtnn_amd.Context(0)
    yield c
    c.close()


def golden_d2(rows):
    return np.array([[np.inf if v is None else v for v in r] for r in rows], dtype=np.float32)
