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
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP kernels through the C ABI")


@pytest.fixture(scope="session")
def refvals():
    with open(os.path.join(GOLDEN, "reference_values.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def itq29():
    with open(os.path.join(GOLDEN, "itq29.json")) as f:
        d = json.load(f)
    a = d["a"]
    frac = np.array(d["fractional"])
    return dict(a=a, pos=frac * a, element=d["element"], charge=np.array(d["charge"]))


@pytest.fixture(scope="session")
def mfi():
    with open(os.path.join(GOLDEN, "mfi.json")) as f:
        d = json.load(f)
    abc = np.array([d["a"], d["b"], d["c"]])
    frac = np.array(d["fractional"])
    return dict(abc=abc, frac=frac, element=d["element"], charge=np.array(d["charge"]))


@pytest.fixture(scope="session")
def gpu_ctx():
    import raspa3_amd as ra
    ctx = ra.Context(0)
    yield ctx
    ctx.close()
