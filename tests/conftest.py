import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hpgl-reborn_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ref(orc):
    if not orc.ref_available():
        pytest.skip("reference checker oracle/_ref/libhpgl_ref.so not built")
    return orc


@pytest.fixture(scope="session")
def golden_7_2():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "sk_problem_7_2.json")) as f:
        return json.load(f)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
