import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "pushpull.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden", "pushpull_golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def golden():
    z = np.load(GOLDEN, allow_pickle=False)
    n = 1 + max(int(k[1:4]) for k in z.files if k.startswith("c"))
    cases = []
    for i in range(n):
        p = f"c{i:03d}_"
        cases.append({k[len(p):]: z[k] for k in z.files if k.startswith(p)})
    return cases


@pytest.fixture(scope="session")
def golden_affine():
    z = np.load(GOLDEN, allow_pickle=False)
    ids = sorted({k[:4] for k in z.files if k.startswith("a")})
    return [{k[5:]: z[k] for k in z.files if k.startswith(i + "_")} for i in ids]
