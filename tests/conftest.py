"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "draugr.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def dh():
    import draugr_hip
    return draugr_hip


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch


def hierarchy_levels(h):
    """Host arrays of every level plus the coarsest operator of a hierarchy."""
    return [h.level(l) for l in range(h.nlevels)], h.coarsest()


def oracle_of(h, orc, smoother, cycle="v", pre=1, post=1, omega=2.0 / 3.0, gs_symmetric=False):
    levels, c = hierarchy_levels(h)
    return orc.OracleHierarchy(levels, c["rp"], c["cv"], c["nz"], smoother=smoother, omega=omega, cycle=cycle,
                               pre=pre, post=post, gs_symmetric=gs_symmetric)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
