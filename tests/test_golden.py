"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the hierarchy the setup produces (P, transpose map, triple map, coarse
pattern and values), the oracle's V-cycle, the dense restatement's V-cycle and
the first five PCG iterates, for 1-D / 2-D Poisson with aggregation and a
3-D heterogeneous 8^3 hypre-default hierarchy.  The CPU tests hold the setup
and the oracle to the fixtures bit for bit; the GPU test holds the device
cycle and PCG to them."""
import glob
import os

import numpy as np
import pytest

from conftest import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
NAMES = [os.path.basename(f)[:-4] for f in FIXTURES]


def _load(name):
    return dict(np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False))


def _case(name):
    import make_golden
    return make_golden.cases()[name]


@pytest.fixture(scope="module", autouse=True)
def _path():
    import sys
    p = os.path.join(HERE, "golden")
    if p not in sys.path:
        sys.path.insert(0, p)


def test_fixtures_present():
    assert len(FIXTURES) == 3


@pytest.mark.parametrize("name", NAMES)
def test_setup_and_oracle_match_fixture(name):
    import make_golden
    A, cfg, sm = _case(name)
    ref = _load(name)
    got = make_golden.build(name, A, cfg, sm)
    assert set(got) == set(ref)
    for k in ref:
        if k == "cycle_dense":
            continue
        assert np.array_equal(got[k], ref[k]), (name, k)
    # the independent dense restatement agrees with the oracle's cycle
    assert rel_err(ref["cycle_dense"], ref["cycle_oracle"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_device_cycle_and_pcg_match_fixture(dh, torch_cuda, name):
    torch = torch_cuda
    A, cfg, sm = _case(name)
    ref = _load(name)
    h = dh.amg_setup(A, cfg)
    assert h.nlevels == int(ref["nlevels"][0])
    x = torch.tensor(ref["x0"], device="cuda")
    dh.amg_cycle_(x, torch.tensor(ref["b"], device="cuda"), h, cfg)
    assert np.array_equal(x.cpu().numpy(), ref["cycle_oracle"])
    b = torch.tensor(ref["b"], device="cuda")
    for k in range(1, 6):
        xk = torch.zeros_like(b)
        dh.pcg_solve_(xk, b, h, cfg, tol=0.0, maxiter=k)
        assert rel_err(xk.cpu().numpy(), ref["pcg_iterates"][k - 1]) < 1e-12, (name, k)
