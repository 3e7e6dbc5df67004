import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as graft  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return graft.load_package()


@pytest.fixture(scope="session")
def orc():
    o = graft.load_oracle()
    o.lib()
    return o


@pytest.fixture(scope="session")
def small_sphere(pkg):
    xyz, nrm = pkg.synth.sphere(n=3000)
    return xyz, nrm, 5


@pytest.fixture(scope="session")
def small_sphere_ref(orc, small_sphere):
    xyz, nrm, d = small_sphere
    return orc.Result(xyz, nrm, d, precision=1)


def bspline_np(u):
    """Closed form of the normalised quadratic B-spline (F(0) = 1)."""
    u = np.abs(np.asarray(u, dtype=np.float64))
    return np.where(u < 0.5, 1 - (4.0 / 3.0) * u * u, np.where(u < 1.5, (2.0 / 3.0) * (1.5 - u) ** 2, 0.0))


def func_cw(idx):
    idx = np.asarray(idx)
    depth = np.floor(np.log2(idx + 1)).astype(np.int64)
    off = idx + 1 - (1 << depth)
    w = 1.0 / (1 << depth)
    return (0.5 + off) * w, w
