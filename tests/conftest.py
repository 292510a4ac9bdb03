import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "dugks-gpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: long-running")


# D3Q19 tables written out for the tests (velocity order of the reference,
# src/lbm_d3q19.cu:10-21); their closed-form moments are checked in
# test_oracle.py before anything relies on them.
C = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
              [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
              [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]]).T  # (19, 3)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = np.array([int(np.where((C == -C[q]).all(1))[0][0]) for q in range(19)])
RT = 1.0 / 3.0


def feq(rho, u):
    """Second-order Maxwellian on D3Q19; rho (...), u (..., 3) -> (..., 19)."""
    cu = np.einsum("qa,...a->...q", C, u) / RT
    uu = (u ** 2).sum(-1)[..., None] / RT
    return W * rho[..., None] * (1 + cu + 0.5 * (cu * cu - uu))


def moments(f):
    rho = f.sum(-1)
    j = np.einsum("...q,qa->...a", f, C)
    return rho, j


def random_state(shape, rng, amp=0.05, drho=0.02):
    """Positive near-equilibrium f (nz, ny, nx, 19) with random moments and
    a random non-equilibrium part."""
    nx, ny, nz = shape
    rho = 1 + drho * rng.standard_normal((nz, ny, nx))
    u = amp * rng.standard_normal((nz, ny, nx, 3))
    f = feq(rho, u)
    f *= 1 + 0.02 * rng.standard_normal(f.shape)
    return f


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def reflect_state(f, axis):
    """Mirror x_axis -> L - x_axis of a state f (nz, ny, nx, 19): the cells
    reverse along the axis and each population takes the slot of the one
    whose velocity has that component negated."""
    cn = C.copy()
    cn[:, axis] *= -1
    q_new = [int(np.where((C == cn[q]).all(1))[0][0]) for q in range(19)]
    sp = np.flip(f, axis=2 - axis)
    out = np.empty_like(sp)
    out[..., q_new] = sp
    return out
