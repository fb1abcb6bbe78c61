import os
import sys
import warnings

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


# The synthetic workload of the parity tests: a seeded Fishbone-Moncrief torus
# (a = 0.9375, h = 0.3, Rout = 40, gamma = 13/9), M_unit 1e23 g, R_high 10,
# R_low 1, 230 GHz, inclination 60 deg, 40 M field of view.
M_UNIT = 1e23
R_HIGH = 10.0
FREQ = 2.3e11
INCL = 60.0


@pytest.fixture(scope="session")
def dump128():
    from raptor_amd import torus
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return torus.fishbone_moncrief(128, 128)


@pytest.fixture(scope="session")
def dump3d_small():
    from raptor_amd import torus
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return torus.fishbone_moncrief(48, 32, 16)


@pytest.fixture(scope="session")
def dump256():
    from raptor_amd import torus
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return torus.fishbone_moncrief(256, 256)


def rel_err(a, b, floor_frac=1e-6):
    """Per-pixel relative error with an absolute floor floor_frac*max|b|."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    floor = floor_frac * np.max(np.abs(b))
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def counters_agree(cnt, ocnt):
    """Per ray, whether the kernels' counters (steps, samples) agree with the
    oracle's (columns steps, samples in the grid, samples past the sigma cut,
    escaped, grid step, opaque step).  A ray whose optical depth passes
    ln 1000 at every frequency ends there in the kernels, since no later
    sample reaches its pixel: its steps lie after the oracle's opaque step
    and at most at its full trace's (the FP32 transfer ends 2e-3 past the
    gate, the FP64 one at it).  Other rays take the oracle's steps and
    samples."""
    steps, samples = cnt[:, 0], cnt[:, 1]
    op = ocnt[:, 5]
    full = (steps == ocnt[:, 0]) & (samples == ocnt[:, 2])
    ended = (op >= 0) & (steps > op) & (steps <= ocnt[:, 0]) & (samples <= ocnt[:, 2])
    return full | ended


# Spacetimes and step sizes beyond the headline workload: Schwarzschild, a
# near-extremal spin, the unsqueezed grid (h = 1), a negative spin (on the
# a = 0.5 torus's primitives: parity between two codes needs no consistent
# fluid), and step sizes whose RK2 points leave the angle-addition range.
SPACETIME_CASES = {
    "a0": dict(a=0.0),
    "a0.99": dict(a=0.99),
    "a0.5-h1": dict(a=0.5, hslope=1.0),
    "a-0.5": dict(a=0.5, grid_a=-0.5),
    "step0.1": dict(stepsize=0.1),
    "step0.01": dict(stepsize=0.01),
}


def spacetime_case(name, n=48):
    """(dump, stepsize) for SPACETIME_CASES[name] on an n x n grid."""
    import dataclasses
    from raptor_amd import torus
    c = dict(SPACETIME_CASES[name])
    stepsize = c.pop("stepsize", 0.03)
    grid_a = c.pop("grid_a", None)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dump = torus.fishbone_moncrief(n, n, **c)
    if grid_a is not None:
        dump = dataclasses.replace(dump, grid=dataclasses.replace(dump.grid, a=grid_a))
    return dump, stepsize
