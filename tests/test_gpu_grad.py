"""Gradients of the image on the GPU (tangent pass) against central
differences, on a 3D dump with azimuthal structure.

The differences are taken of FP64 oracle images (step 1e-3): differences of
GPU images would be dominated by the FP32 transfer's ~1e-6 relative rounding,
which a 1e-3 step amplifies to ~1e-3."""
import math
import warnings

import numpy as np
import pytest

from conftest import FREQ, INCL, M_UNIT, R_HIGH

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dump3d():
    from raptor_amd import torus
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return torus.fishbone_moncrief(64, 48, 24)


@pytest.mark.parametrize("R_high", [1.0, R_HIGH, 40.0, 160.0])
def test_gpu_gradient_against_oracle_differences_step_1e3(dump3d, R_high):
    """GPU tangents against central differences (step 1e-3) of the FP64
    oracle, for R_high across [1, 160]; pixels whose difference quotient
    changes between steps h and h/2 (a gate flips inside the step) are
    excluded."""
    from oracle import oracle as O
    from raptor_amd import rt
    W = 24
    with rt.Tracer(0) as tr:
        tr.load_dump(dump3d)
        tr.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_high)
        img, dimg, _ = tr.render([INCL], [FREQ], grad=True)
    kw = dict(width=W, height=W, inclination=INCL)

    def ora(M, R):
        return O.render(O.params(dump3d.grid, M_unit=M, R_high=R, **kw), dump3d.prims,
                        [FREQ])[:, 0]

    h = 1e-3
    R0 = R_high
    for q in range(2):
        if q == 0:
            f1 = (ora(M_UNIT * math.exp(h), R0) - ora(M_UNIT * math.exp(-h), R0)) / (2 * h)
            f2 = (ora(M_UNIT * math.exp(h / 2), R0) - ora(M_UNIT * math.exp(-h / 2), R0)) / h
        else:
            f1 = (ora(M_UNIT, R0 * (1 + h)) - ora(M_UNIT, R0 * (1 - h))) / (2 * h * R0)
            f2 = (ora(M_UNIT, R0 * (1 + h / 2)) - ora(M_UNIT, R0 * (1 - h / 2))) / (h * R0)
        an = dimg[0, :, 0, q]
        den = np.maximum(np.abs(an), 1e-3 * np.abs(an).max())
        smooth = np.abs(f1 - f2) / den < 1e-3
        err = np.abs(f1 - an) / den
        assert smooth.mean() > 0.9
        assert np.mean(err[smooth] < 1e-3) > 0.98 and err[smooth].max() < 2e-2
        assert f1[smooth].sum() / an[smooth].sum() == pytest.approx(1.0, abs=5e-4)
        # every smooth pixel to 1e-3 of the larger of its tangent and I / p
        # (a unit logarithmic derivative d ln I / d ln p): relative where the
        # tangent is large, as a logarithmic derivative where it is near zero
        p = 1.0 if q == 0 else R0
        ipix = np.maximum(img[0, :, 0], 1e-6 * img[0, :, 0].max())
        e2 = np.abs(f1 - an) / np.maximum(np.abs(an), ipix / p)
        print("R_high %g d%d: max error %.3g (tangent or I/p scale) over %d smooth pixels"
              % (R0, q, e2[smooth].max(), smooth.sum()))
        assert e2[smooth].max() < 1e-3, q
