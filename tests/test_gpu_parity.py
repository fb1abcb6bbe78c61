"""Parity of the HIP path (through the C-ABI) with the FP64 CPU oracle.

Bar: total flux within 1e-4 relative, every pixel within 1e-3
relative with an absolute floor of 1e-6 * max(I); building blocks to
near-FP64 rounding.
"""
import ctypes

import numpy as np
import pytest

from conftest import (FREQ, INCL, M_UNIT, R_HIGH, SPACETIME_CASES, counters_agree, rel_err,
                      spacetime_case)
from oracle import oracle as O
from raptor_amd import rt

pytestmark = pytest.mark.gpu

FLUX_TOL = 1e-4
PIX_TOL = 1e-3
# distance of the FP64 tau from ln 1000 at the gate crossing below which an
# FP32 tau (error up to ~3e-4 there) may decide the gate differently.  The
# thermal fit's rays at the gate are traced again with the FP64 transfer, so
# its images meet the bar on every pixel; the other emission models (no FP64
# form) still hold these pixels apart.
GATE_EPS = 1e-4


@pytest.fixture(scope="module")
def tracer():
    tr = rt.Tracer(0)
    yield tr
    tr.close()


def _setup(tr, dump, W, H=None, **kw):
    H = W if H is None else H
    tr.set_params(width=W, height=H, M_unit=kw.get("M_unit", M_UNIT),
                  R_high=kw.get("R_high", R_HIGH), R_low=kw.get("R_low", 1.0))
    tr.load_dump(dump)
    return O.params(dump.grid, width=W, height=H, inclination=kw.get("incl", INCL),
                    M_unit=kw.get("M_unit", M_UNIT), R_high=kw.get("R_high", R_HIGH),
                    R_low=kw.get("R_low", 1.0))


def _report(name, gpu, ora):
    r = rel_err(gpu, ora)
    worst = np.argsort(r)[-5:][::-1]
    print("%s: flux rel %.3e, max pix rel %.3e, worst %s" %
          (name, gpu.sum() / ora.sum() - 1, r.max(), [(int(i), float(r[i])) for i in worst]))
    return r


def test_c0_image_matches_oracle(tracer, dump128):
    """64x64 image of the 128^2 torus at 230 GHz, i = 60 deg (the CPU config)."""
    P = _setup(tracer, dump128, 64)
    img, spec = tracer.render([INCL], [FREQ])
    ora = O.render(P, dump128.prims, [FREQ])[:, 0]
    gpu = img[0, :, 0]
    r = _report("c0", gpu, ora)
    assert abs(gpu.sum() / ora.sum() - 1) < FLUX_TOL
    assert r.max() < PIX_TOL
    assert spec[0, 0] == pytest.approx(gpu.sum(), rel=1e-14)


@pytest.mark.parametrize("nf", [1, 3, 8])
def test_cold_cells_poison_rays_as_the_reference(tracer, dump128, nf):
    """Cells with u < 0 give Theta_e = 1e-20 (raptor_harm_model.c:274-276):
    j = 0 there, and the reference's FP64 Planck exponent overflows, so its
    absorptivity is 0/0 and tau NaN, which stops the ray's later
    contributions (the gate test fails from then on).  A shell of such cells
    inside the torus (r ~ 6.0-9.2 M, X2 < 3/4): images and tangents at 1,
    3 and 8 frequencies against the oracle, every pixel; the shell changes
    tens of pixels by more than 1e-3 of the image's largest (the flux by
    1-12%), and no output is NaN."""
    import copy
    cold = copy.copy(dump128)
    cold.prims = dump128.prims.copy()
    cold.prims[1, 56:72, :96, :] = -np.abs(cold.prims[1, 56:72, :96, :])
    W = 32
    freqs = [FREQ] if nf == 1 else list(86e9 * (690.0 / 86.0) ** (np.arange(nf) / (nf - 1)))
    P = _setup(tracer, cold, W)
    try:
        img, dimg, _ = tracer.render([INCL], freqs, grad=True)
    finally:
        tracer.load_dump(dump128)
    oimg, odimg = O.render(P, cold.prims, freqs, grad=True)
    warm = O.render(P, dump128.prims, freqs)
    assert np.all(np.isfinite(img)) and np.all(np.isfinite(dimg))
    for f in range(nf):
        gpu, ora = img[0, :, f], oimg[:, f]
        r = rel_err(gpu, ora)
        assert abs(gpu.sum() / ora.sum() - 1) < FLUX_TOL, f
        assert r.max() < PIX_TOL, (f, r.max())
        changed = np.abs(ora - warm[:, f]) > 1e-3 * np.abs(warm[:, f]).max()
        assert changed.sum() >= 20, f
        ipix = np.maximum(ora, 1e-6 * ora.max())
        for q, p in enumerate((1.0, R_HIGH)):
            g, o = dimg[0, :, f, q], odimg[:, f, q]
            assert (np.abs(g - o) * p / ipix).max() < 1e-3, (f, q)


@pytest.mark.parametrize("W,H,incl", [(37, 23, 60.0), (1, 1, 30.0), (16, 16, 1.0),
                                      (16, 16, 179.0), (24, 8, 90.0)])
def test_ragged_sizes_and_inclinations(tracer, dump128, W, H, incl):
    P = _setup(tracer, dump128, W, H, incl=incl)
    img, _ = tracer.render([incl], [FREQ])
    ora = O.render(P, dump128.prims, [FREQ])[:, 0]
    if ora.max() == 0:
        assert np.all(img == 0)
        return
    r = _report("%dx%d i=%g" % (W, H, incl), img[0, :, 0], ora)
    assert r.max() < PIX_TOL


def test_maximum_batch_cameras_and_frequencies(tracer, dump128):
    """The largest batch one call takes: 64 inclinations x 16 frequencies
    (10 to 2000 GHz) of a small image, every camera and frequency against the
    oracle's, with and without the approach kernel; one more camera or
    frequency is refused."""
    W = 8
    P = _setup(tracer, dump128, W)
    incl = list(np.linspace(2.0, 178.0, 64))
    freqs = list(1e10 * 200.0 ** (np.arange(16) / 15.0))
    img, spec = tracer.render(incl, freqs)
    assert img.shape == (64, W * W, 16)
    tracer.set_launch_config(approach=100)
    try:
        img2, _ = tracer.render(incl, freqs)
    finally:
        tracer.set_launch_config()
    assert np.array_equal(img, img2)
    worst = 0.0
    for c in range(0, 64, 7):
        Pc = O.params(dump128.grid, width=W, height=W, inclination=incl[c], M_unit=M_UNIT,
                      R_high=R_HIGH)
        ora, gate = O.render(Pc, dump128.prims, freqs, gate=True)
        for f in range(16):
            if ora[:, f].max() <= 0:
                assert np.all(img[c, :, f] == 0)
                continue
            r = rel_err(img[c, :, f], ora[:, f])
            worst = max(worst, r.max())
            assert spec[c, f] == pytest.approx(img[c, :, f].sum(), rel=1e-13)
    assert worst < PIX_TOL
    with pytest.raises(rt.RtError, match="n_cam"):
        tracer.render(incl + [60.0], [FREQ])
    with pytest.raises(rt.RtError, match="n_freq"):
        tracer.render([60.0], freqs + [3e12])


def test_ray_counters_match_oracle(tracer, dump128):
    W = 32
    P = _setup(tracer, dump128, W)
    n = W * W
    d_img = tracer.alloc(n * 8)
    d_cnt = tracer.alloc(n * 8)
    try:
        tracer.render_device([INCL], [FREQ], d_img, None, d_cnt)
        tracer.synchronize()
        cnt = tracer.d2h(d_cnt, (n, 2), np.int32)
    finally:
        tracer.free(d_img)
        tracer.free(d_cnt)
    _, oc = O.render(P, dump128.prims, [FREQ], counters=True)
    assert counters_agree(cnt, oc).mean() > 0.99


def test_gradients_match_oracle(tracer, dump128):
    P = _setup(tracer, dump128, 32)
    img, dimg, _ = tracer.render([INCL], [FREQ], grad=True)
    oimg, odimg = O.render(P, dump128.prims, [FREQ], grad=True)
    assert rel_err(img[0, :, 0], oimg[:, 0]).max() < PIX_TOL
    ipix = np.maximum(oimg[:, 0], 1e-6 * oimg[:, 0].max())
    for q, p in enumerate((1.0, R_HIGH)):   # d/d ln M_unit, R_high d/d R_high
        g, o = dimg[0, :, 0, q], odimg[:, 0, q]
        assert abs(g.sum() / o.sum() - 1) < 1e-4
        # FP32 transfer: per-pixel gradients are sums of terms of both signs,
        # so a tangent near zero has a large relative error; every pixel's
        # logarithmic derivative d ln I / d ln p is held to 1e-3
        r = rel_err(g, o, 1e-5)
        assert np.mean(r < 1e-3) > 0.99 and r.max() < 1e-2
        assert (np.abs(g - o) * p / ipix).max() < 1e-3, q


@pytest.mark.parametrize("nf", [3, 8])
def test_gradients_several_frequencies_match_oracle(tracer, dump128, nf):
    """The tangent pass of the 4- and 8-frequency kernels against the
    oracle's tangents, every frequency, every pixel (rays at the ln 1000 gate
    take the FP64 transfer's tangents).  At 86 GHz the torus is thick enough
    that some pixels hardly depend on M_unit: their tangent, a sum of terms
    of both signs, is ~1e-5 of the image's largest, and its relative error
    says little; every pixel is held to 1e-3 in the logarithmic derivative
    d ln I / d ln p instead, 99% to 1e-3 relative, and the sums to 1e-4."""
    P = _setup(tracer, dump128, 32)
    freqs = list(86e9 * (690.0 / 86.0) ** (np.arange(nf) / (nf - 1)))
    img, dimg, _ = tracer.render([INCL], freqs, grad=True)
    oimg, odimg = O.render(P, dump128.prims, freqs, grad=True)
    for f in range(nf):
        assert rel_err(img[0, :, f], oimg[:, f]).max() < PIX_TOL, f
        ipix = np.maximum(oimg[:, f], 1e-6 * oimg[:, f].max())
        for q, p in enumerate((1.0, R_HIGH)):   # d/d ln M_unit, R_high d/d R_high
            g, o = dimg[0, :, f, q], odimg[:, f, q]
            assert abs(g.sum() / o.sum() - 1) < 1e-4, (f, q)
            r = rel_err(g, o, 1e-5)
            assert np.mean(r < 1e-3) > 0.99, (f, q)
            assert (np.abs(g - o) * p / ipix).max() < 1e-3, (f, q)


def test_multi_frequency_and_camera_batch(tracer, dump128):
    """One launch over 3 inclinations x 3 frequencies equals the per-image launches."""
    _setup(tracer, dump128, 24)
    incls = [20.0, 60.0, 85.0]
    freqs = [8.6e10, 2.3e11, 6.9e11]
    img, spec = tracer.render(incls, freqs)
    for c, inc in enumerate(incls):
        for f, nu in enumerate(freqs):
            one, _ = tracer.render([inc], [nu])
            np.testing.assert_allclose(img[c, :, f], one[0, :, 0], rtol=1e-12, atol=0)
    P = O.params(dump128.grid, width=24, height=24, inclination=85.0, M_unit=M_UNIT,
                 R_high=R_HIGH)
    ora = O.render(P, dump128.prims, freqs)
    for f in range(3):
        assert rel_err(img[2, :, f], ora[:, f]).max() < PIX_TOL


def test_repeat_runs_are_bitwise_identical(tracer, dump128):
    _setup(tracer, dump128, 48)
    a, _ = tracer.render([INCL], [FREQ])
    b, _ = tracer.render([INCL], [FREQ])
    assert np.array_equal(a, b)


def test_3d_phi_independent_dump_equals_2d(tracer, dump128):
    """A 3D dump whose phi slices are identical renders the 2D image."""
    from raptor_amd.torus import Dump
    import dataclasses
    _setup(tracer, dump128, 32)
    img2, _ = tracer.render([INCL], [FREQ])
    g3 = dataclasses.replace(dump128.grid, n3=6, dx3=2 * np.pi / 6)
    d3 = Dump(grid=g3, prims=np.repeat(dump128.prims, 6, axis=3))
    tracer.load_dump(d3)
    img3, _ = tracer.render([INCL], [FREQ])
    # identical up to the FP32 rounding of (1 - d3) w + d3 w in the sampler
    assert rel_err(img3[0, :, 0], img2[0, :, 0]).max() < 1e-5


def test_3d_dump_matches_oracle(tracer):
    from raptor_amd import torus
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        d3 = torus.fishbone_moncrief(48, 32, 16)
    P = _setup(tracer, d3, 32)
    img, _ = tracer.render([INCL], [FREQ])
    ora = O.render(P, d3.prims, [FREQ])[:, 0]
    r = _report("3d", img[0, :, 0], ora)
    assert abs(img[0, :, 0].sum() / ora.sum() - 1) < FLUX_TOL
    assert r.max() < PIX_TOL


@pytest.mark.parametrize("nf", [3, 8])
def test_3d_dump_several_frequencies_match_oracle(tracer, dump3d_small, nf):
    """The trilinear sampler's multi-frequency kernels (4 and 8 frequencies)
    on the 3D dump, two inclinations (one nearly edge-on, whose rays cross
    many phi cells), every frequency against the oracle."""
    W = 24
    freqs = list(86e9 * (690.0 / 86.0) ** (np.arange(nf) / max(1, nf - 1)))
    incl = [INCL, 85.0]
    _setup(tracer, dump3d_small, W)
    img, spec = tracer.render(incl, freqs)
    for c, inc in enumerate(incl):
        P = O.params(dump3d_small.grid, width=W, height=W, inclination=inc, M_unit=M_UNIT,
                     R_high=R_HIGH)
        ora, gate = O.render(P, dump3d_small.prims, freqs, gate=True)
        for f in range(nf):
            if ora[:, f].max() <= 0:
                assert np.all(img[c, :, f] == 0)
                continue
            r = _report("3d i=%g nu=%.3g" % (inc, freqs[f]), img[c, :, f], ora[:, f])
            assert abs(img[c, :, f].sum() / ora[:, f].sum() - 1) < FLUX_TOL
            assert r.max() < PIX_TOL
            assert spec[c, f] == pytest.approx(img[c, :, f].sum(), rel=1e-13)


def test_c1_full_image(tracer, dump256):
    """The 1024^2 image of the 256^2 dump against the oracle on every pixel,
    with no pixel held apart: rays whose FP32 optical depth passes the
    ln 1000 gate (radiative_transfer.c:90-95) within the margin where an FP32
    and an FP64 tau may decide it differently, at samples that would show in
    the pixel, are traced again with the FP64 transfer.  The gate-adjacent
    pixels (FP64 tau within GATE_EPS of ln 1000 at a sample) are reported."""
    W = 1024
    P = _setup(tracer, dump256, W)
    img, spec = tracer.render([INCL], [FREQ])
    gpu = img[0, :, 0]
    assert np.all(np.isfinite(gpu)) and gpu.min() >= 0
    ora, gate = O.render(P, dump256.prims, [FREQ], gate=True)
    ora, gate = ora[:, 0], gate[:, 0]
    r = _report("c1", gpu, ora)
    near = gate < GATE_EPS
    print("c1: %d pixels cross the gate, %d within %g of it (max rel %.3g there); "
          "elsewhere max rel %.3g" % (np.isfinite(gate).sum(), near.sum(), GATE_EPS,
                                      r[near].max() if near.any() else 0., r[~near].max()))
    assert abs(gpu.sum() / ora.sum() - 1) < FLUX_TOL
    assert r.max() < PIX_TOL
    img2, spec2 = tracer.render([INCL], [FREQ])
    assert spec2[0, 0] == spec[0, 0]
    assert np.array_equal(img2, img)


@pytest.mark.parametrize("case", sorted(SPACETIME_CASES))
def test_other_spacetimes_and_step_sizes(tracer, case):
    """Schwarzschild, a = 0.99, h = 1, negative spin, and step sizes 0.1
    (RK2 points beyond the angle-addition range: full evaluation) and 0.01,
    at 86 GHz with R_high = 1 (the a = 0 torus is too cold to shine at 230 GHz
    within FP32 range): image, flux and step counts against the oracle."""
    dump, stepsize = spacetime_case(case, 48)
    W, nu, incl = 48, 8.6e10, (20.0 if case == "a0.99" else INCL)
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=1.0, stepsize=stepsize)
    tracer.load_dump(dump)
    P = O.params(dump.grid, width=W, height=W, inclination=incl, M_unit=M_UNIT, R_high=1.0,
                 stepsize=stepsize)
    n = W * W
    d, dc = tracer.alloc(n * 8), tracer.alloc(n * 8)
    try:
        tracer.render_device([incl], [nu], d, None, dc)
        tracer.synchronize()
        gpu = tracer.d2h(d, (n,), np.float64)
        cnt = tracer.d2h(dc, (n, 2), np.int32)
    finally:
        tracer.free(d)
        tracer.free(dc)
    ora, ocnt, gate = O.render(P, dump.prims, [nu], counters=True, gate=True)
    assert ora.max() > 1e-12
    r = _report(case, gpu, ora[:, 0])
    assert abs(gpu.sum() / ora[:, 0].sum() - 1) < FLUX_TOL
    assert r.max() < PIX_TOL
    assert counters_agree(cnt, ocnt).mean() > 0.99


@pytest.mark.parametrize("integrator", [rt.INT_RK4, rt.INT_VERLET])
def test_other_integrators(tracer, dump128, integrator):
    """RK4 and velocity Verlet (integrator.c:13-105, chosen by int_method):
    image, flux and step counts against the oracle's."""
    W = 64
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH, integrator=integrator)
    tracer.load_dump(dump128)
    P = O.params(dump128.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=R_HIGH, integrator=integrator)
    n = W * W
    d, dc = tracer.alloc(n * 8), tracer.alloc(n * 8)
    try:
        tracer.render_device([INCL], [FREQ], d, None, dc)
        tracer.synchronize()
        gpu = tracer.d2h(d, (n,), np.float64)
        cnt = tracer.d2h(dc, (n, 2), np.int32)
    finally:
        tracer.free(d)
        tracer.free(dc)
    ora, ocnt, gate = O.render(P, dump128.prims, [FREQ], counters=True, gate=True)
    r = _report("integrator %d" % integrator, gpu, ora[:, 0])
    assert abs(gpu.sum() / ora[:, 0].sum() - 1) < FLUX_TOL
    assert r.max() < PIX_TOL
    assert counters_agree(cnt, ocnt).mean() > 0.99


def test_bad_integrator_rejected(tracer):
    with pytest.raises(rt.RtError, match="integrator"):
        tracer.set_params(width=8, height=8, M_unit=M_UNIT, R_high=R_HIGH, integrator=3)


@pytest.mark.parametrize("emission", [rt.EMIS_THSYNCHAV, rt.EMIS_KAPPA])
def test_other_emission_models(tracer, dump128, emission):
    """The angle-averaged thermal fit and the kappa fits (j_nu.c:28-106,
    157-167, which emission_coeff does not call; chosen here by
    rt_params.emission): images at one frequency and at three (the 16-wide
    kernel) against the oracle's, with and without the approach kernel
    (bit-identical)."""
    W = 48
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH, emission=emission)
    tracer.load_dump(dump128)
    P = O.params(dump128.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=R_HIGH, emission=emission)
    try:
        for freqs in ([FREQ], [8.6e10, FREQ, 6.9e11]):
            imgs = []
            for approach in (rt.APPROACH_NONE, 100):
                tracer.set_launch_config(approach=approach)
                img, _ = tracer.render([INCL], freqs)
                imgs.append(img[0])
            assert np.array_equal(imgs[0], imgs[1])
            ora, gate = O.render(P, dump128.prims, freqs, gate=True)
            for f in range(len(freqs)):
                r = _report("emission %d nu %.3g" % (emission, freqs[f]), imgs[0][:, f],
                            ora[:, f])
                assert ora[:, f].max() > 0
                assert abs(imgs[0][:, f].sum() / ora[:, f].sum() - 1) < FLUX_TOL
                assert r[~(gate[:, f] < GATE_EPS)].max() < PIX_TOL
    finally:
        tracer.set_launch_config()


def test_other_emission_models_3d(tracer, dump3d_small):
    """The kappa fits on a 3D dump (trilinear sampler kernels)."""
    W = 24
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH, emission=rt.EMIS_KAPPA)
    tracer.load_dump(dump3d_small)
    P = O.params(dump3d_small.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=R_HIGH, emission=O.KAPPA)
    img, _ = tracer.render([INCL], [FREQ])
    ora, gate = O.render(P, dump3d_small.prims, [FREQ], gate=True)
    r = _report("kappa 3d", img[0, :, 0], ora[:, 0])
    assert abs(img[0, :, 0].sum() / ora[:, 0].sum() - 1) < FLUX_TOL
    assert r[~(gate[:, 0] < GATE_EPS)].max() < PIX_TOL


def test_emission_model_errors(tracer, dump128):
    with pytest.raises(rt.RtError, match="emission"):
        tracer.set_params(width=8, height=8, M_unit=M_UNIT, R_high=R_HIGH, emission=3)
    tracer.set_params(width=8, height=8, M_unit=M_UNIT, R_high=R_HIGH, emission=rt.EMIS_KAPPA)
    tracer.load_dump(dump128)
    with pytest.raises(rt.RtError, match="thermal synchrotron"):
        tracer.render([INCL], [FREQ], grad=True)
    with pytest.raises(rt.RtError, match="emission model"):
        tracer.probe(rt.PROBE_EMISSION_MODEL, [[10.0, 0.5, 10.0, 2.3e11, 1e5, 7.0]])


def test_probe_emission_models(tracer):
    """Device FP32 fits of every emission model against the FP64 oracle over
    plasma states from the Theta_e floor to 300, including the gates (nu_p
    floor, |cos pitch| = 1, |z| = 1 in the kappa absorptivity)."""
    rng = np.random.default_rng(9)
    n = 600
    for model in (O.THSYNCH, O.THSYNCHAV, O.KAPPA):
        q = np.column_stack([10 ** rng.uniform(-1, 3, n), rng.uniform(-1, 1, n),
                             10 ** rng.uniform(-3, 2.5, n), 10 ** rng.uniform(9, 13, n),
                             10 ** rng.uniform(2, 8, n), np.full(n, float(model))])
        q[:10, 2] = 1e-20
        q[10:20, 2] = 1.0 / 3.5
        q[20:30, 3] = 1.0
        q[30:40, 1] = 1.0
        out = tracer.probe(rt.PROBE_EMISSION_MODEL, q)
        for i in range(n):
            j, a = O.emission(model, *q[i, :5])
            nuc = 4.8032068e-10 * q[i, 0] / (2 * np.pi * 9.1093826e-28 * 2.99792458e10)
            # exp(-X^(1/3)) (thermal) and exp(-1.8899 x_M^(1/3)) (averaged) multiply
            # the FP32 relative error of the cube root by that exponent
            X = q[i, 3] / max(nuc * q[i, 2] ** 2, 1e-300)
            tol = 5e-4 + 2e-6 * np.cbrt(X)
            x = 6.6260693e-27 * q[i, 3] / (q[i, 2] * 9.1093826e-28 * 2.99792458e10 ** 2)
            tola = tol + (4e-16 / x if model != O.KAPPA else 0.0)
            for got, ref, t in ((out[i, 0], j, tol), (out[i, 1], a, tola)):
                if np.isnan(ref) or np.isnan(got):
                    assert np.isnan(ref) and np.isnan(got), (model, q[i], got, ref)
                elif ref == 0.0 or got == 0.0:
                    assert abs(got - ref) <= 1e-30, (model, q[i], got, ref)
                else:
                    assert got == pytest.approx(ref, rel=t), (model, q[i], got, ref)


# ---- building blocks ------------------------------------------------------


def _rand_points(rng, n, grid):
    X = np.zeros((n, 4))
    X[:, 1] = rng.uniform(np.log(1.5), np.log(1e4), n)
    X[:, 2] = rng.uniform(0.02, 0.98, n)
    X[:, 3] = rng.uniform(0, 2 * np.pi, n)
    return X


def test_probe_metric(tracer, dump128):
    P = _setup(tracer, dump128, 8)
    rng = np.random.default_rng(1)
    X = _rand_points(rng, 200, dump128.grid)
    out = tracer.probe(rt.PROBE_METRIC, X)
    for i in range(len(X)):
        np.testing.assert_allclose(out[i, :16].reshape(4, 4), O.metric_dd(P, X[i]),
                                   rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(out[i, 16:].reshape(4, 4), O.metric_uu(P, X[i]),
                                   rtol=1e-13, atol=1e-15)


def test_probe_accel_and_rk2(tracer, dump128):
    P = _setup(tracer, dump128, 8)
    rng = np.random.default_rng(2)
    X = _rand_points(rng, 100, dump128.grid)
    K = rng.normal(size=(100, 4))
    acc = tracer.probe(rt.PROBE_ACCEL, np.hstack([X, K]))
    y = np.hstack([X, K])
    step = tracer.probe(rt.PROBE_RK2, y)
    for i in range(len(X)):
        G = O.connection(P, X[i])
        A = -np.einsum("mab,a,b->m", G, K[i], K[i])
        np.testing.assert_allclose(acc[i], A, rtol=1e-11, atol=1e-11 * np.abs(A).max())
        dl = O.stepsize(P, X[i], K[i])
        assert step[i, 0] == pytest.approx(dl, rel=1e-14)
        # X^0 (coordinate time) is not advanced by the tracer: nothing reads it
        np.testing.assert_allclose(step[i, 2:], O.rk2_step(P, y[i], dl)[1:], rtol=1e-11,
                                   atol=1e-12)


def test_probe_camera(tracer, dump128):
    _setup(tracer, dump128, 8)
    rng = np.random.default_rng(3)
    pts = [(a, b, i) for a, b, i in zip(rng.uniform(-20, 20, 50), rng.uniform(-20, 20, 50),
                                        rng.choice([10.0, 60.0, 85.0], 50))]
    out = tracer.probe(rt.PROBE_CAMERA, np.array(pts))
    for k, (a, b, i) in enumerate(pts):
        P = O.params(dump128.grid, width=8, height=8, inclination=i, M_unit=M_UNIT,
                     R_high=R_HIGH)
        ref = O.init_photon(P, a, b)
        np.testing.assert_allclose(out[k, 1:], ref[1:], rtol=1e-12, atol=1e-16)


def test_probe_emission(tracer):
    """FP32 emissivity and absorptivity of the transfer pass against the FP64
    oracle functions."""
    rng = np.random.default_rng(4)
    n = 300
    q = np.column_stack([10 ** rng.uniform(-1, 3, n), rng.uniform(-1, 1, n),
                         10 ** rng.uniform(-5, 2, n), 10 ** rng.uniform(9, 13, n),
                         10 ** rng.uniform(2, 8, n)])
    # every other row: fields of 1e15-1e17 G just above the Theta_e gate at
    # 3e15-3e16 Hz, where h nu / (m c^2 Theta_e) spans 0.08-2.4 while
    # X^(1/3) stays below ~20, so the Planck factor's exp branch (|x| > 1/4)
    # runs in waves whose other lanes take the series
    far = np.column_stack([10 ** rng.uniform(15, 17, n // 2), rng.uniform(-0.9, 0.9, n // 2),
                           10 ** rng.uniform(-3.99, -3.5, n // 2),
                           10 ** rng.uniform(15.5, 16.5, n // 2), 10 ** rng.uniform(2, 8, n // 2)])
    q[1::2] = far
    out = tracer.probe(rt.PROBE_EMISSION, q)
    xs = 6.6260693e-27 * q[:, 3] / (q[:, 2] * 9.1093826e-28 * 2.99792458e10 ** 2)
    assert (xs > 0.25).sum() > 20 and (xs[1::2] < 0.25).sum() > 20
    lib = O.lib()
    for i in range(n):
        j = lib.or_j_thsynch(q[i, 0], q[i, 1], q[i, 2], q[i, 3], q[i, 4])
        a = lib.or_a_th(j, q[i, 3], q[i, 2])
        # exp(-X^(1/3)) multiplies the FP32 relative error of X^(1/3) (~1e-6:
        # the inputs' rounding and the exp2/log2 estimate) by X^(1/3)
        nuc = 4.8032068e-10 * q[i, 0] / (2 * np.pi * 9.1093826e-28 * 2.99792458e10)
        X = q[i, 3] / ((2.0 / 9.0) * nuc * q[i, 2] ** 2 * np.sqrt(1 - q[i, 1] ** 2))
        tol = 2e-5 + 1e-6 * np.cbrt(X)
        assert out[i, 0] == pytest.approx(j, rel=tol, abs=1e-300)
        # the oracle's Planck factor exp(x) - 1 (j_nu.c:190) loses ~1e-16/x;
        # the kernel's (a series below |x| = 1/4) does not
        x = 6.6260693e-27 * q[i, 3] / (q[i, 2] * 9.1093826e-28 * 2.99792458e10 ** 2)
        if np.isfinite(a):
            assert out[i, 1] == pytest.approx(a, rel=tol + 4e-16 / x, abs=1e-250)
        else:
            assert np.isnan(out[i, 1]) == np.isnan(a)
    assert np.count_nonzero(out[1::2, 1][xs[1::2] > 0.25]) > 20


def test_fluid_params_hook_3d_periodic(tracer, dump3d_small):
    """The 3D sampler's periodic phi: positions at, just below and just above
    multiples of the period, negative and many periods out, against the
    oracle's fmod-based index."""
    P = _setup(tracer, dump3d_small, 8)
    g = dump3d_small.grid
    period = g.n3 * g.dx3
    rng = np.random.default_rng(12)
    offs = [0.0, 1e-15 * period, -1e-15 * period, 0.5 * g.dx3, -0.5 * g.dx3, g.dx3,
            period - 1e-15 * period, 0.25 * period]
    X3 = [g.startx3 + m * period + o for m in (-3, -1, 0, 1, 2, 7) for o in offs]
    X3 += list(g.startx3 + rng.uniform(-20, 20, 52) * period)
    n = len(X3)
    X = np.zeros((n, 4))
    X[:, 1] = rng.uniform(g.startx1 + 0.3 * (g.stopx1 - g.startx1), g.stopx1, n)
    X[:, 2] = rng.uniform(g.startx2 + 0.2, g.stopx2 - 0.2, n)
    X[:, 3] = X3
    out = tracer.fluid_params(X)
    checked = 0
    for i in range(n):
        ref = O.fluid_at(P, dump3d_small.prims, X[i])
        assert bool(out[i, 11]) == ref["ok"]
        if ref["ok"]:
            assert out[i, 0] == pytest.approx(ref["ne"], rel=1e-6), X[i]
            assert out[i, 1] == pytest.approx(ref["thetae"], rel=1e-5), X[i]
            assert out[i, 2] == pytest.approx(ref["B"], rel=1e-4), X[i]
            checked += 1
    assert checked > n // 2


def test_fluid_params_hook(tracer, dump128):
    P = _setup(tracer, dump128, 8)
    rng = np.random.default_rng(5)
    g = dump128.grid
    X = np.zeros((300, 4))
    X[:, 1] = rng.uniform(g.startx1, g.stopx1, 300)
    X[:, 2] = rng.uniform(g.startx2, g.stopx2, 300)
    out = tracer.fluid_params(X)
    for i in range(len(X)):
        ref = O.fluid_at(P, dump128.prims, X[i])
        assert bool(out[i, 11]) == ref["ok"]
        # the sampler runs in FP32 from FP32 primitives
        np.testing.assert_allclose(out[i, 3:7], ref["ucon"], rtol=2e-6,
                                   atol=2e-6 * np.abs(ref["ucon"]).max())
        np.testing.assert_allclose(out[i, 7:11], ref["bcon"], rtol=1e-4,
                                   atol=1e-4 * np.abs(ref["bcon"]).max() + 1e-30)
        if ref["ok"]:
            assert out[i, 0] == pytest.approx(ref["ne"], rel=1e-6)
            assert out[i, 1] == pytest.approx(ref["thetae"], rel=1e-5)
            assert out[i, 2] == pytest.approx(ref["B"], rel=1e-4)


# ---- error paths ----------------------------------------------------------


def test_error_codes(dump128):
    with rt.Tracer(0) as tr:
        with pytest.raises(rt.RtError, match="rt_set_params"):
            tr.width, tr.height = 4, 4
            tr.render([60.0], [FREQ])
        tr.set_params(width=4, height=4, M_unit=M_UNIT, R_high=R_HIGH)
        with pytest.raises(rt.RtError, match="rt_load_prims"):
            tr.render([60.0], [FREQ])
        # a grid past the samplers' 32-bit cell offsets is refused before its
        # primitives are read (the buffer passed is one cell)
        big = rt.RtGrid(n1=1024, n2=512, n3=512, startx1=0.3, startx2=0., startx3=0.,
                        dx1=0.01, dx2=1. / 512, dx3=2 * np.pi / 512, a=0.9375, hslope=0.3, R0=0.)
        one = np.zeros(8, np.float32)
        rc = tr.lib.rt_load_prims(tr.ctx, ctypes.byref(big),
                                  one.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        assert rc != rt.RT_OK and b"2^27" in tr.lib.rt_last_error(tr.ctx)
        tr.load_dump(dump128)
        with pytest.raises(rt.RtError, match="n_freq"):
            tr.render([60.0], np.full(17, FREQ))
        with pytest.raises(rt.RtError, match="inclination"):
            tr.render([0.0], [FREQ])
        with pytest.raises(rt.RtError, match="positive"):
            tr.render([60.0], [FREQ, 0.0])
        with pytest.raises(rt.RtError, match="n_cam"):
            tr.render([], [FREQ])
        with pytest.raises(rt.RtError, match="n_freq"):
            tr.render([60.0], [])
        with pytest.raises(rt.RtError, match="1e17 Hz"):
            tr.render([60.0], [FREQ, 2e17])
        tr.render([60.0], [1e17])
        with pytest.raises(rt.RtError):
            tr.set_params(width=0, height=4, M_unit=M_UNIT, R_high=R_HIGH)
