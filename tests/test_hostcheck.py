"""CPU build of the tracer's per-ray physics (rt_physics.hpp compiled for the
host, tests/hostcheck) against the oracle.  This checks the kernel's
arithmetic before any GPU run; the GPU path itself is checked in
test_gpu_parity.py through the C-ABI."""
import ctypes
import os
import subprocess
import warnings

import numpy as np
import pytest

from conftest import (FREQ, INCL, M_UNIT, R_HIGH, SPACETIME_CASES, counters_agree, rel_err,
                      spacetime_case)
from oracle import oracle as O
from raptor_amd import rt

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "hostcheck", "hostcheck.hip")
LIB = os.path.join(HERE, "hostcheck", "libhostcheck.so")
PHYS = os.path.join(os.path.dirname(HERE), "raptor_amd", "csrc", "rt_physics.hpp")


def _lib():
    newest = max(os.path.getmtime(SRC), os.path.getmtime(PHYS))
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < newest:
        hipcc = "/opt/rocm/bin/hipcc"
        if not os.path.exists(hipcc):
            pytest.skip("hipcc not available to build the host check")
        subprocess.run([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
                        "-o", LIB, SRC], check=True, capture_output=True)
    return ctypes.CDLL(LIB)


def test_kappa_hypergeometric_table():
    """The kappa absorptivity's hypergeometric factor as the kernels evaluate
    it: log2 of 2F1(kappa - 1/3, kappa + 1; kappa + 2/3; z), z = -2^u, by
    cubic Hermite interpolation of a table built in FP64 over u in [-32, 32]
    (the series beyond), against scipy's hyp2f1 at random points, at the
    table's nodes and midpoints, across z = -1 where its two series forms
    meet, and outside the table (a wrong node spacing, derivative scale or
    branch is off by far more than the FP32 tolerance)."""
    from scipy import special
    L = _lib()
    L.hc_kappa_lhyp.restype = ctypes.c_double
    L.hc_kappa_lhyp.argtypes = [ctypes.c_double]
    k = 3.5
    rng = np.random.default_rng(7)
    nodes = np.arange(-32.0, 32.0, 1.0 / 16.0)
    us = np.concatenate([rng.uniform(-32, 32, 3000), nodes, nodes + 1.0 / 32.0,
                         np.linspace(-0.05, 0.05, 41), [-40.0, -33.0, 32.5, 40.0]])
    # the kernels' argument is the FP32 log2(kappa Theta_e)
    us = us.astype(np.float32).astype(np.float64)
    worst = 0.0
    for u in us:
        want = np.log2(special.hyp2f1(k - 1 / 3, k + 1, k + 2 / 3, -2.0 ** u))
        got = L.hc_kappa_lhyp(u)
        err = abs(got - want) / (1.0 + abs(want))
        worst = max(worst, err)
        assert err <= 2e-6, (u, got, want)
    print("kappa table: worst error %.2e" % worst)


def _render(dump, W, incl, freqs, grad=False, R_high=R_HIGH, stepsize=0.03, integrator=0,
            near_off=None, emission=0, mode=None):
    L = _lib()
    g = dump.grid
    grid = rt.RtGrid(n1=g.n1, n2=g.n2, n3=g.n3, startx1=g.startx1, startx2=g.startx2,
                     startx3=g.startx3, dx1=g.dx1, dx2=g.dx2, dx3=g.dx3, a=g.a, hslope=g.hslope,
                     R0=g.R0)
    p = rt.RtParams(MBH=8.9505e39, M_unit=M_UNIT, R_high=R_high, R_low=1.0, width=W, height=W,
                    cam_size_x=40.0, cam_size_y=40.0, stepsize=stepsize, max_steps=100000,
                    integrator=integrator, emission=emission)
    prims = np.ascontiguousarray(dump.prims, np.float32)
    incl = np.ascontiguousarray(incl, float)
    fr = np.ascontiguousarray(freqs, float)
    n = incl.size * W * W
    img = np.zeros(n * fr.size)
    dimg = np.zeros(2 * n * fr.size) if grad else None
    cnt = np.zeros(2 * n, np.int32)
    vp = ctypes.c_void_p
    if mode is not None:
        rc = L.hc_render_mode(ctypes.byref(p), ctypes.byref(grid), vp(prims.ctypes.data),
                              vp(incl.ctypes.data), incl.size, vp(fr.ctypes.data), fr.size,
                              mode, vp(img.ctypes.data), vp(dimg.ctypes.data) if grad else None,
                              vp(cnt.ctypes.data))
    elif near_off is None:
        rc = L.hc_render(ctypes.byref(p), ctypes.byref(grid), vp(prims.ctypes.data),
                         vp(incl.ctypes.data), incl.size, vp(fr.ctypes.data), fr.size,
                         vp(img.ctypes.data), vp(dimg.ctypes.data) if grad else None,
                         vp(cnt.ctypes.data))
    else:
        rc = L.hc_render_split(ctypes.byref(p), ctypes.byref(grid), vp(prims.ctypes.data),
                               vp(incl.ctypes.data), incl.size, vp(fr.ctypes.data), fr.size,
                               1, ctypes.c_double(near_off), vp(img.ctypes.data),
                               vp(dimg.ctypes.data) if grad else None, vp(cnt.ctypes.data))
    assert rc == 0
    return img.reshape(n, fr.size), dimg, cnt.reshape(n, 2)


def test_fp64_transfer_matches_oracle(dump128):
    """The FP64 transfer the kernels use for rays at the tau gate
    (transfer_sample_f64), here on every ray of an image at three
    frequencies, against the oracle: both FP64, so the images and their
    tangents agree to rounding (the geodesics differ at ~1e-13 by their
    formulation), far inside the FP32 transfer's 1e-4."""
    W = 24
    freqs = [8.6e10, FREQ, 6.9e11]
    img, dimg, cnt = _render(dump128, W, [INCL], freqs, grad=True, mode=1)
    P = O.params(dump128.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=R_HIGH)
    oimg, odimg, ocnt = O.render(P, dump128.prims, freqs, grad=True, counters=True)
    # all-FP64 rays end at the sample where tau passes the gate, as the
    # oracle's opaque step
    op = ocnt[:, 5]
    same = np.where(op >= 0, cnt[:, 0] == op + 1,
                    (cnt[:, 0] == ocnt[:, 0]) & (cnt[:, 1] == ocnt[:, 2]))
    assert same.mean() > 0.99
    for f in range(3):
        r = rel_err(img[same, f], oimg[same, f], 1e-9)
        print("nu=%.3g: FP64 transfer max rel %.2e" % (freqs[f], r.max()))
        assert r.max() < 1e-5
        d = dimg.reshape(-1, 3, 2)
        for q in range(2):
            assert rel_err(d[same, f, q], odimg[same, f, q], 1e-7).max() < 1e-8


def test_rays_at_the_gate_are_traced_again_in_fp64(dump256):
    """The FP32 transfer flags rays whose tau passes ln 1000 within the
    margin at samples carrying more than kGateFlip of the pixel; those, and
    only those, are traced again with the FP64 transfer: every pixel of the
    kernels' image equals either the FP32-only image's or the all-FP64
    image's, some take the FP64 value, and the image meets the per-pixel bar
    against the oracle with no pixel held apart (64x64 of the 1024^2
    workload's field of view, 230 GHz)."""
    W = 128
    got, _, _ = _render(dump256, W, [INCL], [FREQ], mode=0)
    f32, _, _ = _render(dump256, W, [INCL], [FREQ], mode=2)
    f64, _, _ = _render(dump256, W, [INCL], [FREQ], mode=1)
    g, a, b = got[:, 0], f32[:, 0], f64[:, 0]
    retraced = g != a
    assert np.all((g == a) | (g == b))
    print("rays traced again in FP64: %d of %d" % (retraced.sum(), g.size))
    assert 0 < retraced.sum() < 1e-3 * g.size
    P = O.params(dump256.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=R_HIGH)
    ora = O.render(P, dump256.prims, [FREQ])[:, 0]
    assert rel_err(g, ora).max() < 1e-3
    assert rel_err(b[retraced], ora[retraced]).max() < 1e-5


def test_host_build_matches_oracle_image_and_counters(dump128):
    W = 32
    img, dimg, cnt = _render(dump128, W, [INCL], [FREQ], grad=True)
    P = O.params(dump128.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=R_HIGH)
    oimg, odimg, ocnt = O.render(P, dump128.prims, [FREQ], grad=True, counters=True)
    # FP32 transfer against the FP64 oracle (the geodesics are FP64 in both)
    assert rel_err(img[:, 0], oimg[:, 0]).max() < 1e-4
    assert abs(img[:, 0].sum() / oimg[:, 0].sum() - 1) < 1e-6
    # rounding differences (sincospi, the step-size form) may move a rare ray
    # across a discrete event (grid edge, termination) by one step
    assert counters_agree(cnt, ocnt).mean() > 0.99
    d = dimg.reshape(-1, 2)
    for q in range(2):
        # a pixel's gradient is a sum of terms of both signs; FP32 rounding of
        # the terms shows up relative to a small net value on optically thick rays
        r = rel_err(d[:, q], odimg[:, 0, q], 1e-5)
        assert np.mean(r < 1e-3) > 0.99 and r.max() < 1e-2
        assert d[:, q].sum() / odimg[:, 0, q].sum() == pytest.approx(1.0, abs=1e-4)


def test_host_build_multi_frequency(dump128):
    W = 12
    freqs = [8.6e10, 2.3e11, 6.9e11]
    img, _, _ = _render(dump128, W, [30.0], freqs)
    P = O.params(dump128.grid, width=W, height=W, inclination=30.0, M_unit=M_UNIT,
                 R_high=R_HIGH)
    oimg = O.render(P, dump128.prims, freqs)
    for f in range(3):
        assert rel_err(img[:, f], oimg[:, f]).max() < 1e-4


def test_batched_frequencies_equal_single_frequency_renders_with_retrace(dump256):
    """Eight frequencies in one render equal eight single-frequency renders
    bit for bit, also where rays are traced again with the FP64 transfer:
    the second trace decides per frequency, as a launch of that frequency
    alone does (c2's frequencies; some rays are flagged at 115-380 GHz)."""
    W = 64
    freqs = list(86e9 * (690.0 / 86.0) ** (np.arange(8) / 7.0))
    many, _, _ = _render(dump256, W, [INCL], freqs)
    f32, _, _ = _render(dump256, W, [INCL], freqs, mode=2)
    assert (many != f32).any()
    for f, nu in enumerate(freqs):
        one, _, _ = _render(dump256, W, [INCL], [nu])
        np.testing.assert_array_equal(many[:, f], one[:, 0])


def test_midpoint_trigonometry_matches_full_evaluation():
    """shifted_point (Taylor + angle addition from the start point) against a
    full exp/sincospi evaluation at the shifted point, over the half and full
    steps the integrator produces (|dl k^1| <= 0.03, |dl k^2| <= 0.03
    min(X2, 1-X2)), at their bounds, and beyond the guarded range (full
    fallback)."""
    L = _lib()
    L.hc_shifted_point.argtypes = [ctypes.c_double] * 5 + [ctypes.c_void_p]
    rng = np.random.default_rng(11)
    out = np.zeros(10)
    for _ in range(2000):
        X1 = rng.uniform(0.3, 9.3)
        X2 = rng.uniform(1e-4, 1 - 1e-4)
        e1 = rng.choice([rng.uniform(-0.03, 0.03), 0.03, -0.03]) * rng.choice([1, 1, 3])
        e2 = (rng.choice([rng.uniform(-0.03, 0.03), 0.03, -0.03]) * min(X2, 1 - X2) *
              rng.choice([1, 1, 3]))
        h = rng.choice([0.0, 0.3, 0.7, 1.0])
        L.hc_shifted_point(X1, X2, e1, e2, h, ctypes.c_void_p(out.ctypes.data))
        a, b = out[:5], out[5:]
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("case", sorted(SPACETIME_CASES))
def test_host_build_other_spacetimes(case):
    """The tracer's physics on the CPU against the oracle for other spins,
    h = 1 and step sizes beyond the angle-addition range (86 GHz, R_high = 1:
    the a = 0 torus is too cold to shine at 230 GHz within FP32 range)."""
    dump, stepsize = spacetime_case(case, 32)
    W = 20
    img, _, cnt = _render(dump, W, [INCL], [8.6e10], R_high=1.0, stepsize=stepsize)
    P = O.params(dump.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=1.0, stepsize=stepsize)
    oimg, ocnt = O.render(P, dump.prims, [8.6e10], counters=True)
    assert oimg.max() > 1e-12
    assert rel_err(img[:, 0], oimg[:, 0]).max() < 1e-3
    assert abs(img[:, 0].sum() / oimg[:, 0].sum() - 1) < 1e-4
    assert counters_agree(cnt, ocnt).mean() > 0.99


@pytest.mark.parametrize("integrator", [O.RK4, O.VERLET])
def test_host_build_other_integrators(dump128, integrator):
    """RK4 and velocity Verlet (integrator.c:13-105), the steppers int_method
    selects, on the CPU build against the oracle's."""
    W = 24
    img, _, cnt = _render(dump128, W, [INCL], [FREQ], integrator=integrator)
    P = O.params(dump128.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=R_HIGH, integrator=integrator)
    oimg, ocnt = O.render(P, dump128.prims, [FREQ], counters=True)
    assert rel_err(img[:, 0], oimg[:, 0]).max() < 1e-3
    assert abs(img[:, 0].sum() / oimg[:, 0].sum() - 1) < 1e-4
    assert counters_agree(cnt, ocnt).mean() > 0.99
    # a different stepper gives a measurably different image
    P2 = O.params(dump128.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                  R_high=R_HIGH)
    rk2 = O.render(P2, dump128.prims, [FREQ])
    assert np.abs(rk2[:, 0] - oimg[:, 0]).max() > 1e-6 * np.abs(oimg).max()


@pytest.mark.parametrize("integrator", [O.RK2, O.RK4, O.VERLET])
def test_split_approach_equals_one_loop(dump128, integrator):
    """The split launch's path (approach_ray to the grid, the ray saved before
    its first step inside and resumed from there) gives the same bits as one
    loop from the camera: with the kernel's save margin (4 STEPSIZE), with a
    margin of 0 and 1 STEPSIZE, and with a negative margin below every X1,
    which forces the fallback that traces the ray again from the camera and
    saves before every step."""
    W = 16
    stepsize = 0.03
    ref = _render(dump128, W, [INCL, 10.0], [FREQ], grad=True, integrator=integrator)
    for off in (4 * stepsize, stepsize, 0.0, 1e-300):
        got = _render(dump128, W, [INCL, 10.0], [FREQ], grad=True, integrator=integrator,
                      near_off=off)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), off
    forced = _render(dump128, W, [INCL, 10.0], [FREQ], grad=True, integrator=integrator,
                     near_off=-100.0)
    for a, b in zip(forced, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("model", [O.THSYNCHAV, O.KAPPA])
def test_host_build_emission_models_unit(model):
    """The FP32 fits of the other emission models (host build) against the
    oracle's FP64 restatement over the plasma states a torus sample meets and
    beyond (Theta_e from the 1e-20 floor to 300, B, n_e, nu_p over decades)."""
    L = _lib()
    L.hc_emission.argtypes = [ctypes.c_int] + [ctypes.c_double] * 5 + [ctypes.c_void_p]
    rng = np.random.default_rng(7 + model)
    out = np.zeros(2)
    n = 4000
    q = np.column_stack([10 ** rng.uniform(-1, 3, n), rng.uniform(-1, 1, n),
                         10 ** rng.uniform(-3, 2.5, n), 10 ** rng.uniform(9, 13, n),
                         10 ** rng.uniform(2, 8, n)])
    q[:20, 2] = 1e-20
    q[20:40, 2] = 1.0 / 3.5          # |z| = 1 in the kappa absorptivity
    q[40:60, 3] = 1.0                # nu_p floor
    for row in q:
        L.hc_emission(model, *row, ctypes.c_void_p(out.ctypes.data))
        j, a = O.emission(model, *row)
        # Kirchhoff's absorptivity: the oracle's Planck factor exp(x) - 1
        # (j_nu.c:190) loses ~1e-16/x, the kernel's e^x - 1 does not
        x = 6.6260693e-27 * row[3] / (row[2] * 9.1093826e-28 * 2.99792458e10 ** 2)
        tola = 2e-4 + (4e-16 / x if model == O.THSYNCHAV else 0.0)
        for got, ref, tol in ((out[0], j, 2e-4), (out[1], a, tola)):
            if np.isnan(ref) or np.isnan(got):
                assert np.isnan(ref) and np.isnan(got), (row, got, ref)
            elif ref == 0.0 or got == 0.0:
                # FP32 underflow against FP64 values far below any pixel's noise
                assert abs(got - ref) <= 1e-30, (row, got, ref)
            else:
                assert got == pytest.approx(ref, rel=tol), (row, got, ref)


@pytest.mark.parametrize("model", [O.THSYNCHAV, O.KAPPA])
def test_host_build_emission_models_image(dump128, model):
    """Whole images with the other emission models: the tracer's physics on
    the CPU against the oracle (1e-3 per pixel off the tau gate, 1e-4 in
    flux), and different from the thermal image."""
    W = 24
    img, _, cnt = _render(dump128, W, [INCL], [FREQ], emission=model)
    P = O.params(dump128.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=R_HIGH, emission=model)
    oimg, ocnt, gate = O.render(P, dump128.prims, [FREQ], counters=True, gate=True)
    assert oimg.max() > 0
    near = gate[:, 0] < 1e-4
    assert rel_err(img[:, 0], oimg[:, 0])[~near].max() < 1e-3
    assert abs(img[:, 0].sum() / oimg[:, 0].sum() - 1) < 1e-4
    th = O.render(O.params(dump128.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                           R_high=R_HIGH), dump128.prims, [FREQ])
    assert abs(th[:, 0].sum() / oimg[:, 0].sum() - 1) > 1e-3
