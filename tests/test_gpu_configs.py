"""The benchmark workloads at their full sizes on the GPU, through the C-ABI,
against the FP64 oracle on pixel samples:

  c2  the 256^2 torus at 512^2 x 6 inclinations (10-85 deg) x 8 frequencies
      (86-690 GHz, log-spaced) in one batched launch: 48 images
  c3  the 192x96x96 torus with m = 1..4 azimuthal modes (10%) at 1024^2
  c4  the image and its gradient d/d(ln M_unit, R_high) on the c3 dump at
      512^2, R_high in {1, 10, 40, 160}, against central differences (step
      1e-3) of the oracle

Bar: flux within 1e-4 relative of the oracle, every pixel within
1e-3 relative with an absolute floor of 1e-6 max(I); gradients as in
test_gpu_grad.py.  The flux of a sampled image is the sum over its sampled
pixels."""
import math
import warnings

import numpy as np
import pytest

from conftest import FREQ, INCL, M_UNIT, R_HIGH, rel_err
from oracle import oracle as O
from raptor_amd import rt

pytestmark = pytest.mark.gpu

C2_INCL = [10.0, 25.0, 40.0, 55.0, 70.0, 85.0]
C2_FREQ = list(86e9 * (690.0 / 86.0) ** (np.arange(8) / 7.0))


@pytest.fixture(scope="module")
def dump3d_full():
    from raptor_amd import torus
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return torus.fishbone_moncrief(192, 96, 96)


@pytest.fixture(scope="module")
def tracer():
    tr = rt.Tracer(0)
    yield tr
    tr.close()


def _every64(npix):
    return np.arange(0, npix, 64, dtype=np.int32)


def test_c2_full_workload(tracer, dump256):
    """c2: all 48 images in one launch, every 64th pixel of each inclination
    at all 8 frequencies against the oracle; the batched launch equals the 48
    single-image launches bit for bit; the device spectrum is each image's
    pixel sum."""
    W = 512
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH)
    tracer.load_dump(dump256)
    npix = W * W
    n = len(C2_INCL) * npix * len(C2_FREQ)
    d_img = tracer.alloc(n * 8)
    d_spec = tracer.alloc(len(C2_INCL) * len(C2_FREQ) * 8)
    d_one = tracer.alloc(npix * 8)
    try:
        tracer.render_device(C2_INCL, C2_FREQ, d_img, d_spectrum=d_spec)
        tracer.synchronize()
        img = tracer.d2h(d_img, (len(C2_INCL), npix, len(C2_FREQ)), np.float64)
        spec = tracer.d2h(d_spec, (len(C2_INCL), len(C2_FREQ)), np.float64)
        for c, inc in enumerate(C2_INCL[::5]):
            for f, nu in enumerate(C2_FREQ[::3]):
                tracer.render_device([inc], [nu], d_one)
                tracer.synchronize()
                one = tracer.d2h(d_one, (npix,), np.float64)
                assert np.array_equal(one, img[5 * c, :, 3 * f]), (inc, nu)
    finally:
        for d in (d_img, d_spec, d_one):
            tracer.free(d)
    np.testing.assert_allclose(spec, img.sum(axis=1), rtol=1e-12)
    pix = _every64(npix)
    worst = 0.0
    for c, inc in enumerate(C2_INCL):
        P = O.params(dump256.grid, width=W, height=W, inclination=inc, M_unit=M_UNIT,
                     R_high=R_HIGH)
        ora = O.render(P, dump256.prims, C2_FREQ, pixels=pix)
        for f in range(len(C2_FREQ)):
            g, o = img[c, pix, f], ora[:, f]
            r = rel_err(g, o)
            worst = max(worst, r.max())
            assert abs(g.sum() / o.sum() - 1) < 1e-4, (inc, C2_FREQ[f])
            assert r.max() < 1e-3, (inc, C2_FREQ[f], int(pix[r.argmax()]), r.max())
    print("c2: 48 images, %d sampled pixels each, max pixel rel %.3g" % (pix.size, worst))


def test_c3_full_workload(tracer, dump3d_full):
    """c3: the 1024^2 image of the 192x96x96 dump (56.6 MB of primitives),
    every 64th pixel against the oracle's trilinear, phi-periodic sampler."""
    W = 1024
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH)
    tracer.load_dump(dump3d_full)
    img, spec = tracer.render([INCL], [FREQ])
    g = img[0, :, 0]
    assert np.all(np.isfinite(g)) and g.min() >= 0
    assert spec[0, 0] == pytest.approx(g.sum(), rel=1e-12)
    pix = _every64(W * W)
    P = O.params(dump3d_full.grid, width=W, height=W, inclination=INCL, M_unit=M_UNIT,
                 R_high=R_HIGH)
    ora = O.render(P, dump3d_full.prims, [FREQ], pixels=pix)[:, 0]
    r = rel_err(g[pix], ora)
    print("c3: %d sampled pixels, flux rel %.3g, max pixel rel %.3g"
          % (pix.size, g[pix].sum() / ora.sum() - 1, r.max()))
    assert abs(g[pix].sum() / ora.sum() - 1) < 1e-4
    assert r.max() < 1e-3


@pytest.mark.parametrize("R_high", [1.0, R_HIGH, 40.0, 160.0])
def test_c4_full_workload(tracer, dump3d_full, R_high):
    """c4: image and gradient of the 512^2 image of the 3D dump.  The image on
    every 64th pixel against the oracle; the tangents on 2048 sampled pixels
    against central differences (step 1e-3) of the oracle, pixels whose
    quotient changes between steps h and h/2 (a gate flips inside the step)
    excluded, as in test_gpu_grad.py."""
    W = 512
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_high)
    tracer.load_dump(dump3d_full)
    img, dimg, _ = tracer.render([INCL], [FREQ], grad=True)
    kw = dict(width=W, height=W, inclination=INCL)
    pix = _every64(W * W)
    ora = O.render(O.params(dump3d_full.grid, M_unit=M_UNIT, R_high=R_high, **kw),
                   dump3d_full.prims, [FREQ], pixels=pix)[:, 0]
    r = rel_err(img[0, pix, 0], ora)
    assert abs(img[0, pix, 0].sum() / ora.sum() - 1) < 1e-4
    assert r.max() < 1e-3
    fpix = np.arange(0, W * W, W * W // 2048, dtype=np.int32)[:2048]

    def orad(M, R):
        return O.render(O.params(dump3d_full.grid, M_unit=M, R_high=R, **kw),
                        dump3d_full.prims, [FREQ], pixels=fpix)[:, 0]

    h = 1e-3
    for q in range(2):
        if q == 0:
            f1 = (orad(M_UNIT * math.exp(h), R_high) - orad(M_UNIT * math.exp(-h), R_high)) / (2 * h)
            f2 = (orad(M_UNIT * math.exp(h / 2), R_high) -
                  orad(M_UNIT * math.exp(-h / 2), R_high)) / h
        else:
            f1 = (orad(M_UNIT, R_high * (1 + h)) - orad(M_UNIT, R_high * (1 - h))) / (2 * h * R_high)
            f2 = (orad(M_UNIT, R_high * (1 + h / 2)) - orad(M_UNIT, R_high * (1 - h / 2))) / (h * R_high)
        an = dimg[0, fpix, 0, q]
        den = np.maximum(np.abs(an), 1e-3 * np.abs(an).max())
        smooth = np.abs(f1 - f2) / den < 1e-3
        err = np.abs(f1 - an) / den
        print("c4 R_high=%g d%d: smooth %.3f, within 1e-3 %.4f, max %.3g"
              % (R_high, q, smooth.mean(), np.mean(err[smooth] < 1e-3), err[smooth].max()))
        p = 1.0 if q == 0 else R_high
        ipix = np.maximum(img[0, fpix, 0], 1e-6 * img[0, :, 0].max())
        e2 = np.abs(f1 - an) / np.maximum(np.abs(an), ipix / p)
        print("   max error %.3g of the larger of tangent and I/p over smooth pixels"
              % e2[smooth].max())
        assert smooth.mean() > 0.9
        assert np.mean(err[smooth] < 1e-3) > 0.98 and err[smooth].max() < 2e-2
        assert f1[smooth].sum() / an[smooth].sum() == pytest.approx(1.0, abs=5e-4)
        # every smooth pixel to 1e-3 of the larger of its tangent and I / p
        # (a unit logarithmic derivative)
        assert e2[smooth].max() < 1e-3, q


def test_device_spectrum_and_kernel_times(tracer, dump128):
    """rt_render_device's spectrum is each image's sum over its pixels
    (core.c:146-150) for several cameras and frequencies, with and without
    the gradient, bit-identical across calls; rt_render's host spectrum is the
    same; the per-kernel times of a one-chunk render add up to its total and
    are refused after a graph replay."""
    W = 48
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH)
    tracer.load_dump(dump128)
    incl, freqs = [INCL, 20.0], [8.6e10, FREQ, 6.9e11]
    npix = W * W
    n = len(incl) * npix * len(freqs)
    d_img, d_dimg = tracer.alloc(n * 8), tracer.alloc(2 * n * 8)
    d_spec = tracer.alloc(len(incl) * len(freqs) * 8)
    try:
        specs = []
        for grad in (False, True, False):
            tracer.render_device(incl, freqs, d_img, d_dimg if grad else None,
                                 d_spectrum=d_spec)
            tracer.synchronize()
            img = tracer.d2h(d_img, (len(incl), npix, len(freqs)), np.float64)
            spec = tracer.d2h(d_spec, (len(incl), len(freqs)), np.float64)
            np.testing.assert_allclose(spec, img.sum(axis=1), rtol=1e-13)
            specs.append(spec)
            t = tracer.last_kernel_times()
            assert all(v >= 0 for v in t.values()) and t["trace"] > 0
            assert sum(t.values()) == pytest.approx(tracer.last_kernel_ms(), rel=0.02, abs=0.02)
        assert np.array_equal(specs[0], specs[2])
        himg, hspec = tracer.render(incl, freqs)
        assert np.array_equal(hspec, specs[0])
        tracer.capture_begin()
        tracer.render_device(incl, freqs, d_img, d_spectrum=d_spec)
        g = tracer.capture_end()
        try:
            g.launch()
            tracer.synchronize()
            assert np.array_equal(tracer.d2h(d_spec, (len(incl), len(freqs)), np.float64),
                                  specs[0])
            with pytest.raises(rt.RtError):
                tracer.last_kernel_times()
        finally:
            g.destroy()
    finally:
        for d in (d_img, d_dimg, d_spec):
            tracer.free(d)
