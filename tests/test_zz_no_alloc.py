"""rt_render_device_noalloc through the C-ABI against rt_render_device: the
same image, tangents, per-ray counters and spectrum, bit for bit, on two
calls in a row -- the second smaller, so that nothing the first left in the
library's scratch array may reach it -- and again once rt_render_device has
built the centre-out claim order it then reuses; and its size limit."""
import numpy as np
import pytest

from conftest import FREQ, INCL, M_UNIT, R_HIGH
from raptor_amd import rt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tracer():
    tr = rt.Tracer(0)
    yield tr
    tr.close()


def _render(tr, fn, incl, freqs, grad):
    """Outputs of one call of fn (render_device or render_device_noalloc):
    image, tangents (or None), counters, spectrum."""
    npix = tr.width * tr.height
    ncam, nf = len(incl), len(freqs)
    n = ncam * npix * nf
    d_img = tr.alloc(n * 8)
    d_dimg = tr.alloc(2 * n * 8) if grad else None
    d_cnt = tr.alloc(ncam * npix * 8)
    d_spec = tr.alloc(ncam * nf * 8)
    try:
        fn(incl, freqs, d_img, d_dimg, d_cnt, d_spec)
        tr.synchronize()
        return (tr.d2h(d_img, (ncam, npix, nf), np.float64),
                tr.d2h(d_dimg, (ncam, npix, nf, 2), np.float64) if grad else None,
                tr.d2h(d_cnt, (ncam * npix, 2), np.int32),
                tr.d2h(d_spec, (ncam, nf), np.float64))
    finally:
        for d in (d_img, d_dimg, d_cnt, d_spec):
            if d:
                tr.free(d)


def _same(a, b):
    for x, y in zip(a, b):
        if x is None:
            assert y is None
        else:
            assert np.array_equal(x, y)


def test_noalloc_equals_render_device(tracer, dump256, dump128):
    # 1: the 1024^2 image and its tangents (rt_render_device: approach +
    # resumed trace, centre-out order; noalloc: one launch, tiled scan order)
    print("noalloc: 1024^2, 1 frequency, with tangents", flush=True)
    tracer.set_params(width=1024, height=1024, M_unit=M_UNIT, R_high=R_HIGH)
    tracer.load_dump(dump256)
    first = _render(tracer, tracer.render_device_noalloc, [INCL], [FREQ], True)
    ref = _render(tracer, tracer.render_device, [INCL], [FREQ], True)
    _same(first, ref)
    assert first[0].max() > 0 and np.isfinite(first[1]).all()
    # 2: smaller, 2 inclinations x 3 frequencies (the FP64 re-trace's list in
    # the scratch); noalloc first, then after rt_render_device built the
    # centre-out order for this shape
    print("noalloc: 96^2, 2 inclinations x 3 frequencies", flush=True)
    tracer.set_params(width=96, height=96, M_unit=M_UNIT, R_high=R_HIGH)
    tracer.load_dump(dump128)
    incl, freqs = [INCL, 20.0], [8.6e10, FREQ, 6.9e11]
    second = _render(tracer, tracer.render_device_noalloc, incl, freqs, False)
    ref = _render(tracer, tracer.render_device, incl, freqs, False)
    _same(second, ref)
    again = _render(tracer, tracer.render_device_noalloc, incl, freqs, False)
    _same(again, ref)
    np.testing.assert_allclose(second[3], second[0].sum(axis=1), rtol=1e-13)


def test_noalloc_size_limit(tracer, dump128):
    """More work items than the scratch holds are refused before any launch
    (the buffers passed are full-sized all the same)."""
    W = 2048
    assert 2 * W * W > rt.NOALLOC_MAX_RAYS >= W * W
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH)
    tracer.load_dump(dump128)
    d_img = tracer.alloc(2 * W * W * 8)
    try:
        with pytest.raises(rt.RtError, match="scratch"):
            tracer.render_device_noalloc([INCL, 20.0], [FREQ], d_img)
        tracer.synchronize()
    finally:
        tracer.free(d_img)
    # and the next call is served
    print("noalloc: 48^2 after a refused call", flush=True)
    tracer.set_params(width=48, height=48, M_unit=M_UNIT, R_high=R_HIGH)
    out = _render(tracer, tracer.render_device_noalloc, [INCL], [FREQ], False)
    ref = _render(tracer, tracer.render_device, [INCL], [FREQ], False)
    _same(out, ref)
