"""Batched launches, launch-shape knobs, graph replay and the autotuner.

A ray's arithmetic does not depend on which lane, wave or launch traces it,
so every launch shape must give bit-identical images, and a batched launch
over cameras x frequencies must equal the per-image launches."""
import warnings

import numpy as np
import pytest

from conftest import FREQ, INCL, M_UNIT, R_HIGH
from raptor_amd import autotune, rt, torus

pytestmark = pytest.mark.gpu

C2_INCL = [10.0, 25.0, 40.0, 55.0, 70.0, 85.0]
C2_FREQ = list(86e9 * (690.0 / 86.0) ** (np.arange(8) / 7.0))


@pytest.fixture(scope="module")
def tracer(dump128):
    tr = rt.Tracer(0)
    tr.load_dump(dump128)
    yield tr
    tr.close()


def _params(tr, W, H=None):
    tr.set_params(width=W, height=W if H is None else H, M_unit=M_UNIT, R_high=R_HIGH)


@pytest.mark.parametrize("cfg", [dict(tile=4), dict(tile=8), dict(tile=16),
                                 dict(refill=64), dict(refill=17, tile=5),
                                 dict(blocks_per_cu=1), dict(blocks_per_cu=12, tile=3),
                                 dict(tile=1), dict(order=rt.ORDER_SCAN, tile=1),
                                 dict(order=rt.ORDER_SCAN), dict(order=rt.ORDER_SCAN, tile=8),
                                 dict(order=rt.ORDER_SCAN, tile=5, refill=9),
                                 dict(order=rt.ORDER_CENTRE, tile=64),
                                 dict(approach=rt.APPROACH_NONE),
                                 dict(approach=rt.APPROACH_NONE, order=rt.ORDER_SCAN, tile=5),
                                 dict(approach=100, refill=1, tile=1),
                                 dict(approach=1, order=rt.ORDER_SCAN, tile=5, refill=9),
                                 dict(approach=73, tile=3)])
def test_launch_shapes_are_bit_identical(tracer, cfg):
    _params(tracer, 45, 37)                      # ragged against every tile size
    tracer.set_launch_config()
    ref, _ = tracer.render([INCL], [FREQ])
    tracer.set_launch_config(**cfg)
    try:
        got, _ = tracer.render([INCL], [FREQ])
    finally:
        tracer.set_launch_config()
    assert np.array_equal(got, ref)


def test_bad_launch_config_rejected(tracer):
    with pytest.raises(rt.RtError):
        tracer.set_launch_config(refill=65)
    with pytest.raises(rt.RtError):
        tracer.set_launch_config(tile=-1)
    with pytest.raises(rt.RtError):
        tracer.set_launch_config(order=4)
    with pytest.raises(rt.RtError):
        tracer.set_launch_config(approach=101)
    with pytest.raises(rt.RtError):
        tracer.set_launch_config(approach=-2)
    cfg = tracer.get_launch_config()
    assert cfg["order"] == rt.ORDER_RING and cfg["approach"] == 0


def test_split_equals_fused_with_gradients_and_counters(tracer):
    """Launches with an approach kernel (all, some or none of the rays saved
    where they reach the grid and resumed by the trace kernel) give the same
    image, gradient and per-ray counters, for one and several frequencies."""
    W = 48
    _params(tracer, W)
    n = 2 * W * W
    d = tracer.alloc(n * 3 * 8)
    dd = tracer.alloc(n * 3 * 16)
    dc = tracer.alloc(n * 8)
    try:
        for freqs in ([FREQ], [FREQ, 86e9, 345e9]):
            out = []
            for approach in (rt.APPROACH_NONE, 100, 50, 7):
                tracer.set_launch_config(approach=approach)
                tracer.render_device([INCL, 20.0], freqs, d, dd, dc)
                tracer.synchronize()
                nf = len(freqs)
                out.append((tracer.d2h(d, (n, nf), np.float64),
                            tracer.d2h(dd, (n, nf, 2), np.float64),
                            tracer.d2h(dc, (n, 2), np.int32)))
            for o in out[1:]:
                for a, b in zip(out[0], o):
                    assert np.array_equal(a, b), len(freqs)
            assert out[0][0].max() > 0 and (out[0][2][:, 1] > 0).any()
    finally:
        tracer.set_launch_config()
        for p in (d, dd, dc):
            tracer.free(p)


@pytest.mark.parametrize("grad", [False, True])
def test_split_equals_fused_where_rays_are_traced_again(tracer, grad):
    """One frequency at 256^2, where a few tens of rays meet the gate window:
    the split launch lists them for exact_kernel, the one-kernel launch
    traces them again in its own lanes; image, tangents and per-ray
    counters agree bit for bit."""
    W = 256
    _params(tracer, W)
    n = W * W
    d, dc = tracer.alloc(n * 8), tracer.alloc(n * 8)
    dd = tracer.alloc(n * 16) if grad else None
    try:
        out = []
        for approach in (rt.APPROACH_NONE, 100):
            tracer.set_launch_config(approach=approach)
            tracer.render_device([INCL], [FREQ], d, dd, dc)
            tracer.synchronize()
            out.append([tracer.d2h(d, (n,), np.float64), tracer.d2h(dc, (n, 2), np.int32)] +
                       ([tracer.d2h(dd, (n, 2), np.float64)] if grad else []))
        for a, b in zip(out[0], out[1]):
            assert np.array_equal(a, b)
        assert out[0][0].max() > 0
    finally:
        tracer.set_launch_config()
        for p in (d, dc, dd):
            if p is not None:
                tracer.free(p)


def test_split_capture_chunks_and_pins_its_ray_buffer(dump128):
    """The buffer of saved rays is sized by eager renders.  A capture larger
    than the buffer runs in chunks that fit it; a later eager render that
    grows the buffer leaves the graph's one in place until the graph is
    destroyed; a capture before any eager render is refused."""
    tr = rt.Tracer(0)
    try:
        tr.load_dump(dump128)
        W = 40
        tr.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH)
        cams = [INCL, 30.0, 75.0]
        n = len(cams) * W * W
        d = tr.alloc(n * 8)
        try:
            tr.set_launch_config(order=rt.ORDER_SCAN, approach=50)
            tr.capture_begin()
            try:
                with pytest.raises(rt.RtError, match="eager"):
                    tr.render_device([INCL], [FREQ], d)
            finally:
                tr.capture_end().destroy()
            tr.set_launch_config(approach=50)
            one, _ = tr.render([INCL], [FREQ])      # buffer: half a camera's rays
            tr.capture_begin()
            tr.render_device(cams, [FREQ], d)       # chunks of at most one camera
            g = tr.capture_end()
            try:
                three, _ = tr.render(cams, [FREQ])  # grows the buffer
                assert np.array_equal(three[0], one[0])
                for _ in range(2):
                    tr.render_device([5.0, 15.0, 85.0], [FREQ], d)   # other values in d
                    tr.synchronize()
                    g.launch()
                    tr.synchronize()
                    got = tr.d2h(d, (n,), np.float64).reshape(len(cams), W * W)
                    assert np.array_equal(got, three[:, :, 0])
            finally:
                g.destroy()
            again, _ = tr.render(cams, [FREQ])
            assert np.array_equal(again, three)
        finally:
            tr.free(d)
    finally:
        tr.close()


def test_centre_order_covers_every_pixel_once(tracer):
    """Per-ray counters in centre order equal those in scan order on ragged
    shapes (a pixel never traced would keep what the buffer held before: the
    shapes alternate so that this is another image's data)."""
    shapes = [(45, 37, 1), (7, 64, 5), (45, 37, 8), (64, 7, 3), (1, 1, 1)]
    nmax = max(W * H for W, H, _ in shapes)
    d = tracer.alloc(nmax * 8)
    dc_scan = tracer.alloc(nmax * 8)
    dc_centre = tracer.alloc(nmax * 8)
    try:
        for W, H, tile in shapes:
            _params(tracer, W, H)
            n = W * H
            got = []
            for order, dc in ((rt.ORDER_SCAN, dc_scan), (rt.ORDER_CENTRE, dc_centre)):
                tracer.set_launch_config(tile=tile, order=order)
                tracer.render_device([INCL], [FREQ], d, None, dc)
                tracer.synchronize()
                got.append(tracer.d2h(dc, (n, 2), np.int32))
            assert (got[0][:, 0] > 0).all(), (W, H, tile)
            assert np.array_equal(got[0], got[1]), (W, H, tile)
    finally:
        tracer.set_launch_config()
        for p in (d, dc_scan, dc_centre):
            tracer.free(p)


@pytest.mark.parametrize("nf", [1, 3, 8])
def test_three_wave_build_equals_two_wave_build(tracer, nf):
    """Launches of at least 2^19 rays run the kernel built for 3 waves per
    SIMD (up to 8 frequencies), smaller ones the 2-wave build; the same rays
    give the same bits, also when the approach kernel hands them over."""
    freqs = [FREQ] if nf == 1 else C2_FREQ[:nf]
    _params(tracer, 512)
    small, _ = tracer.render([INCL], freqs)          # 262144 rays: 2-wave build
    big, _ = tracer.render([INCL, 30.0], freqs)      # 524288 rays: 3-wave build
    assert np.array_equal(big[0], small[0])
    tracer.set_launch_config(approach=100)
    try:
        split, _ = tracer.render([INCL, 30.0], freqs)
    finally:
        tracer.set_launch_config()
    assert np.array_equal(split, big)
    assert big[0].max() > 0


def test_c2_batched_equals_per_image(tracer):
    """Six inclinations x eight log-spaced frequencies (86-690 GHz) in one
    launch equal 48 single-image launches."""
    _params(tracer, 32)
    batched, spec = tracer.render(C2_INCL, C2_FREQ)
    for c, inc in enumerate(C2_INCL):
        for f, nu in enumerate(C2_FREQ):
            one, s1 = tracer.render([inc], [nu])
            np.testing.assert_array_equal(batched[c, :, f], one[0, :, 0])
            assert spec[c, f] == s1[0, 0]


@pytest.mark.parametrize("dim3,grad,integrator,approach", [
    (False, False, rt.INT_RK2, 100), (False, True, rt.INT_RK2, None),
    (True, True, rt.INT_RK2, 100), (False, False, rt.INT_RK4, None),
    (True, False, rt.INT_VERLET, 100)])
def test_batched_frequencies_equal_single_launches(tracer, dump3d_small, dump128, dim3, grad,
                                                   integrator, approach):
    """Three frequencies in one launch equal three one-frequency launches,
    image and tangents, bit for bit: the rays whose FP32 optical depth met
    the gate where the decision shows are traced again with the FP64
    transfer inline in a one-frequency launch, and by exact_kernel after a
    batched one (a lane per (ray, frequency), or a wave per ray with one
    frequency in the gate window).  256^2 rays (on the 2D dump the FP64
    oracle's optical depth crosses the gate within 1e-3 of ln 1000 on 33-36
    rays per frequency), 2D and 3D dumps, the three integrators, the split
    launch (approach kernel, resumed rays) and the one-kernel launch."""
    freqs = [1.3e11, 2.3e11, 3.45e11]
    W = 256
    tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH, integrator=integrator)
    tracer.load_dump(dump3d_small if dim3 else dump128)
    try:
        if approach is not None:
            tracer.set_launch_config(approach=approach)
        out = tracer.render([INCL], freqs, grad=grad)
        for f, nu in enumerate(freqs):
            one = tracer.render([INCL], [nu], grad=grad)
            np.testing.assert_array_equal(out[0][0, :, f], one[0][0, :, 0])
            if grad:
                np.testing.assert_array_equal(out[1][0, :, f], one[1][0, :, 0])
        assert out[0].max() > 0
    finally:
        tracer.set_launch_config()
        tracer.set_params(width=W, height=W, M_unit=M_UNIT, R_high=R_HIGH)
        tracer.load_dump(dump128)


def test_graph_replay_matches_eager(tracer):
    """A captured graph of two renders, replayed several times with eager
    work in between; every replay equals the eager result."""
    W = 40
    _params(tracer, W)
    n = W * W
    d1 = tracer.alloc(n * 8)
    d2 = tracer.alloc(n * 3 * 8)
    dx = tracer.alloc(n * 8)
    try:
        tracer.render_device([INCL], [FREQ], d1)
        tracer.render_device([30.0], C2_FREQ[:3], d2)
        tracer.synchronize()
        e1 = tracer.d2h(d1, (n,), np.float64)
        e2 = tracer.d2h(d2, (n, 3), np.float64)
        tracer.capture_begin()
        tracer.render_device([INCL], [FREQ], d1)
        tracer.render_device([30.0], C2_FREQ[:3], d2)
        g = tracer.capture_end()
        try:
            for k in range(4):
                # eager work that touches the same queue counter and stream
                tracer.render_device([80.0], [FREQ], dx)
                tracer.synchronize()
                tracer.d2h(dx, (n,), np.float64)
                g.launch()
                tracer.synchronize()
                assert np.array_equal(tracer.d2h(d1, (n,), np.float64), e1), k
                assert np.array_equal(tracer.d2h(d2, (n, 3), np.float64), e2), k
        finally:
            g.destroy()
    finally:
        for d in (d1, d2, dx):
            tracer.free(d)


def test_ring_order_lead_launch_is_bit_identical(tracer):
    """The ring order traces the pixels near the critical curve from the camera
    in a lead kernel on a second stream while the split launch takes the
    rest: same image, gradient and per-ray counters as the centre order,
    at 1024^2 (the default split launch) and two inclinations (the curve
    moves with the inclination), and the lead kernel ran.  A new context
    starts with the ring order."""
    with rt.Tracer(0) as fresh:
        assert fresh.get_launch_config()["order"] == rt.ORDER_RING
    W = 1024
    _params(tracer, W)
    n = W * W
    d = tracer.alloc(n * 8)
    dd = tracer.alloc(n * 16)
    dc = tracer.alloc(n * 8)
    try:
        for incl in (INCL, 17.0):
            out = {}
            for order in (rt.ORDER_CENTRE, rt.ORDER_RING):
                tracer.set_launch_config(order=order)
                tracer.render_device([incl], [FREQ], d, dd, dc)
                tracer.synchronize()
                lead_ms = tracer.last_kernel_times()["lead"]
                out[order] = (tracer.d2h(d, (n,), np.float64), tracer.d2h(dd, (n, 2), np.float64),
                              tracer.d2h(dc, (n, 2), np.int32), lead_ms)
            for k in range(3):
                assert np.array_equal(out[rt.ORDER_CENTRE][k], out[rt.ORDER_RING][k]), (incl, k)
            assert out[rt.ORDER_CENTRE][3] == 0.0 and out[rt.ORDER_RING][3] > 0.0
    finally:
        tracer.set_launch_config()
        for p in (d, dd, dc):
            tracer.free(p)


def test_ring_order_lead_launch_non_square_image(tracer):
    """The lead set of a 2048 x 512 image (2^20 rays, the split launch; the
    ring and polar band in camera units of non-square pixel counts) leaves
    the image bit-identical to the centre order's, at 60 and 150 deg."""
    W, H = 2048, 512
    _params(tracer, W, H)
    n = W * H
    d = tracer.alloc(n * 8)
    try:
        for incl in (INCL, 150.0):
            out = {}
            for order in (rt.ORDER_CENTRE, rt.ORDER_RING):
                tracer.set_launch_config(order=order)
                tracer.render_device([incl], [FREQ], d)
                tracer.synchronize()
                out[order] = (tracer.d2h(d, (n,), np.float64), tracer.last_kernel_times()["lead"])
            assert np.array_equal(out[rt.ORDER_CENTRE][0], out[rt.ORDER_RING][0]), incl
            assert out[rt.ORDER_RING][1] > 0.0 and tracer.claim_order(incl)[1] > 1000
    finally:
        tracer.set_launch_config()
        tracer.free(d)


def _critical_curve(a, incl_deg, n=200001, ds=0.002):
    """Bardeen's critical curve (impact parameters alpha, beta of the
    spherical photon orbits) for a camera at incl_deg around spin a, with
    this code's camera convention alpha = -xi / sin i, as points every ds
    (M) of arc length (even spacing keeps the k-d tree below fast)."""
    if a == 0.0:   # Schwarzschild: the circle of radius sqrt(27)
        ph = np.linspace(0.0, 2 * np.pi, n)
        A, B = np.sqrt(27.0) * np.cos(ph), np.sqrt(27.0) * np.sin(ph)
    else:
        inc = np.radians(incl_deg)
        r1 = 2 * (1 + np.cos(2 / 3 * np.arccos(-a)))
        r2 = 2 * (1 + np.cos(2 / 3 * np.arccos(a)))
        r = np.linspace(min(r1, r2), max(r1, r2), n)
        xi = (r**2 * (3 - r) - a**2 * (r + 1)) / (a * (r - 1))
        eta = r**3 * (4 * a**2 - r * (r - 3)**2) / (a**2 * (r - 1)**2)
        b2 = eta + a**2 * np.cos(inc)**2 - xi**2 / np.tan(inc)**2
        ok = b2 >= 0
        al, be = -xi[ok] / np.sin(inc), np.sqrt(b2[ok])
        A, B = np.r_[al, al[::-1]], np.r_[be, -be[::-1]]   # upper branch, lower reversed
    s = np.r_[0.0, np.cumsum(np.hypot(np.diff(A), np.diff(B)))]
    u = np.arange(0.0, s[-1], ds)
    return np.interp(u, s, A), np.interp(u, s, B)


def _polar_sin_min(a, incl_deg, al, be):
    """sin of the polar angle at which each ray (camera constants of motion
    E = 1, L = -alpha sin i, Q = beta^2 + cos^2 i (alpha^2 - a^2)) turns in
    theta, by bisection of the theta potential Q + a^2 cos^2 - L^2 cot^2 on
    (0, pi/2], where it is positive at pi/2 (Q > 0) and falls to -inf at 0;
    inf where Q <= 0 (the ray never leaves the equatorial band)."""
    inc = np.radians(incl_deg)
    L = -al * np.sin(inc)
    Q = be**2 + np.cos(inc)**2 * (al**2 - a**2)
    lo, hi = np.full(al.shape, 1e-12), np.full(al.shape, np.pi / 2)
    for _ in range(60):
        th = 0.5 * (lo + hi)
        pot = Q + a**2 * np.cos(th)**2 - L**2 / np.tan(th)**2
        neg = pot < 0
        lo, hi = np.where(neg, th, lo), np.where(neg, hi, th)
    return np.where(Q > 0, np.sin(hi), np.inf)


def _check_lead(lead, a, incl, al, be, ring_width=0.3, polar=0.02, share=0.05):
    """The lead set against its definition: the pixels within ring_width of
    the critical curve (a margin of 0.01 M for the library's sampling of the
    curve) and those turning in theta with sin theta_min < polar (a relative
    margin of 1e-6 for rounding), at most share of the image: beyond that
    the band keeps its pixels of smallest sin theta_min.  Returns the number
    of pixels in the band."""
    from scipy.spatial import cKDTree
    ca, cb = _critical_curve(a, incl)
    d, _ = cKDTree(np.c_[ca, cb]).query(np.c_[al, be], distance_upper_bound=0.4)
    sp = _polar_sin_min(a, incl, al, be)
    ring_in, ring_out = d < ring_width - 0.01, d > ring_width + 0.01
    assert np.all(lead[ring_in]), (a, incl)
    assert lead.sum() <= int(share * lead.size), (a, incl)
    # outside the ring, the lead pixels are band pixels, each of them closer
    # to the axis than every band pixel left out
    assert np.all(sp[lead & ring_out] < polar * (1 + 1e-6)), (a, incl)
    band = (sp < polar * (1 - 1e-6)) & ring_out
    if lead.sum() < int(share * lead.size):
        assert np.all(lead[band]), (a, incl)
    else:
        kept, left = sp[lead & ring_out], sp[band & ~lead]
        assert left.size == 0 or kept.max() <= left.min() * (1 + 1e-6), (a, incl)
    return int((sp < polar).sum())


def test_ring_order_leads_with_the_longest_rays(tracer, dump128):
    """rt_claim_order: the ring order is a permutation of the pixels whose first
    n_lead are those within 0.3 M of the critical curve (here from an
    independent dense sampling, with a margin for the library's sampling)
    or whose rays turn in theta within arcsin 0.02 of the spin axis (here
    by bisection of the theta potential), and the rest the centre order
    without them; the centre order has no lead; images below the split
    launch's size have none either.  The curve moves with the inclination,
    and its pixels are where the longest rays are (the photon ring); the
    polar band is where the step-size rule shrinks the steps.  At 1 deg the
    band would take nearly the whole image and is cut to its pixels closest
    to the axis (5% of the image in all)."""
    W = 1024
    _params(tracer, W)
    fov = 40.0
    yy, xx = np.divmod(np.arange(W * W), W)
    al = -fov / 2 + (xx + 0.5) * fov / W
    be = -fov / 2 + (yy + 0.5) * fov / W
    try:
        for incl in (INCL, 17.0, 150.0, 1.0):
            ring, nl = tracer.claim_order(incl)
            assert np.array_equal(np.sort(ring), np.arange(W * W))
            lead = np.zeros(W * W, bool)
            lead[ring[:nl]] = True
            assert _check_lead(lead, dump128.grid.a, incl, al, be) > 2000
            assert nl > 5000
            tracer.set_launch_config(order=rt.ORDER_CENTRE)
            centre, nc = tracer.claim_order(incl)
            tracer.set_launch_config()
            assert nc == 0
            assert np.array_equal(ring[nl:], centre[~lead[centre]])
            assert np.array_equal(ring[:nl], centre[lead[centre]])
        _params(tracer, 512)
        assert tracer.claim_order(INCL)[1] == 0
        # other spins, Schwarzschild's circle included (the ring follows the
        # loaded dump's spin)
        _params(tracer, W)
        for a in (0.0, -0.7, 0.5):
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                tracer.load_dump(torus.fishbone_moncrief(32, 32, a=a))
            ring, nl = tracer.claim_order(INCL)
            lead = np.zeros(W * W, bool)
            lead[ring[:nl]] = True
            _check_lead(lead, a, INCL, al, be)
            assert np.array_equal(np.sort(ring), np.arange(W * W)) and nl > 5000
    finally:
        tracer.set_launch_config()
        tracer.load_dump(dump128)


def test_graph_replays_the_ring_order_lead_launch(tracer):
    """A graph captured around a ring-order render forks the lead kernel onto
    the second stream and joins it back: every replay, with eager renders in
    between, equals the eager image."""
    W = 1024
    _params(tracer, W)
    n = W * W
    d = tracer.alloc(n * 8)
    dx = tracer.alloc(n * 8)
    try:
        tracer.render_device([INCL], [FREQ], d)
        tracer.synchronize()
        e = tracer.d2h(d, (n,), np.float64)
        tracer.capture_begin()
        tracer.render_device([INCL], [FREQ], d)
        g = tracer.capture_end()
        try:
            for k in range(3):
                tracer.render_device([80.0], [FREQ], dx)
                tracer.synchronize()
                tracer.d2h(dx, (n,), np.float64)
                g.launch()
                tracer.synchronize()
                assert np.array_equal(tracer.d2h(d, (n,), np.float64), e), k
        finally:
            g.destroy()
    finally:
        tracer.free(d)
        tracer.free(dx)


def test_graph_keeps_its_claim_order_after_other_shapes(tracer):
    """A graph captured at one image shape replays correctly after renders at
    other shapes have built new claim-order tables -- more shapes than the
    context keeps, so unpinned tables are evicted around the graph's pinned
    one, and evicted shapes render identically when their table is rebuilt;
    a shape first seen during a capture is refused."""
    W = 24
    _params(tracer, W)
    n = W * W
    d1 = tracer.alloc(n * 8)
    try:
        tracer.render_device([INCL], [FREQ], d1)
        tracer.synchronize()
        e1 = tracer.d2h(d1, (n,), np.float64)
        tracer.capture_begin()
        tracer.render_device([INCL], [FREQ], d1)
        g = tracer.capture_end()
        try:
            first = {}
            for W2 in [31, 64] + list(range(10, 22)) + [31, 64, 10]:
                _params(tracer, W2)
                img, _ = tracer.render([INCL], [FREQ])
                if W2 in first:
                    assert np.array_equal(img, first[W2]), W2
                first.setdefault(W2, img)
                g.launch()
                tracer.synchronize()
                assert np.array_equal(tracer.d2h(d1, (n,), np.float64), e1), W2
            _params(tracer, 29)
            tracer.capture_begin()
            try:
                with pytest.raises(rt.RtError, match="eager"):
                    tracer.render_device([INCL], [FREQ], d1)
            finally:
                tracer.capture_end().destroy()
        finally:
            g.destroy()
    finally:
        tracer.free(d1)


def test_graph_keeps_its_primitives_after_a_new_dump(tracer, dump128):
    """rt_load_prims while graphs still launch with the old primitives: the
    graphs go on rendering the dump they were captured with (the buffer is
    kept until the last of them is destroyed), eager renders use the new
    dump, and a graph captured after the load renders the new one."""
    import dataclasses
    W = 32
    _params(tracer, W)
    n = W * W
    other = dataclasses.replace(dump128, prims=dump128.prims.copy())
    other.prims[0] *= 1.5   # denser: a different image
    d1, d2 = tracer.alloc(n * 8), tracer.alloc(n * 8)
    graphs = []
    try:
        tracer.load_dump(dump128)
        e_old, _ = tracer.render([INCL], [FREQ])
        for _ in range(2):   # two graphs on the old buffer
            tracer.capture_begin()
            tracer.render_device([INCL], [FREQ], d1)
            graphs.append(tracer.capture_end())
        tracer.load_dump(other)
        e_new, _ = tracer.render([INCL], [FREQ])
        assert not np.array_equal(e_old, e_new)
        tracer.capture_begin()
        tracer.render_device([INCL], [FREQ], d2)
        graphs.append(tracer.capture_end())
        for k, g in enumerate(graphs):
            g.launch()
            tracer.synchronize()
            want = e_new if k == 2 else e_old
            got = tracer.d2h(d2 if k == 2 else d1, (n,), np.float64)
            assert np.array_equal(got, want[0, :, 0]), k
        graphs.pop(0).destroy()   # one graph still holds the old buffer
        graphs[0].launch()
        tracer.synchronize()
        assert np.array_equal(tracer.d2h(d1, (n,), np.float64), e_old[0, :, 0])
        tracer.load_dump(dump128)   # the old dump again, in a new buffer
        img, _ = tracer.render([INCL], [FREQ])
        assert np.array_equal(img, e_old)
        graphs[1].launch()
        tracer.synchronize()
        assert np.array_equal(tracer.d2h(d2, (n,), np.float64), e_new[0, :, 0])
    finally:
        for g in graphs:
            g.destroy()
        tracer.free(d1)
        tracer.free(d2)
        tracer.load_dump(dump128)


def test_calls_that_synchronise_are_refused_during_capture(tracer):
    _params(tracer, 8)
    tracer.capture_begin()
    try:
        with pytest.raises(rt.RtError, match="capturing"):
            tracer.render([INCL], [FREQ])
        with pytest.raises(rt.RtError, match="capturing"):
            tracer.synchronize()
    finally:
        tracer.capture_end().destroy()


def test_autotune_picks_a_valid_config_and_caches_it(tracer, tmp_path):
    _params(tracer, 48)
    path = str(tmp_path / "tune.json")
    res = autotune.tune(tracer, [INCL], [FREQ], space=dict(blocks_per_cu=(0, 4),
                                                           refill=(1,), tile=(1, 8),
                                                           order=(rt.ORDER_CENTRE,),
                                                           approach=(100,)),
                        reps=1, path=path)
    assert len(res["table"]) == 4 and res["ms"] > 0
    # the leading shapes timed again; the result is the fastest of those
    assert len(res["refined"]) == 4
    assert res["ms"] == min(ms for _, ms in res["refined"])
    assert tracer.get_launch_config()["tile"] == res["config"]["tile"]
    tracer.set_launch_config()
    assert autotune.apply_cached(tracer, 1, 1, path=path) == res["config"]
    img, _ = tracer.render([INCL], [FREQ])
    tracer.set_launch_config()
    ref, _ = tracer.render([INCL], [FREQ])
    assert np.array_equal(img, ref)
