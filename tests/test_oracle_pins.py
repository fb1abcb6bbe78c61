"""Pins for the FP64 CPU oracle (oracle/raptor_oracle.c).

The reference ships no tests, fixtures or golden vectors, and its path cannot
be built here (it needs GSL headers the image lacks), so every oracle function
is pinned by something that does not come from the oracle: symbolic or
closed-form values (tests/golden, made by tests/golden/make_golden.py),
conservation laws, textbook limits and brute-force sums.  Each test names the
mistake it would catch.
"""
import json
import math
import os
import warnings

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# constants.h:16-28 of the reference
E = 4.8032068e-10
ME = 9.1093826e-28
MP = 1.67262171e-24
KB = 1.3806505e-16
C = 2.99792458e10
H = 6.6260693e-27
GG = 6.6742e-8


def _grid(a=0.9375, h=0.3, n1=64, n2=64):
    from raptor_amd.torus import make_grid
    return make_grid(n1, n2, 1, a=a, hslope=h)


def _P(a=0.9375, h=0.3, incl=60.0, **kw):
    return O.params(_grid(a, h), width=kw.pop("width", 16), height=kw.pop("height", 16),
                    inclination=incl, M_unit=kw.pop("M_unit", 1e23),
                    R_high=kw.pop("R_high", 10.0), **kw)


# ---------------------------------------------------------------- metric


@pytest.fixture(scope="module")
def kerr_gold():
    return json.load(open(os.path.join(GOLD, "kerr_mks.json")))


def test_metric_matches_kerr_schild_pullback(kerr_gold):
    """g_dd and g^uu against the Cartesian Kerr-Schild form pulled back to MKS
    (catches a wrong Jacobian factor, sign of a, or a dropped term)."""
    for case in kerr_gold["cases"]:
        P = _P(case["a"], case["hslope"])
        for p, gd, gu in zip(kerr_gold["points"], case["gdd"], case["guu"]):
            X = np.array(p)
            gdd = O.metric_dd(P, X)
            guu = O.metric_uu(P, X)
            np.testing.assert_allclose(gdd, np.reshape(gd, (4, 4)), rtol=1e-12,
                                       atol=1e-12 * np.abs(gd).max())
            np.testing.assert_allclose(guu, np.reshape(gu, (4, 4)), rtol=1e-11,
                                       atol=1e-12 * np.abs(gu).max())


def test_metric_inverse_identity():
    rng = np.random.default_rng(0)
    P = _P()
    for _ in range(50):
        X = np.array([0, rng.uniform(np.log(1.4), np.log(1e4)), rng.uniform(0.01, 0.99), 1.0])
        np.testing.assert_allclose(O.metric_dd(P, X) @ O.metric_uu(P, X), np.eye(4),
                                   atol=1e-11)


def test_connection_matches_exact_derivatives(kerr_gold):
    """Gamma^m_ab against 1/2 g^md (d g) from exact symbolic derivatives
    (catches any wrong entry, sign or symmetric copy)."""
    for case in kerr_gold["cases"]:
        P = _P(case["a"], case["hslope"])
        for p, gam in zip(kerr_gold["points"], case["gamma"]):
            G = O.connection(P, np.array(p))
            ref = np.reshape(gam, (4, 4, 4))
            scale = np.abs(ref).max()
            np.testing.assert_allclose(G, ref, rtol=1e-9, atol=1e-11 * scale)


def test_connection_symmetric_lower_indices():
    P = _P()
    G = O.connection(P, np.array([0, 1.3, 0.37, 0.0]))
    np.testing.assert_array_equal(G, np.transpose(G, (0, 2, 1)))


# ---------------------------------------------------------------- camera


def _lower(P, y):
    return O.metric_dd(P, y[:4]) @ y[4:]


@pytest.mark.parametrize("incl", [10.0, 60.0, 85.0, 120.0])
def test_camera_constants_of_motion(incl):
    """At the camera k_t = -1 (E = 1), k_phi = -alpha sin i, k_theta = beta and
    g(k,k) = 0 (Cunningham & Bardeen 1973; catches swapped alpha/beta or a
    wrong BL->KS transform, which would change k_t or k_phi)."""
    P = _P(incl=incl)
    si = math.sin(math.radians(incl))
    h = P.hslope
    for alpha, beta in [(3.0, 4.0), (-7.5, 1.25), (0.3, -11.0), (15.0, 0.0)]:
        y = O.init_photon(P, alpha, beta)
        kd = _lower(P, y)
        thp = math.pi * (1 + (1 - h) * math.cos(2 * math.pi * y[2]))
        assert kd[0] == pytest.approx(-1.0, rel=1e-10)
        assert kd[3] == pytest.approx(-alpha * si, rel=1e-9, abs=1e-9)
        assert kd[2] / thp == pytest.approx(beta, rel=1e-8, abs=1e-8)
        knorm = y[4:] @ O.metric_dd(P, y[:4]) @ y[4:]
        assert abs(knorm) < 1e-10 * y[4] ** 2
        assert y[5] > 0  # travelling inwards for negative d lambda


def test_camera_theta_inversion():
    """theta(X2) = theta_cam to rounding (newtonraphson.c)."""
    lib = O.lib()
    for h in (0.3, 0.7, 1.0):
        for th in np.linspace(0.01, math.pi - 0.01, 25):
            x = lib.or_theta_to_x2(th, h)
            assert math.pi * x + 0.5 * (1 - h) * math.sin(2 * math.pi * x) == pytest.approx(
                th, abs=2e-15)


# ---------------------------------------------------------------- geodesics


def _carter(P, y):
    """Carter constant from the covariant momentum in KS/MKS."""
    h = P.hslope
    kd = _lower(P, y)
    x2 = y[2]
    th = math.pi * x2 + 0.5 * (1 - h) * math.sin(2 * math.pi * x2)
    thp = math.pi * (1 + (1 - h) * math.cos(2 * math.pi * x2))
    kth = kd[2] / thp
    a = P.a
    return kth ** 2 + math.cos(th) ** 2 * (-a * a * kd[0] ** 2 + kd[3] ** 2 / math.sin(th) ** 2)


def _drift(P, alpha, beta):
    path = O.ray_path(P, alpha, beta)
    E0, L0, Q0 = _lower(P, path[0])[0], _lower(P, path[0])[3], _carter(P, path[0])
    kd = np.array([_lower(P, y) for y in path])
    Q = np.array([_carter(P, y) for y in path])
    return (np.abs(kd[:, 0] - E0).max(), np.abs(kd[:, 3] - L0).max(),
            np.abs(Q - Q0).max() / abs(Q0), len(path))


def test_conserved_quantities_and_second_order_convergence():
    """E = -k_t, L = k_phi and Carter's Q are constants of Kerr geodesics; the
    RK2 midpoint error must fall ~4x when the step scale halves (an Euler or
    misweighted step would fall only ~2x)."""
    alpha, beta = 4.5, 3.0
    P1 = _P(stepsize=0.03)
    P2 = _P(stepsize=0.015)
    e1, l1, q1, n1 = _drift(P1, alpha, beta)
    e2, l2, q2, n2 = _drift(P2, alpha, beta)
    L0 = 4.5 * math.sin(math.radians(60.0))
    assert e1 < 1e-2 and l1 < 0.03 * L0 and q1 < 0.05
    assert n2 > 1.8 * n1
    for d1, d2 in ((e1, e2), (l1, l2), (q1, q2)):
        ratio = d1 / d2
        assert 3.5 < ratio < 4.5, (d1, d2)


def test_rk4_fourth_order_convergence():
    """integrator.c:13-39 RK4: the conserved-quantity drift falls ~16x per
    halving of the step scale (a wrong weight or a dropped stage would show
    second or third order)."""
    d1 = _drift(_P(stepsize=0.015, integrator=O.RK4), 4.5, 3.0)
    d2 = _drift(_P(stepsize=0.0075, integrator=O.RK4), 4.5, 3.0)
    assert d1[0] < 1e-7
    for a, b in zip(d1[:3], d2[:3]):
        assert 14.0 < a / b < 19.0, (a, b)


def test_verlet_second_order_convergence():
    """integrator.c:71-105 velocity Verlet: second order (drift ~4x per
    halving), and a different error than RK2's on the same ray."""
    P1 = _P(stepsize=0.03, integrator=O.VERLET)
    d1 = _drift(P1, 4.5, 3.0)
    d2 = _drift(_P(stepsize=0.015, integrator=O.VERLET), 4.5, 3.0)
    for a, b in zip(d1[:3], d2[:3]):
        assert 3.5 < a / b < 4.5, (a, b)
    r1 = _drift(_P(stepsize=0.03), 4.5, 3.0)
    assert abs(d1[0] / r1[0] - 1) > 0.5


def test_null_condition_along_ray():
    P = _P()
    path = O.ray_path(P, 2.0, -6.0)
    for y in path[:: max(1, len(path) // 40)]:
        g = O.metric_dd(P, y[:4])
        k = y[4:]
        scale = np.abs(np.outer(k, k) * g).sum()
        assert abs(k @ g @ k) < 2e-3 * scale


def _captured(P, alpha, beta):
    path = O.ray_path(P, alpha, beta)
    return math.exp(path[-1][1]) < 2.0 * (1 + math.sqrt(1 - P.a ** 2))


def test_schwarzschild_capture_threshold():
    """a = 0: rays with impact parameter below sqrt(27) are captured."""
    P = _P(a=0.0, h=1.0)
    bc = math.sqrt(27.0)
    for ang in np.linspace(0, 2 * math.pi, 7)[:-1]:
        for b, cap in ((0.96 * bc, True), (1.04 * bc, False)):
            assert _captured(P, b * math.cos(ang), b * math.sin(ang)) == cap, (b, ang)


def test_kerr_shadow_boundary():
    """a = 0.9375, i = 60 deg: points 5% inside / outside Bardeen's critical
    curve are captured / escape (catches a wrong spin sign or frame drag)."""
    a, i = 0.9375, math.radians(60.0)
    P = _P(a=a, incl=60.0)
    rp1 = 2 * (1 + math.cos(2 / 3 * math.acos(-a)))
    rp2 = 2 * (1 + math.cos(2 / 3 * math.acos(a)))
    n_checked = 0
    for rr in np.linspace(rp1 + 1e-3, rp2 - 1e-3, 9):
        xi = (rr * rr * (3 - rr) - a * a * (rr + 1)) / (a * (rr - 1))
        eta = rr ** 3 * (4 * a * a - rr * (rr - 3) ** 2) / (a * a * (rr - 1) ** 2)
        bsq = eta + a * a * math.cos(i) ** 2 - xi * xi / math.tan(i) ** 2
        if bsq <= 0:
            continue
        alpha, beta = -xi / math.sin(i), math.sqrt(bsq)
        assert _captured(P, 0.95 * alpha, 0.95 * beta)
        assert not _captured(P, 1.05 * alpha, 1.05 * beta)
        n_checked += 1
    assert n_checked >= 5


def test_stepsize_formula_limits():
    """Dolence-style step: with only k^1 nonzero, dl = -eps/|k^1|; at the pole
    (X2 = 0) the step collapses to the 1e-12 floor."""
    P = _P()
    assert O.stepsize(P, [0, 1, 0.5, 0], [0, 2.0, 0, 0]) == pytest.approx(-0.03 / 2.0,
                                                                          rel=1e-12)
    assert O.stepsize(P, [0, 1, 0.0, 0], [0, 2.0, 1.0, 0]) == -1e-12


# ---------------------------------------------------------------- fluid


def _linear_prims(grid, coef):
    i = np.arange(grid.n1)[:, None, None]
    j = np.arange(grid.n2)[None, :, None]
    p = np.zeros((8, grid.n1, grid.n2, 1))
    for v in range(8):
        p[v] = coef[v][0] + coef[v][1] * i + coef[v][2] * j
    return p


def test_bilinear_exact_on_linear_fields_including_edge_cells():
    """Interpolation is exact on fields linear in the cell index, including
    the linear extrapolation over the first half cell (C truncation in Xtoij)
    and the constant clamp over the last half cell (grmhd.c:46-71)."""
    g = _grid(n1=16, n2=12)
    P = O.params(g, width=4, height=4, inclination=60.0, M_unit=1e23, R_high=10.0)
    coef = [(1.0, 0.01, 0.02), (0.1, 0.001, 0.003), (0, 0, 0), (0, 0, 0), (0, 0, 0),
            (0, 0, 0), (0, 0, 0), (0, 0, 0)]
    prims = _linear_prims(g, coef)
    for fi, fj in ((3.3, 4.7), (0.2, 5.0), (7.0, 0.3), (0.1, 0.1), (15.7, 6.0), (8.0, 11.8)):
        X = np.array([0, g.startx1 + fi * g.dx1, g.startx2 + fj * g.dx2, 0])
        ti, tj = fi - 0.5, fj - 0.5          # continuous cell-centre index
        ti = min(ti, g.n1 - 1.0)             # clamp past the last centre
        tj = min(tj, g.n2 - 1.0)
        want_rho = coef[0][0] + coef[0][1] * ti + coef[0][2] * tj
        out = O.fluid_at(P, prims, X)
        assert out["ne"] / O.units(P)["Ne_unit"] == pytest.approx(want_rho, rel=1e-13)


def test_fluid_frame_invariants(dump128):
    """u.u = -1, u.b = 0, b.b >= 0 for sampled torus points."""
    P = O.params(dump128.grid, width=4, height=4, inclination=60.0, M_unit=1e23, R_high=10.0)
    g = dump128.grid
    rng = np.random.default_rng(3)
    for _ in range(100):
        X = np.array([0, rng.uniform(g.startx1, g.stopx1), rng.uniform(0.02, 0.98), 0])
        f = O.fluid_at(P, dump128.prims, X)
        gd = O.metric_dd(P, X)
        u, b = f["ucon"], f["bcon"]
        assert u @ gd @ u == pytest.approx(-1.0, abs=1e-10)
        assert abs(u @ gd @ b) < 1e-10 * (1 + abs(b).max())
        assert b @ gd @ b >= -1e-18


def _uniform_prims(grid, rho, uu, B1):
    p = np.zeros((8, grid.n1, grid.n2, 1))
    p[0], p[1], p[5] = rho, uu, B1
    return p


def test_electron_temperature_r_beta_limits():
    """R-beta model (Moscibrodzka et al. 2016): for beta >> 1 the ratio is
    R_high, for beta << 1 it is R_low; Theta_e = u/rho (gam2-1) m_p/m_e/(1+R)
    with gam2 = (1 + 2/3 (R+1)/(R+2) + 4/3)/2 (catches swapped R_high/R_low)."""
    g = _grid(n1=8, n2=8)
    X = np.array([0, g.startx1 + 3.3 * g.dx1, 0.41, 0])
    for Rh, Rl in ((40.0, 1.0), (3.0, 7.0)):
        P = O.params(g, width=4, height=4, inclination=60.0, M_unit=1e23, R_high=Rh, R_low=Rl)
        for B1, R in ((0.0, Rh), (1e3, Rl)):
            f = O.fluid_at(P, _uniform_prims(g, 1.0, 0.05, B1), X)
            gam2 = 0.5 * ((1 + 2.0 / 3.0 * (R + 1) / (R + 2)) + 4.0 / 3.0)
            want = 0.05 * (gam2 - 1) * (MP / ME) / (1 + R)
            assert f["thetae"] == pytest.approx(want, rel=1e-6)


def test_sigma_cut():
    """b^2/rho > 1 makes the hook return 0 (raptor_harm_model.c:285-290)."""
    g = _grid(n1=8, n2=8)
    P = O.params(g, width=4, height=4, inclination=60.0, M_unit=1e23, R_high=10.0)
    X = np.array([0, g.startx1 + 3.3 * g.dx1, 0.41, 0])
    assert O.fluid_at(P, _uniform_prims(g, 1.0, 0.05, 1e-3), X)["ok"]
    assert not O.fluid_at(P, _uniform_prims(g, 1.0, 0.05, 10.0), X)["ok"]


# ------------------------------------------------- plasma-frame quantities
#
# The transfer sample's Doppler factor nu_p / nu and pitch-angle cosine
# (radiative_transfer.c:17-47, composed at :61-84) against an independent
# construction: the Kerr-Schild metric written here in closed form and pulled
# back to MKS, the fluid 4-velocity built as u = gamma n + u~ from the normal
# observer n (HARM's primitive velocity), and the angle measured in an
# orthonormal fluid-frame tetrad.  A mistake in the oracle's lowering of k
# (the inverse metric, a transposed index), a sign of nu_p, a dropped
# m_e c^2 / h or B_unit, or u^mu in place of u_mu fails these.


def _mks_metric_np(X, a, h):
    r = math.exp(X[1])
    th = math.pi * X[2] + 0.5 * (1 - h) * math.sin(2 * math.pi * X[2])
    dth = math.pi * (1 + (1 - h) * math.cos(2 * math.pi * X[2]))
    s2, c2 = math.sin(th) ** 2, math.cos(th) ** 2
    S = r * r + a * a * c2
    z = 2 * r / S
    g = np.zeros((4, 4))
    g[0, 0] = -(1 - z)
    g[0, 1] = g[1, 0] = z
    g[0, 3] = g[3, 0] = -z * a * s2
    g[1, 1] = 1 + z
    g[1, 3] = g[3, 1] = -a * s2 * (1 + z)
    g[2, 2] = S
    g[3, 3] = s2 * (S + a * a * s2 * (1 + z))
    J = np.diag([1.0, r, dth, 1.0])
    return J @ g @ J


def _null_k(g, kspace):
    """k^t > 0 with g(k, k) = 0 for the spatial components kspace."""
    ks = np.asarray(kspace, float)
    A = g[0, 0]
    B = 2 * g[0, 1:] @ ks
    Cq = ks @ g[1:, 1:] @ ks
    kt = (-B - math.sqrt(B * B - 4 * A * Cq)) / (2 * A)
    if kt < 0:
        kt = (-B + math.sqrt(B * B - 4 * A * Cq)) / (2 * A)
    return np.concatenate([[kt], ks])


def _fluid_np(g, vt, Bp):
    """u = gamma n + u~ (n the normal observer, gamma^2 = 1 + g_ij u~^i u~^j)
    and b^t = u_i B^i, b^i = (B^i + u^i b^t) / u^t."""
    gi = np.linalg.inv(g)
    alpha = 1.0 / math.sqrt(-gi[0, 0])
    n = -alpha * gi[:, 0]
    vt = np.asarray(vt, float)
    gamma = math.sqrt(1.0 + vt @ g[1:, 1:] @ vt)
    u = gamma * n + np.concatenate([[0.0], vt])
    ul = g @ u
    bt = ul[1:] @ np.asarray(Bp, float)
    b = np.concatenate([[bt], (np.asarray(Bp, float) + u[1:] * bt) / u[0]])
    return u, b


def _tetrad_cos(g, u, k, b):
    """cos of the angle between the photon's direction and the field, both
    expressed in an orthonormal tetrad whose time leg is u (Gram-Schmidt of
    the coordinate legs d/dX1, d/dX2, d/dX3 orthogonal to u)."""
    dot = lambda p, q: p @ g @ q
    legs = []
    for i in (1, 2, 3):
        v = np.zeros(4)
        v[i] = 1.0
        v = v + dot(v, u) * u
        for e in legs:
            v = v - dot(v, e) * e
        legs.append(v / math.sqrt(dot(v, v)))
    kk = np.array([dot(k, e) for e in legs])
    bb = np.array([dot(b, e) for e in legs])
    return kk @ bb / (np.linalg.norm(kk) * np.linalg.norm(bb)), np.linalg.norm(bb)


def _uniform_fluid(grid, vt, Bp, rho=1.0, uu=0.05):
    p = np.zeros((8, grid.n1, grid.n2, 1))
    p[0], p[1] = rho, uu
    p[2], p[3], p[4] = vt
    p[5], p[6], p[7] = Bp
    return p


def _random_points(grid, rng, n):
    for _ in range(n):
        X = np.array([0.0, rng.uniform(grid.startx1 + 2 * grid.dx1, grid.stopx1 - 2 * grid.dx1),
                      rng.uniform(0.05, 0.95), rng.uniform(0, 2 * math.pi)])
        ks = rng.normal(size=3) * np.array([0.3, 0.05, 0.2])
        yield X, ks


def test_doppler_factor_of_the_normal_observer():
    """With u~ = 0 the fluid is the normal observer, u_mu = (-alpha, 0, 0, 0),
    so nu_p / nu = -u.k = alpha k^t = k^t / sqrt(-g^tt), g^tt = -(1 + 2r/Sigma)
    in (M)KS; at several frequencies and spins."""
    rng = np.random.default_rng(11)
    for a in (0.9375, 0.0, -0.5):
        g0 = _grid(a=a, n1=16, n2=16)
        P = O.params(g0, width=4, height=4, inclination=60.0, M_unit=1e23, R_high=10.0)
        prims = _uniform_fluid(g0, (0, 0, 0), (1e-3, 0, 0))
        for X, ks in _random_points(g0, rng, 40):
            g = _mks_metric_np(X, a, g0.hslope)
            k = _null_k(g, ks)
            r = math.exp(X[1])
            th = math.pi * X[2] + 0.5 * (1 - g0.hslope) * math.sin(2 * math.pi * X[2])
            S = r * r + a * a * math.cos(th) ** 2
            want = k[0] / math.sqrt(1 + 2 * r / S)
            s = O.sample_at(P, prims, np.concatenate([X, k]), [2.3e11, 8.6e10])
            assert s["ok"]
            assert s["nu_p"] / np.array([2.3e11, 8.6e10]) == pytest.approx(
                [want, want], rel=1e-12)


def test_doppler_factor_of_a_moving_fluid():
    """For HARM's primitive velocity u~ the fluid moves with u = gamma n + u~;
    nu_p / nu = -u_mu k^mu with u lowered by the closed-form metric, and the
    1 Hz floor where that is <= 0 (radiative_transfer.c:24-26)."""
    rng = np.random.default_rng(12)
    g0 = _grid(n1=16, n2=16)
    P = O.params(g0, width=4, height=4, inclination=60.0, M_unit=1e23, R_high=10.0)
    n = 0
    for vt in ((0.3, -0.02, 0.1), (-0.5, 0.05, 0.4), (0.05, 0.0, -0.3)):
        prims = _uniform_fluid(g0, vt, (1e-3, 2e-4, -1e-4))
        for X, ks in _random_points(g0, rng, 30):
            g = _mks_metric_np(X, g0.a, g0.hslope)
            k = _null_k(g, ks)
            u, _ = _fluid_np(g, vt, (0, 0, 0))
            want = -(g @ u) @ k
            got = O.sample_at(P, prims, np.concatenate([X, k]), [2.3e11])["nu_p"][0] / 2.3e11
            if want > 0:
                assert got == pytest.approx(want, rel=1e-11)
                n += 1
            else:
                assert got == 1.0 / 2.3e11
    assert n > 60


def test_pitch_angle_in_the_fluid_frame():
    """mu = cos of the angle between the photon and b measured in the fluid's
    rest frame (an orthonormal tetrad with time leg u), for random fluid
    velocities, fields and photon directions; and |b| there times B_unit is
    the oracle's B (the field strength the emissivity uses)."""
    rng = np.random.default_rng(13)
    g0 = _grid(n1=16, n2=16)
    P = O.params(g0, width=4, height=4, inclination=60.0, M_unit=1e23, R_high=10.0)
    B_unit = O.units(P)["B_unit"]
    for _ in range(6):
        vt = rng.normal(size=3) * np.array([0.3, 0.03, 0.3])
        Bp = rng.normal(size=3) * np.array([1e-3, 1e-4, 1e-3])
        prims = _uniform_fluid(g0, vt, Bp)
        for X, ks in _random_points(g0, rng, 20):
            g = _mks_metric_np(X, g0.a, g0.hslope)
            k = _null_k(g, ks)
            u, b = _fluid_np(g, vt, Bp)
            want, bnorm = _tetrad_cos(g, u, k, b)
            s = O.sample_at(P, prims, np.concatenate([X, k]), [2.3e11])
            assert s["ok"]
            assert s["mu"] == pytest.approx(want, abs=1e-10)
            assert s["B"] == pytest.approx(bnorm * B_unit, rel=1e-10)


def test_units_closed_form():
    P = _P(M_unit=3e21)
    u = O.units(P)
    L = GG * 8.9505e39 / C ** 2
    assert u["L_unit"] == pytest.approx(L, rel=1e-14)
    assert u["RHO_unit"] * L ** 3 == pytest.approx(3e21, rel=1e-13)
    assert u["B_unit"] ** 2 / (4 * math.pi * C ** 2) == pytest.approx(u["RHO_unit"], rel=1e-13)
    assert u["Ne_unit"] * (MP + ME) == pytest.approx(u["RHO_unit"], rel=1e-13)
    assert u["cutoff_inner"] == pytest.approx(1.1 * (1 + math.sqrt(1 - 0.9375 ** 2)), rel=1e-14)
    jf = O.lib().or_jansky_factor(P)
    assert jf == pytest.approx(1e23 * (40 * L / 2.6228263e22 / 16) ** 2, rel=1e-13)


# ---------------------------------------------------------------- emission


def test_thsynch_fit_against_exact_integral():
    """The emissivity against the Maxwell-Juettner integral of the single
    electron synchrotron spectrum (tests/golden/thsynch.json); the fit is good
    to ~2% at Theta_e >= 30 (a dropped sqrt(2), pi or 3 is off by >= 40%)."""
    rows = json.load(open(os.path.join(GOLD, "thsynch.json")))["rows"]
    lib = O.lib()
    for r in rows:
        Th, X, sth = r["Theta"], r["X"], r["sin_theta"]
        B, ne = 10.0, 1e5
        nuc = E * B / (2 * math.pi * ME * C)
        nu = X * (2.0 / 9.0) * nuc * Th ** 2 * sth
        j = lib.or_j_thsynch(B, math.sqrt(1 - sth * sth), Th, nu, ne)
        exact = r["j_over_ne_e2nuc_c"] * ne * (E * E / C) * nuc
        # the reference's K2(1/Theta) ~ 2 Theta^2 (j_nu.c:146); remove that approximation
        exact_k2approx = exact * r["K2_exact"] / (2 * Th * Th)
        assert j / exact_k2approx == pytest.approx(1.0, abs=0.03), r


def test_thsynch_low_frequency_slope_and_gates():
    lib = O.lib()
    B, Th, ne = 10.0, 20.0, 1e5
    nuc = E * B / (2 * math.pi * ME * C)
    nus = (2.0 / 9.0) * nuc * Th ** 2 * 0.8
    j1 = lib.or_j_thsynch(B, 0.6, Th, 1e-9 * nus, ne)
    j2 = lib.or_j_thsynch(B, 0.6, Th, 2e-9 * nus, ne)
    assert math.log(j2 / j1) / math.log(2) == pytest.approx(1.0 / 3.0, abs=0.01)
    assert lib.or_j_thsynch(B, 0.6, 0.9e-4, 1e11, ne) == 0.0   # Theta_e gate
    assert lib.or_j_thsynch(B, 1.0, Th, 1e11, ne) == 0.0       # sin(pitch) gate
    assert lib.or_j_thsynch(B, 0.6, Th, 1.0, ne) == 0.0        # nu_p floor gate
    assert lib.or_j_thsynch(2 * B, 0.6, Th, 1e11, 2 * ne) > 0


# ------------------------------------------------ other emission models


def test_hyp2f1_against_scipy_and_closed_forms():
    """or_hyp2f1 (Pfaff's transformation for -1 <= z <= 0, the 1/(1-z)
    continuation below) against scipy.special.hyp2f1 at the kappa model's
    parameters (a, b, c) = (kappa - 1/3, kappa + 1, kappa + 2/3) over the
    z = -kappa Theta_e it meets, and at other parameters with b - a not an
    integer; and the closed forms 2F1(a, b; b; z) = (1-z)^-a and
    2F1(1, 1; 2; z) = -ln(1-z)/z (a swapped argument or a wrong
    transformation is off by far more than 1e-12)."""
    from scipy import special
    lib = O.lib()
    k = 3.5
    for z in -np.concatenate([np.logspace(-4, 3.5, 40), [1.0, 0.999, 1.001]]):
        assert lib.or_hyp2f1(k - 1 / 3, k + 1, k + 2 / 3, z) == pytest.approx(
            special.hyp2f1(k - 1 / 3, k + 1, k + 2 / 3, z), rel=1e-12), z
    for a, b, c in ((0.5, 1.7, 2.3), (1.2, 0.4, 3.1), (2.5, 0.25, 4.5), (3.0, 5.5, 4.0)):
        for z in (-0.05, -0.5, -0.97, -1.3, -7.0, -80.0):
            assert lib.or_hyp2f1(a, b, c, z) == pytest.approx(special.hyp2f1(a, b, c, z),
                                                              rel=1e-11), (a, b, c, z)
    for z in (-0.3, -2.0, -40.0):
        assert lib.or_hyp2f1(0.7, 1.9, 1.9, z) == pytest.approx((1 - z) ** -0.7, rel=1e-13)
    for z in (-0.2, -0.9):
        assert lib.or_hyp2f1(1.0, 1.0, 2.0, z) == pytest.approx(-math.log(1 - z) / z, rel=1e-13)


def _kappa_fit(r):
    lib = O.lib()
    B, ne = 30.0, 1e6
    w, X, sth = r["w"], r["X"], r["sin_theta"]
    mu = math.sqrt(1 - sth * sth)
    nuc = E * B / (2 * math.pi * ME * C)
    nu = X * nuc * sth * (w * 3.5) ** 2
    J = lib.or_j_kappa(nu, ne, w, B, mu) / (ne * E * E * nuc * sth / C)
    A = lib.or_a_kappa(nu, ne, w, B, mu) * B * sth / (ne * E)
    return J, A


def test_kappa_fits_against_distribution_integrals():
    """The kappa = 3.5 emissivity and absorptivity fits (j_nu.c:28-106,
    Pandya et al. 2016) against integrals over the kappa distribution of the
    single-electron spectrum (tests/golden/kappa.json; the absorptivity
    integral reproduces Kirchhoff's law for a thermal distribution to 3e-4).
    Emissivity: within 6% for widths w = Theta_e >= 10, the fits' regime,
    and 50% at w = 1.  Absorptivity: the fit, as the reference writes it,
    runs from 0.94 of the integral at X = nu/nu_s = 0.01 (the A_low term with
    the hypergeometric factor) to 0.48 at X = 1e4 (the A_high term); the band
    [0.45, 0.96] and the monotone fall in X catch a wrong exponent, a dropped
    2F1 or Gamma factor, or swapped terms (each moves the ratio by a factor of
    2 or more, or by orders of magnitude across X).  Both depend on X alone,
    not on the pitch angle, like the integrals."""
    rows = json.load(open(os.path.join(GOLD, "kappa.json")))["rows"]
    by_w = {}
    for r in rows:
        J, A = _kappa_fit(r)
        rj, ra = J / r["J"], A / r["A"]
        if r["w"] >= 10:
            assert rj == pytest.approx(1.0, abs=0.06), r
            assert 0.45 <= ra <= 0.96, r
            if r["sin_theta"] == 0.8:
                by_w.setdefault(r["w"], []).append((r["X"], ra))
        elif r["w"] == 1.0:
            assert 0.5 <= rj <= 1.5 and 0.4 <= ra <= 0.8, r
    for w, pts in by_w.items():
        ra = [q for _, q in sorted(pts)]
        assert all(b < a for a, b in zip(ra, ra[1:])), (w, ra)
    same = [r for r in rows if r["X"] == 100.0]
    for w in {r["w"] for r in same}:
        a, b = [_kappa_fit(r) for r in same if r["w"] == w]
        assert a == pytest.approx(b, rel=1e-12)


def test_kappa_fits_asymptotic_slopes_and_scalings():
    """What the kappa distribution fixes exactly in the fits of j_nu.c:28-106
    (Pandya et al. 2016, kappa = 3.5), independent of their fitted constants:
    the power-law slopes in X = nu / nu_s, nu_s = nu_c sin(theta) (w kappa)^2
    -- emissivity X^(1/3) below the peak and X^(-(kappa-2)/2) above it,
    absorptivity X^(-5/3) and X^(-(kappa+3)/2) -- and the scalings at fixed X:
    j proportional to n_e nu_c sin(theta) with no other dependence on the
    width w; alpha proportional to n_e / (B sin(theta)) times (w kappa)^-5 in
    the high-X regime and (w kappa)^(kappa - 16/3) 2F1(kappa - 1/3, kappa + 1;
    kappa + 2/3; -kappa w) in the low-X regime (2F1 from scipy).  A wrong
    exponent, a swapped regime, X formed with the wrong power of w or a
    dropped hypergeometric factor fails one of these.  The fits' constant
    factors are pinned only by the distribution integrals above."""
    from scipy.special import hyp2f1
    lib = O.lib()
    kap = 3.5

    def nu_of(X, w, B, mu):
        nuc = E * B / (2 * math.pi * ME * C)
        return X * nuc * math.sqrt(1 - mu * mu) * (w * kap) ** 2

    B, mu, ne = 30.0, 0.4, 1e6

    def slope(fn, X, w, h=1e-3):
        lo = fn(nu_of(X * math.exp(-h), w, B, mu), ne, w, B, mu)
        hi = fn(nu_of(X * math.exp(h), w, B, mu), ne, w, B, mu)
        return (math.log(hi) - math.log(lo)) / (2 * h)

    for w in (3.0, 30.0):
        assert slope(lib.or_j_kappa, 1e-12, w) == pytest.approx(1 / 3, abs=2e-3)
        assert slope(lib.or_j_kappa, 1e12, w) == pytest.approx(-(kap - 2) / 2, abs=2e-3)
        assert slope(lib.or_a_kappa, 1e-12, w) == pytest.approx(-5 / 3, abs=2e-3)
        assert slope(lib.or_a_kappa, 1e14, w) == pytest.approx(-(kap + 3) / 2, abs=2e-3)
    hyp = lambda w: hyp2f1(kap - 1 / 3, kap + 1, kap + 2 / 3, -kap * w)
    for X in (1e-12, 1.0, 1e12):
        for (B1, mu1, w1), (B2, mu2, w2) in (((30.0, 0.4, 3.0), (7.0, -0.8, 40.0)),
                                             ((2.0, 0.1, 0.5), (90.0, 0.6, 12.0))):
            j1 = lib.or_j_kappa(nu_of(X, w1, B1, mu1), ne, w1, B1, mu1)
            j2 = lib.or_j_kappa(nu_of(X, w2, B2, mu2), 3 * ne, w2, B2, mu2)
            s1, s2 = math.sqrt(1 - mu1 ** 2), math.sqrt(1 - mu2 ** 2)
            assert j2 / j1 == pytest.approx(3 * (B2 * s2) / (B1 * s1), rel=1e-12)
    for X, w1, w2 in ((1e14, 3.0, 40.0), (1e14, 0.5, 12.0)):
        a1 = lib.or_a_kappa(nu_of(X, w1, B, mu), ne, w1, B, mu)
        a2 = lib.or_a_kappa(nu_of(X, w2, 2 * B, -0.8), 5 * ne, w2, 2 * B, -0.8)
        sin_ratio = math.sqrt(1 - mu * mu) / 0.6
        assert a2 / a1 == pytest.approx(5 / 2 * sin_ratio * (w2 / w1) ** -5, rel=2e-3)
    for X, w1, w2 in ((1e-14, 3.0, 40.0), (1e-14, 0.5, 12.0)):
        a1 = lib.or_a_kappa(nu_of(X, w1, B, mu), ne, w1, B, mu)
        a2 = lib.or_a_kappa(nu_of(X, w2, B, mu), ne, w2, B, mu)
        want = (w2 / w1) ** (kap - 16 / 3) * hyp(w2) / hyp(w1)
        assert a2 / a1 == pytest.approx(want, rel=2e-3)


def test_kappa_gates():
    lib = O.lib()
    B, ne, w = 30.0, 1e6, 5.0
    assert lib.or_j_kappa(1.0, ne, w, B, 0.5) == 0.0          # nu_p floor
    assert lib.or_j_kappa(1e11, ne, w, B, 1.0) == 0.0         # |cos pitch| = 1
    assert lib.or_a_kappa(1e11, ne, w, B, -1.0) == 0.0
    assert lib.or_a_kappa(1e11, ne, 1.0 / 3.5, B, 0.5) == 0.0  # |z| = 1 (j_nu.c:79)
    assert lib.or_j_kappa(1e11, ne, 1.0 / 3.5, B, 0.5) > 0.0
    # emissivity and absorptivity scale with n_e
    assert lib.or_j_kappa(1e11, 2 * ne, w, B, 0.5) == pytest.approx(
        2 * lib.or_j_kappa(1e11, ne, w, B, 0.5), rel=1e-14)


def test_thsynchav_against_angle_averaged_integral():
    """The angle-averaged thermal fit (j_nu.c:157-167, Dexter 2009) times e^2/c
    against the thermal emissivity integral averaged over isotropic pitch
    angles (tests/golden/thsynchav.json): within 3% for Theta_e >= 3 and
    x_M = 1..1000 (a dropped 2 sqrt(3), a wrong x_M power or the Mahadevan
    coefficients mistyped are off by 20% or more)."""
    rows = json.load(open(os.path.join(GOLD, "thsynchav.json")))["rows"]
    lib = O.lib()
    B, ne = 30.0, 1e6
    nuc = E * B / (2 * math.pi * ME * C)
    for r in rows:
        nu = r["x_M"] * 1.5 * nuc * r["Theta"] ** 2
        j = lib.or_j_thsynchav(B, r["Theta"], nu, ne) * (E * E / C)
        assert j / (ne * E * E * nuc / C) / r["j_over_ne_e2nuc_c"] == pytest.approx(
            1.0, abs=0.03), r


def test_emission_model_composition():
    """or_emission: model 0 is the path's thermal fit with Kirchhoff's
    absorptivity; model 1 the angle-averaged fit in cgs with Kirchhoff's;
    model 2 the kappa emissivity and absorptivity fits."""
    lib = O.lib()
    q = (30.0, 0.4, 12.0, 2.3e11, 1e6)
    j0, a0 = O.emission(O.THSYNCH, *q)
    assert j0 == lib.or_j_thsynch(*q) and a0 == lib.or_a_th(j0, q[3], q[2])
    j1, a1 = O.emission(O.THSYNCHAV, *q)
    assert j1 == pytest.approx(lib.or_j_thsynchav(q[0], q[2], q[3], q[4]) * (E * E / C),
                               rel=1e-15)
    assert a1 == lib.or_a_th(j1, q[3], q[2])
    j2, a2 = O.emission(O.KAPPA, *q)
    assert j2 == lib.or_j_kappa(q[3], q[4], q[2], q[0], q[1])
    assert a2 == lib.or_a_kappa(q[3], q[4], q[2], q[0], q[1])


def test_kirchhoff_and_planck_limits():
    """alpha = j / B_nu with B_nu = 2 h nu^3/c^2 / (exp(h nu/k T) - 1): the
    reference's Planck function omits h and e2_ch/e2_c restores it."""
    lib = O.lib()
    assert 7.695589048e-30 == pytest.approx(E * E / C, rel=1e-7)
    assert 1.161410e-03 == pytest.approx(E * E / (C * H), rel=1e-5)
    for nu, Th in ((1e11, 10.0), (3e12, 0.5), (1e9, 3e-3)):
        T = Th * ME * C * C / KB
        x = H * nu / (KB * T)
        Bnu = 2 * H * nu ** 3 / C ** 2 / math.expm1(x)
        # exp(x) - 1 as in j_nu.c:190 loses ~1e-16/x relative (expm1 does not)
        assert H * lib.or_planck(nu, Th) == pytest.approx(Bnu, rel=max(1e-9, 4e-16 / x))
        assert lib.or_a_th(1e-20, nu, Th) * Bnu == pytest.approx(1e-20, rel=2e-5)
    # Rayleigh-Jeans limit (x = 8e-9: correction x/2 plus cancellation 1e-16/x)
    nu, Th = 1e12, 1.0
    T = Th * ME * C * C / KB
    assert H * lib.or_planck(nu, Th) == pytest.approx(2 * nu * nu * KB * T / C ** 2, rel=2e-7)


# ---------------------------------------------------------------- transfer


def _series(js, als, nu_p, dl):
    lib = O.lib()
    import ctypes
    tau = ctypes.c_double(0.0)
    I = ctypes.c_double(0.0)
    dt = ctypes.c_double(0.0)
    di = ctypes.c_double(0.0)
    for j, a in zip(js, als):
        lib.or_transfer_step(ctypes.byref(tau), ctypes.byref(I), j, a, nu_p, dl,
                             ctypes.byref(dt), ctypes.byref(di))
    return tau.value, I.value


def test_transfer_slab_closed_form():
    """Constant j, alpha over n end-point samples: tau_k = k d, I = (j/nu^2) dl
    sum_k exp(-k d) -- a geometric series (radiative_transfer.c:88-96)."""
    j, a, nu, dl, n = 2e-5, 3e-14, 1e11, 0.05, 200
    d = nu * a * dl
    tau, I = _series([j] * n, [a] * n, nu, dl)
    q = math.exp(-d)
    want = (j / nu ** 2) * dl * q * (1 - q ** n) / (1 - q)
    assert tau == pytest.approx(n * d, rel=1e-12)
    assert I == pytest.approx(want, rel=1e-12)
    # optically thin limit: plain sum of j dl / nu^2
    tau0, I0 = _series([j] * n, [0.0] * n, nu, dl)
    assert tau0 == 0 and I0 == pytest.approx(n * j * dl / nu ** 2, rel=1e-12)


def test_transfer_tau_gate_and_nan_poisoning():
    """Nothing is added once tau >= ln 1000; a NaN tau (B_nu underflow with
    j = 0) closes the gate for the rest of the ray."""
    j, nu, dl = 1e-5, 1e11, 1.0
    d = 0.5                                 # tau per sample
    a = d / (nu * dl)
    n = 30
    tau, I = _series([j] * n, [a] * n, nu, dl)
    kmax = max(k for k in range(1, n + 1) if k * d < math.log(1000.0))
    want = sum((j / nu ** 2) * math.exp(-k * d) * dl for k in range(1, kmax + 1))
    assert I == pytest.approx(want, rel=1e-12)
    tau, I = _series([j, 0.0, j, j], [0.0, float("nan"), 0.0, 0.0], nu, dl)
    assert math.isnan(tau) and I == pytest.approx(j / nu ** 2 * dl, rel=1e-12)


# ---------------------------------------------------------------- tangents


def test_tangent_pass_against_central_differences(dump128):
    """d image / d ln(M_unit) and d image / d R_high from the tangent pass
    against central differences.  Pixels whose image jumps between p-h and
    p+h (a sample crossing the tau < ln 1000 or Theta_e > 1e-4 gate) are found
    by comparing steps h and h/2 -- a jump makes the difference quotient grow
    like 1/h -- and are excluded; they must be few."""
    kw = dict(width=20, height=20, inclination=60.0, R_low=1.0)
    M0, R0 = 1e23, 10.0
    _, dimg = O.render(O.params(dump128.grid, M_unit=M0, R_high=R0, **kw), dump128.prims,
                       [2.3e11], grad=True)

    def fd(q, h):
        if q == 0:
            pa, pb, st = (M0 * math.exp(h), R0), (M0 * math.exp(-h), R0), 2 * h
        else:
            pa, pb, st = (M0, R0 * (1 + h)), (M0, R0 * (1 - h)), 2 * h * R0
        ip = O.render(O.params(dump128.grid, M_unit=pa[0], R_high=pa[1], **kw),
                      dump128.prims, [2.3e11])[:, 0]
        im = O.render(O.params(dump128.grid, M_unit=pb[0], R_high=pb[1], **kw),
                      dump128.prims, [2.3e11])[:, 0]
        return (ip - im) / st

    for q in range(2):
        an = dimg[:, 0, q]
        f1, f2 = fd(q, 1e-3), fd(q, 5e-4)
        den = np.maximum(np.abs(an), 1e-3 * np.abs(an).max())
        smooth = np.abs(f1 - f2) / den < 1e-3
        assert smooth.mean() > 0.9, q
        err = np.abs(f2 - an) / den
        assert err[smooth].max() < 2e-3, (q, err[smooth].max())
        assert f2[smooth].sum() / an[smooth].sum() == pytest.approx(1.0, abs=1e-4)


# ---------------------------------------------------------------- data


def test_torus_generator_properties(dump128):
    from raptor_amd import torus
    g = dump128.grid
    p = dump128.prims[..., 0]
    ex = dump128.extras
    # density maximum near r_max = 12 on the equator
    i, j = np.unravel_index(np.argmax(p[0]), p[0].shape)
    assert abs(ex["r"][i, j] - 12.0) < 1.5 and abs(ex["th"][i, j] - math.pi / 2) < 0.2
    # u.u = -1 of the stored 4-velocities
    gd, _, _ = torus._mks_metric(ex["r"], ex["th"], ex["x2"], g.a, g.hslope)
    uu = np.einsum("i...,...ij,j...->...", ex["ucon"], gd, ex["ucon"])
    assert np.abs(uu + 1).max() < 1e-6
    # plasma beta normalisation (before noise) is within the noise of 100
    bsq = np.einsum("i...,i...->...", ex["bcon"], ex["bcov"])
    beta_act = (g.gam - 1) * p[1].max() / (0.5 * bsq.max())
    assert 80 < beta_act < 125
    # every primitive is float32-representable
    assert np.array_equal(dump128.prims.astype(np.float32).astype(np.float64), dump128.prims)


def test_harm_ascii_and_binary_round_trip(tmp_path):
    from raptor_amd import harm, torus
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        d = torus.fishbone_moncrief(24, 16)
    f = tmp_path / "dump.txt"
    harm.write_harm_ascii(str(f), d)
    r = harm.read_harm_ascii(str(f))
    assert np.array_equal(r.prims, d.prims)
    assert (r.grid.n1, r.grid.n2, r.grid.a, r.grid.hslope) == (24, 16, d.grid.a, d.grid.hslope)
    assert r.grid.startx1 == d.grid.startx1 and r.grid.dx1 == d.grid.dx1
    line = open(f).readline().split()
    assert len(line) == 26
    body = open(f).read().splitlines()[1].split()
    assert len(body) == 34
    fb = tmp_path / "dump.bin"
    harm.write_prims_bin(str(fb), d)
    rb = harm.read_prims_bin(str(fb))
    assert np.array_equal(rb.prims, d.prims) and rb.grid == d.grid


def test_tau_gate_margin(dump128):
    """The oracle's tau-gate margin: no ray crosses ln 1000 in an optically
    thin setting (small M_unit), many do in a thick one, margins are
    non-negative, and a ray's margin is the distance from ln 1000 of tau on
    one side of the crossing, which shrinks as the samples get denser."""
    g = dump128.grid

    def margins(M_unit, stepsize=0.03):
        P = O.params(g, width=24, height=24, inclination=60.0, M_unit=M_unit, R_high=10.0,
                     stepsize=stepsize)
        img, gate = O.render(P, dump128.prims, [2.3e11], gate=True)
        return img[:, 0], gate[:, 0]

    _, thin = margins(1e17)
    assert np.all(np.isinf(thin))
    img, thick = margins(1e25)
    fin = np.isfinite(thick)
    assert fin.sum() > 50
    assert np.all(thick[fin] >= 0)
    # the crossing falls between two samples, so the margin is at most half
    # of the largest optical depth a single step adds; with the step halved the
    # median margin shrinks accordingly
    _, fine = margins(1e25, stepsize=0.015)
    both = fin & np.isfinite(fine)
    assert np.median(fine[both]) < 0.75 * np.median(thick[both])
