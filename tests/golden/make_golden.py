#!/usr/bin/env python3
"""Writes tests/golden/*.json: reference values for the oracle's building
blocks that do not come from the oracle (or from RAPTOR) itself.

1. kerr_mks.json -- the Kerr metric in HARM's modified Kerr-Schild coordinates
   derived symbolically (sympy) by pulling back the Cartesian Kerr-Schild form
       g = eta + f l (x) l,  f = 2 r^3 / (r^4 + a^2 z^2),
       l = (1, (r x + a y)/(r^2+a^2), (r y - a x)/(r^2+a^2), z/r),
       x + i y = (r + i a) e^{i phi} sin(theta), z = r cos(theta),
   through r = exp(X1), theta = pi X2 + (1-h)/2 sin(2 pi X2), phi = X3.  The
   orientation of phi is the one for which the hole rotates towards +phi
   (g_t_phi < 0 for a > 0).  From it, at sample points: g_dd, g_uu (30-digit
   inverse) and the connection Gamma^m_ab = 1/2 g^md (d_a g_db + d_b g_da -
   d_d g_ab) from exact symbolic derivatives.

2. thsynch.json -- the thermal synchrotron emissivity as the integral over a
   Maxwell-Juettner distribution of the single-electron spectrum
   (Rybicki & Lightman 1979, eq. 6.18; ultrarelativistic, pitch angle = angle
   to the line of sight):
       j_nu = n_e sqrt(3) e^2 nu_c sin(theta) / (2 c)
              * int dgamma gamma^2 beta exp(-gamma/Theta) / (Theta K2(1/Theta))
                F(2 nu / (3 gamma^2 nu_c sin(theta))),
       F(x) = x int_x^inf K_{5/3}(t) dt,  nu_c = e B / (2 pi m_e c),
   stored as j_nu / (n_e e^2 nu_c / c), with the exact K2.

3. kappa.json -- emissivity and absorptivity of the relativistic kappa
   distribution n(gamma) = N gamma sqrt(gamma^2-1) (1 + (gamma-1)/(kappa w))^-(kappa+1)
   (normalised to n_e numerically), kappa = 3.5, from the single-electron
   spectrum P_nu = 2 pi sqrt(3) e^2 nu_c sin(theta)/c F(x), x = 2 nu / (3
   gamma^2 nu_c sin(theta)) (Rybicki & Lightman 1979, eq. 6.18):
       j_nu     = 1/(4 pi) int dgamma n P_nu,
       alpha_nu = -c^2 / (8 pi nu^2 m_e c^2) int dgamma P_nu gamma^2 d/dgamma(n / gamma^2)
   (R&L eq. 6.50), stored as J = j_nu / (n_e e^2 nu_c sin(theta) / c) and
   A = alpha_nu B sin(theta) / (n_e e), the normalisations of the fits of
   Pandya et al. (2016).

4. thsynchav.json -- the thermal emissivity of item 2 averaged over an
   isotropic distribution of pitch angles (Gauss-Legendre in cos theta),
   stored as j_nu / (n_e e^2 nu_c / c) against x_M = 2 nu / (3 nu_c Theta^2).

Items 3-4 tabulate F on a logarithmic grid (cubic interpolation in log-log,
relative error below 1e-7) to keep the nested integrals fast.

Run:  python tests/golden/make_golden.py [kerr] [thsynch] [kappa] [thsynchav]
(about two minutes each)
"""
import json
import os

import mpmath as mp
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def kerr_mks_golden():
    t, X1, X2, X3 = sp.symbols("t X1 X2 X3", real=True)
    a, h = sp.symbols("a h", real=True)
    r = sp.exp(X1)
    th = sp.pi * X2 + (1 - h) / 2 * sp.sin(2 * sp.pi * X2)
    ph = X3

    def metric(sgn):
        x = (r * sp.cos(ph) - sgn * a * sp.sin(ph)) * sp.sin(th)
        y = (r * sp.sin(ph) + sgn * a * sp.cos(ph)) * sp.sin(th)
        z = r * sp.cos(th)
        f = 2 * r ** 3 / (r ** 4 + a ** 2 * z ** 2)
        l = [sp.Integer(1), (r * x + sgn * a * y) / (r ** 2 + a ** 2),
             (r * y - sgn * a * x) / (r ** 2 + a ** 2), z / r]
        eta = sp.diag(-1, 1, 1, 1)
        gc = sp.Matrix(4, 4, lambda i, j: eta[i, j] + f * l[i] * l[j])
        J = sp.Matrix([t, x, y, z]).jacobian(sp.Matrix([t, X1, X2, X3]))
        return J.T * gc * J

    mp.mp.dps = 30
    pts = []
    import random
    rnd = random.Random(20261203)
    for _ in range(16):
        pts.append((0.0, rnd.uniform(float(mp.log(1.5)), float(mp.log(1e4))),
                    rnd.uniform(0.02, 0.98), rnd.uniform(0.0, 6.28)))
    out = {"points": pts, "cases": []}
    for aval, hval in ((0.9375, 0.3), (0.5, 0.7), (0.0, 1.0)):
        chosen = None
        for sgn in (1, -1):
            g = metric(sgn)
            gnum = sp.lambdify((X1, X2, X3, a, h), g, "mpmath")
            G = gnum(mp.mpf(1.0), mp.mpf(0.3), mp.mpf(0.7), mp.mpf(0.5), mp.mpf(0.3))
            # physical orientation: frame dragging towards +phi, ingoing t
            if G[0, 3] < 0 and G[0, 1] > 0:
                chosen = (sgn, g)
                break
        assert chosen is not None, "no orientation with g_tphi < 0"
        g = chosen[1]
        dg = [sp.zeros(4, 4)] + [g.diff(v) for v in (X1, X2, X3)]
        fg = sp.lambdify((X1, X2, X3, a, h), g, "mpmath")
        fdg = [None] + [sp.lambdify((X1, X2, X3, a, h), d, "mpmath") for d in dg[1:]]
        case = {"a": aval, "hslope": hval, "gdd": [], "guu": [], "gamma": [], "dphi_max": 0.0}
        for p in pts:
            args = (mp.mpf(p[1]), mp.mpf(p[2]), mp.mpf(p[3]), mp.mpf(aval), mp.mpf(hval))
            gm = mp.matrix(fg(*args))
            gi = gm ** -1
            d = [mp.matrix(4, 4)] + [mp.matrix(fdg[k](*args)) for k in (1, 2, 3)]
            case["dphi_max"] = max(case["dphi_max"],
                                   float(max(abs(d[3][i, j]) for i in range(4) for j in range(4))))
            Gam = [[[mp.mpf(0)] * 4 for _ in range(4)] for _ in range(4)]
            for m_ in range(4):
                for al in range(4):
                    for be in range(4):
                        s = mp.mpf(0)
                        for de in range(4):
                            s += gi[m_, de] * (d[al][de, be] + d[be][de, al] - d[de][al, be])
                        Gam[m_][al][be] = s / 2
            case["gdd"].append([float(gm[i, j]) for i in range(4) for j in range(4)])
            case["guu"].append([float(gi[i, j]) for i in range(4) for j in range(4)])
            case["gamma"].append([float(Gam[m_][al][be]) for m_ in range(4) for al in range(4)
                                  for be in range(4)])
        out["cases"].append(case)
        print("kerr a=%g h=%g done (phi-derivative max %.2e)" % (aval, hval, case["dphi_max"]))
    with open(os.path.join(HERE, "kerr_mks.json"), "w") as fh:
        json.dump(out, fh)


def thsynch_golden():
    import numpy as np
    from scipy import integrate, special

    def Fsync(x):
        v, _ = integrate.quad(lambda t: special.kv(5.0 / 3.0, t), x, np.inf, limit=200)
        return x * v

    rows = []
    for Theta in (30.0, 100.0):
        for X in (3.0, 30.0, 300.0, 3000.0):
            for sth in (0.8, 0.3):
                K2 = special.kv(2, 1.0 / Theta)
                nu_over_nuc = X * 2.0 / 9.0 * Theta ** 2 * sth

                def integrand(lg):
                    g = np.exp(lg)          # integrate in ln(gamma)
                    beta = np.sqrt(1 - 1 / g ** 2)
                    x = 2 * nu_over_nuc / (3 * g ** 2 * sth)
                    return g * g ** 2 * beta * np.exp(-g / Theta) / (Theta * K2) * Fsync(x)

                gpk = np.sqrt(2 * nu_over_nuc / (3 * sth))
                pts = sorted(set(np.log([max(1.0 + 1e-9, v) for v in
                                         (gpk / 10, gpk, gpk * 10, Theta, 10 * Theta)])))
                I, err = integrate.quad(integrand, 0.0, np.log(200 * Theta), points=pts,
                                        limit=400, epsrel=1e-9)
                val = np.sqrt(3) * sth / 2 * I
                rows.append({"Theta": Theta, "X": X, "sin_theta": sth,
                             "j_over_ne_e2nuc_c": float(val), "K2_exact": float(K2)})
                print("thsynch Theta=%g X=%g sth=%g -> %.6e (quad err %.1e)"
                      % (Theta, X, sth, val, err), flush=True)
    with open(os.path.join(HERE, "thsynch.json"), "w") as fh:
        json.dump({"rows": rows}, fh, indent=1)


def _fsync_table():
    import numpy as np
    from scipy import integrate, interpolate, special

    xs = np.logspace(-8, np.log10(60.0), 700)
    # int_x^inf K_5/3(t) dt in ln t (K_5/3 ~ t^-5/3 near 0; K_5/3(160) < 1e-70)
    F = np.array([x * integrate.quad(lambda u: special.kv(5.0 / 3.0, np.exp(u)) * np.exp(u),
                                     np.log(x), np.log(160.0), limit=400, epsabs=0,
                                     epsrel=1e-12)[0] for x in xs])
    spl = interpolate.CubicSpline(np.log(xs), np.log(F))

    def Fsync(x):
        x = np.asarray(x, float)
        out = np.zeros_like(x)
        m = (x >= xs[0]) & (x <= xs[-1])
        out[m] = np.exp(spl(np.log(x[m])))
        lo = x < xs[0]   # F -> 2^(2/3) pi / (sqrt(3) Gamma(1/3)) x^(1/3)
        out[lo] = 2 ** (2 / 3) * np.pi / (np.sqrt(3) * special.gamma(1 / 3)) * np.cbrt(x[lo])
        return out
    return Fsync


def kappa_golden():
    import numpy as np
    from scipy import integrate

    Fsync = _fsync_table()
    kappa = 3.5
    rows = []
    for w in (0.2, 1.0, 10.0, 50.0):
        shape = lambda g: g * np.sqrt(g * g - 1) * (1 + (g - 1) / (kappa * w)) ** (-kappa - 1)
        norm, _ = integrate.quad(shape, 1.0, np.inf, limit=400, epsrel=1e-11)
        for X in (0.01, 0.1, 1.0, 10.0, 100.0, 1e3, 1e4):
            for sth in ((0.8, 0.3) if X == 100.0 else (0.8,)):
                nu_over_nuc = X * sth * (w * kappa) ** 2      # X = nu / nu_s
                def n_over(g):
                    return shape(g) / norm
                def dn_g2(g):                                  # d/dgamma (n / gamma^2)
                    q = 1 + (g - 1) / (kappa * w)
                    b = np.sqrt(g * g - 1)
                    return (q ** (-kappa - 1) / (g * g * b)
                            - b / g * (kappa + 1) / (kappa * w) * q ** (-kappa - 2)) / norm
                def xarg(g):
                    return 2 * nu_over_nuc / (3 * g * g * sth)
                lo, hi = np.log(1 + 1e-10), np.log(1e9)
                gpk = np.sqrt(2 * nu_over_nuc / (3 * sth))
                pts = sorted(set(np.log(np.clip([gpk / 30, gpk / 3, gpk, 3 * gpk, 30 * gpk,
                                                 1 + kappa * w], 1 + 1e-9, 1e8))))
                Ij, _ = integrate.quad(lambda lg: np.exp(lg) * n_over(np.exp(lg))
                                       * Fsync(xarg(np.exp(lg))), lo, hi, points=pts,
                                       limit=800, epsrel=1e-9)
                Ia, _ = integrate.quad(lambda lg: np.exp(lg) * Fsync(xarg(np.exp(lg)))
                                       * np.exp(2 * lg) * dn_g2(np.exp(lg)), lo, hi, points=pts,
                                       limit=800, epsrel=1e-9)
                J = np.sqrt(3) / 2 * Ij
                # alpha B sin / (n e) with nu = X nu_s and e B = 2 pi m c nu_c:
                # -sqrt(3) e^2 nu_c sin / (4 nu^2 m c) * Ia * B sin / e
                #   = -sqrt(3) pi sin^2 / 2 * (nu_c / nu)^2 * Ia
                A = -np.sqrt(3) * np.pi * sth * sth / 2 / nu_over_nuc ** 2 * Ia
                rows.append({"w": w, "X": X, "sin_theta": sth, "J": float(J), "A": float(A)})
                print("kappa w=%g X=%g sth=%g -> J %.6e A %.6e" % (w, X, sth, J, A), flush=True)
    with open(os.path.join(HERE, "kappa.json"), "w") as fh:
        json.dump({"kappa": kappa, "rows": rows}, fh, indent=1)


def thsynchav_golden():
    import numpy as np
    from scipy import integrate, special

    Fsync = _fsync_table()
    mus, wts = np.polynomial.legendre.leggauss(24)   # cos(theta) on [-1, 1]
    rows = []
    for Theta in (3.0, 10.0, 50.0):
        K2 = special.kv(2, 1.0 / Theta)
        for xM in (1.0, 10.0, 100.0, 1000.0):
            nu_over_nuc = xM * 1.5 * Theta ** 2
            acc = 0.0
            for mu, wt in zip(mus, wts):
                sth = np.sqrt(1 - mu * mu)
                def integrand(lg):
                    g = np.exp(lg)
                    beta = np.sqrt(1 - 1 / g ** 2)
                    return (g * g ** 2 * beta * np.exp(-g / Theta) / (Theta * K2)
                            * Fsync(2 * nu_over_nuc / (3 * g ** 2 * sth)))
                gpk = np.sqrt(2 * nu_over_nuc / (3 * sth))
                pts = sorted(set(np.log([max(1.0 + 1e-9, v) for v in
                                         (gpk / 10, gpk, gpk * 10, Theta, 10 * Theta)])))
                I, _ = integrate.quad(integrand, 0.0, np.log(300 * Theta), points=pts,
                                      limit=400, epsrel=1e-10)
                acc += 0.5 * wt * np.sqrt(3) * sth / 2 * I
            rows.append({"Theta": Theta, "x_M": xM, "j_over_ne_e2nuc_c": float(acc)})
            print("thsynchav Theta=%g x_M=%g -> %.6e" % (Theta, xM, acc), flush=True)
    with open(os.path.join(HERE, "thsynchav.json"), "w") as fh:
        json.dump({"rows": rows}, fh, indent=1)


if __name__ == "__main__":
    import sys
    what = sys.argv[1:] or ["kerr", "thsynch", "kappa", "thsynchav"]
    if "kerr" in what:
        kerr_mks_golden()
    if "thsynch" in what:
        thsynch_golden()
    if "kappa" in what:
        kappa_golden()
    if "thsynchav" in what:
        thsynchav_golden()
