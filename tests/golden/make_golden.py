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

Run:  python tests/golden/make_golden.py [kerr] [thsynch]   (about two minutes)
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


if __name__ == "__main__":
    import sys
    what = sys.argv[1:] or ["kerr", "thsynch"]
    if "kerr" in what:
        kerr_mks_golden()
    if "thsynch" in what:
        thsynch_golden()
