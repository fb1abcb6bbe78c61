"""ctypes binding of liboracle.so, the FP64 CPU restatement (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product (raptor_amd) never does.
"""
from __future__ import annotations

import ctypes
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrParams(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int), ("n2", ctypes.c_int), ("n3", ctypes.c_int),
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("max_steps", ctypes.c_int),
        ("startx1", ctypes.c_double), ("startx2", ctypes.c_double), ("startx3", ctypes.c_double),
        ("dx1", ctypes.c_double), ("dx2", ctypes.c_double), ("dx3", ctypes.c_double),
        ("a", ctypes.c_double), ("hslope", ctypes.c_double), ("R0", ctypes.c_double),
        ("MBH", ctypes.c_double), ("M_unit", ctypes.c_double),
        ("R_high", ctypes.c_double), ("R_low", ctypes.c_double),
        ("cam_size_x", ctypes.c_double), ("cam_size_y", ctypes.c_double),
        ("inclination", ctypes.c_double), ("stepsize", ctypes.c_double),
        ("integrator", ctypes.c_int), ("emission", ctypes.c_int),
    ]


# int_method of parameters.h:186-189
RK2, RK4, VERLET = 0, 1, 2
# emission models (or_emission): the thermal fit of the path, the angle-averaged
# thermal fit and the kappa fits of j_nu.c
THSYNCH, THSYNCHAV, KAPPA = 0, 1, 2


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(OrParams)
        L.or_metric_dd.argtypes = [P, _dp, _dp]
        L.or_metric_uu.argtypes = [P, _dp, _dp]
        L.or_connection.argtypes = [P, _dp, _dp]
        L.or_rk2_step.argtypes = [P, _dp, ctypes.c_double]
        L.or_step.argtypes = [P, _dp, ctypes.c_double]
        L.or_stepsize.argtypes = [P, _dp, _dp]
        L.or_stepsize.restype = ctypes.c_double
        L.or_theta_to_x2.argtypes = [ctypes.c_double, ctypes.c_double]
        L.or_theta_to_x2.restype = ctypes.c_double
        L.or_init_photon.argtypes = [P, ctypes.c_double, ctypes.c_double, _dp]
        L.or_j_thsynch.argtypes = [ctypes.c_double] * 5
        L.or_j_thsynch.restype = ctypes.c_double
        L.or_planck.argtypes = [ctypes.c_double] * 2
        L.or_planck.restype = ctypes.c_double
        L.or_a_th.argtypes = [ctypes.c_double] * 3
        L.or_a_th.restype = ctypes.c_double
        L.or_hyp2f1.argtypes = [ctypes.c_double] * 4
        L.or_hyp2f1.restype = ctypes.c_double
        for fn in (L.or_j_kappa, L.or_a_kappa):
            fn.argtypes = [ctypes.c_double] * 5
            fn.restype = ctypes.c_double
        L.or_j_thsynchav.argtypes = [ctypes.c_double] * 4
        L.or_j_thsynchav.restype = ctypes.c_double
        L.or_emission.argtypes = [ctypes.c_int] + [ctypes.c_double] * 5 + [_dp, _dp]
        L.or_jansky_factor.argtypes = [P]
        L.or_jansky_factor.restype = ctypes.c_double
        L.or_units.argtypes = [P, _dp]
        L.or_fluid_at.argtypes = [P, _dp, _dp, _dp]
        L.or_fluid_at.restype = ctypes.c_int
        L.or_sample_at.argtypes = [P, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.or_sample_at.restype = ctypes.c_int
        L.or_ray_path.argtypes = [P, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int]
        L.or_ray_path.restype = ctypes.c_int
        L.or_render.argtypes = [P, _dp, ctypes.c_int, _dp, _ip, ctypes.c_int, _dp, _dp, _ip,
                                _dp, ctypes.c_int]
        L.or_render.restype = ctypes.c_int
        L.or_transfer_step.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, _dp, _dp]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def params(grid, *, width, height, inclination, M_unit, R_high, R_low=1.0, MBH=8.9505e39,
           cam_size=40.0, stepsize=0.03, max_steps=100000, integrator=RK2, emission=THSYNCH):
    return OrParams(n1=grid.n1, n2=grid.n2, n3=grid.n3, width=width, height=height,
                    max_steps=max_steps, startx1=grid.startx1, startx2=grid.startx2,
                    startx3=grid.startx3, dx1=grid.dx1, dx2=grid.dx2, dx3=grid.dx3, a=grid.a,
                    hslope=grid.hslope, R0=grid.R0, MBH=MBH, M_unit=M_unit, R_high=R_high,
                    R_low=R_low, cam_size_x=cam_size, cam_size_y=cam_size,
                    inclination=inclination, stepsize=stepsize, integrator=integrator,
                    emission=emission)


def emission(model, B, mu, Te, nu_p, ne):
    """(j_nu, alpha_nu) in cgs of one sample for an emission model."""
    j, a = ctypes.c_double(), ctypes.c_double()
    lib().or_emission(model, B, mu, Te, nu_p, ne, ctypes.byref(j), ctypes.byref(a))
    return j.value, a.value


def metric_dd(P, X):
    g = np.zeros(16)
    lib().or_metric_dd(ctypes.byref(P), _d(np.ascontiguousarray(X, float)), _d(g))
    return g.reshape(4, 4)


def metric_uu(P, X):
    g = np.zeros(16)
    lib().or_metric_uu(ctypes.byref(P), _d(np.ascontiguousarray(X, float)), _d(g))
    return g.reshape(4, 4)


def connection(P, X):
    G = np.zeros(64)
    lib().or_connection(ctypes.byref(P), _d(np.ascontiguousarray(X, float)), _d(G))
    return G.reshape(4, 4, 4)


def rk2_step(P, y, dl):
    y = np.array(y, float)
    lib().or_rk2_step(ctypes.byref(P), _d(y), dl)
    return y


def step(P, y, dl):
    """One step of the integrator P.integrator selects (RK2, RK4, VERLET)."""
    y = np.array(y, float)
    lib().or_step(ctypes.byref(P), _d(y), dl)
    return y


def stepsize(P, X, U):
    return lib().or_stepsize(ctypes.byref(P), _d(np.ascontiguousarray(X, float)),
                             _d(np.ascontiguousarray(U, float)))


def init_photon(P, alpha, beta):
    y = np.zeros(8)
    lib().or_init_photon(ctypes.byref(P), alpha, beta, _d(y))
    return y


def units(P):
    out = np.zeros(6)
    lib().or_units(ctypes.byref(P), _d(out))
    return dict(zip(["L_unit", "RHO_unit", "B_unit", "Ne_unit", "Rg", "cutoff_inner"], out))


def fluid_at(P, prims, X):
    out = np.zeros(13)
    p = np.ascontiguousarray(prims, float)
    ok = lib().or_fluid_at(ctypes.byref(P), _d(p), _d(np.ascontiguousarray(X, float)), _d(out))
    return dict(ne=out[0], thetae=out[1], B=out[2], ucon=out[3:7], bcon=out[7:11],
                dthetae_dRhigh=out[11], ok=bool(ok))


def sample_at(P, prims, y, freqs):
    """One transfer sample at y = (X, k) composed as the oracle's ray loop
    composes it: dict(ok, ne, thetae, B, mu, nu_p[nf], j[nf], alpha[nf])."""
    freqs = np.ascontiguousarray(np.atleast_1d(freqs), float)
    out = np.zeros(5 + 3 * freqs.size)
    p = np.ascontiguousarray(prims, float)
    lib().or_sample_at(ctypes.byref(P), _d(p), _d(np.ascontiguousarray(y, float)), freqs.size,
                       _d(freqs), _d(out))
    q = out[5:].reshape(-1, 3)
    return dict(ok=bool(out[0]), ne=out[1], thetae=out[2], B=out[3], mu=out[4],
                nu_p=q[:, 0], j=q[:, 1], alpha=q[:, 2])


def ray_path(P, alpha, beta, maxpts=200000):
    out = np.zeros((maxpts, 8))
    n = lib().or_ray_path(ctypes.byref(P), alpha, beta, _d(out), maxpts)
    return out[:n]


def render(P, prims, freqs, pixels=None, grad=False, counters=False, gate=False, nthreads=0):
    """Trace pixels (all by default).  Returns image [npix, nf] (I_nu, cgs),
    optionally dimage [npix, nf, 2], counters [npix, 6] (steps, samples in the
    grid, samples past the sigma cut, escaped, steps before the first one that
    lands inside the grid or -1, steps taken when every frequency's tau had
    passed ln 1000 or -1) and gate [npix, nf]: the
    distance of tau from ln 1000 at the sample where it crosses that gate
    (inf if it never does)."""
    p = np.ascontiguousarray(prims, float)
    freqs = np.ascontiguousarray(np.atleast_1d(freqs), float)
    nf = freqs.size
    if pixels is None:
        npix = P.width * P.height
        pix_ptr = None
    else:
        pixels = np.ascontiguousarray(pixels, dtype=np.int32)
        npix = pixels.size
        pix_ptr = pixels.ctypes.data_as(_ip)
    img = np.zeros((npix, nf))
    dimg = np.zeros((npix, nf, 2)) if grad else None
    cnt = np.zeros((npix, 6), dtype=np.int32) if counters else None
    gt = np.zeros((npix, nf)) if gate else None
    rc = lib().or_render(ctypes.byref(P), _d(p), nf, _d(freqs), pix_ptr, npix, _d(img),
                         _d(dimg) if grad else None,
                         cnt.ctypes.data_as(_ip) if counters else None,
                         _d(gt) if gate else None, nthreads)
    if rc != 0:
        raise RuntimeError("or_render failed (%d)" % rc)
    out = [img]
    if grad:
        out.append(dimg)
    if counters:
        out.append(cnt)
    if gate:
        out.append(gt)
    return out[0] if len(out) == 1 else tuple(out)
