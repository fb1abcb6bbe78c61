"""Synthetic HARM-format GRMHD dumps: a Fishbone-Moncrief torus on an MKS grid.

The reference ships no GRMHD data (its only dump, ``dump040`` in run.sh:5, is
missing), so every workload is generated here, seeded and reproducible.

Conventions follow what the reference's HARM reader and fluid sampler assume
(raptor_harm_model.c:90-176, :235-259):
  * MKS coordinates r = exp(X1) (R0 = 0), theta = pi X2 + (1-h)/2 sin(2 pi X2),
    X3 = phi; cell centres at startx + (i + 1/2) dx.
  * primitives KRHO, UU, U1, U2, U3, B1, B2, B3 (raptor_harm_model.h:67-74):
    rest-mass density, internal energy density, the HARM relative 4-velocity
    u~^i = u^i - u^0 g^{0i}/g^{00} and the field B^i = *F^{i0}, all in code units.

Physics (written from the papers, not from any simulation code):
  * Fishbone & Moncrief (1976) constant-l torus, l = u_phi u^t fixed by a
    circular Keplerian orbit at r_max (Bardeen, Press & Teukolsky 1972).
  * polytrope p = kappa rho^gamma, normalised to max(rho) = 1.
  * atmosphere rho = 1e-4 r^-3/2, u = 1e-6 r^-5/2 at rest in the normal frame.
  * poloidal field from A_phi = max(rho/rho_max - 0.2, 0) at cell corners,
    scaled so that max(p_gas) / max(b^2/2) = beta.
  * multiplicative log-normal noise exp(sigma N(0,1)) on rho and u; for 3D an
    m = 1..4 azimuthal modulation (1 + amp * sum_m w_m cos(m phi + psi_m)),
    sum |w_m| = 1.

Every primitive is rounded to float32 before it is returned, so an FP64
consumer and an FP32 consumer see exactly the same numbers.
"""
from __future__ import annotations

import dataclasses
import numpy as np

NPRIM = 8
KRHO, UU, U1, U2, U3, B1, B2, B3 = range(8)


@dataclasses.dataclass
class Grid:
    """HARM MKS grid header (raptor_harm_model.c:90-135)."""
    n1: int
    n2: int
    n3: int
    startx1: float
    startx2: float
    startx3: float
    dx1: float
    dx2: float
    dx3: float
    a: float
    gam: float
    hslope: float
    R0: float
    Rin: float
    Rout: float

    @property
    def stopx1(self):
        return self.startx1 + self.n1 * self.dx1

    @property
    def stopx2(self):
        return self.startx2 + self.n2 * self.dx2


@dataclasses.dataclass
class Dump:
    grid: Grid
    prims: np.ndarray          # float64 [8, n1, n2, n3], float32-representable
    t: float = 0.0
    # derived per-cell fields for the ASCII writer (u^mu, u_mu, b^mu, b_mu, gdet)
    extras: dict | None = None


def make_grid(n1, n2, n3=1, a=0.9375, hslope=0.3, Rout=40.0, gam=13.0 / 9.0):
    Rin = 1.0 + np.sqrt(1.0 - a * a)          # the horizon
    startx1 = np.log(Rin)
    dx1 = (np.log(Rout) - np.log(Rin)) / n1
    return Grid(n1=n1, n2=n2, n3=n3, startx1=float(startx1), startx2=0.0, startx3=0.0,
                dx1=float(dx1), dx2=1.0 / n2, dx3=2.0 * np.pi / n3, a=a, gam=gam,
                hslope=hslope, R0=0.0, Rin=float(Rin), Rout=Rout)


def _theta(x2, h):
    return np.pi * x2 + 0.5 * (1.0 - h) * np.sin(2.0 * np.pi * x2)


def _mks_metric(r, th, x2, a, h):
    """Covariant and contravariant MKS metric (R0 = 0) from Kerr-Schild by the
    Jacobian dr/dX1 = r, dtheta/dX2 = pi (1 + (1-h) cos 2 pi X2)."""
    s, c = np.sin(th), np.cos(th)
    S = r * r + a * a * c * c
    z = 2.0 * r / S
    J1 = r
    J2 = np.pi * (1.0 + (1.0 - h) * np.cos(2.0 * np.pi * x2))
    sh = r.shape
    g = np.zeros(sh + (4, 4))
    g[..., 0, 0] = -1.0 + z
    g[..., 0, 1] = g[..., 1, 0] = z * J1
    g[..., 0, 3] = g[..., 3, 0] = -a * z * s * s
    g[..., 1, 1] = (1.0 + z) * J1 * J1
    g[..., 1, 3] = g[..., 3, 1] = -a * (1.0 + z) * s * s * J1
    g[..., 2, 2] = S * J2 * J2
    g[..., 3, 3] = s * s * (S + a * a * s * s * (1.0 + z))
    gi = np.linalg.inv(g)
    gdet = S * np.abs(s) * J1 * J2
    return g, gi, gdet


def _kepler_l(rmax, a):
    """l = u_phi u^t of a prograde circular equatorial geodesic (BPT 1972)."""
    sr = np.sqrt(rmax)
    den = np.sqrt(rmax * rmax - 3.0 * rmax + 2.0 * a * sr)
    u_phi = (rmax * rmax - 2.0 * a * sr + a * a) / (sr * den)
    u_t_up = (rmax ** 1.5 + a) / (rmax ** 0.75 * np.sqrt(rmax ** 1.5 - 3.0 * sr + 2.0 * a))
    return u_phi * u_t_up


def _fm_lnh(r, th, l, a, rin):
    """Fishbone-Moncrief ln(h) relative to its inner edge (FM 1976, eq. 3.6)."""
    def part(r, th):
        s, c = np.sin(th), np.cos(th)
        S = r * r + a * a * c * c
        D = r * r - 2.0 * r + a * a
        A = (r * r + a * a) ** 2 - D * a * a * s * s
        q = np.sqrt(1.0 + 4.0 * l * l * S * S * D / (A * A * s * s))
        return 0.5 * np.log((1.0 + q) / (S * D / A)) - 0.5 * q - 2.0 * a * r * l / A
    return part(r, th) - part(np.asarray(rin, float), np.asarray(np.pi / 2, float))


def _fm_state(r, th, x2, grid, l, rin, kappa):
    """rho, u and BL u^t, u^phi of the torus (zero outside)."""
    a, gam = grid.a, grid.gam
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        lnh = _fm_lnh(r, th, l, a, rin)
        inside = (lnh > 0.0) & (r >= rin)
        hm1 = np.where(inside, np.expm1(np.where(inside, lnh, 0.0)), 0.0)
        rho = np.where(inside, (hm1 * (gam - 1.0) / (kappa * gam)) ** (1.0 / (gam - 1.0)), 0.0)
        u = kappa * rho ** gam / (gam - 1.0)
        s, c = np.sin(th), np.cos(th)
        S = r * r + a * a * c * c
        D = r * r - 2.0 * r + a * a
        A = (r * r + a * a) ** 2 - D * a * a * s * s
        e2 = S * S * D / (A * A * s * s)
        up1 = np.sqrt((-1.0 + np.sqrt(1.0 + 4.0 * l * l * e2)) / 2.0)
        uphi = 2.0 * a * r * np.sqrt(1.0 + up1 * up1) / np.sqrt(A * S * D) + np.sqrt(S / A) * up1 / s
    return rho, u, np.where(inside, uphi, 0.0), inside


def _bl_ut(r, th, uphi, a):
    """u^t from u.u = -1 with u^r = u^theta = 0 in Boyer-Lindquist."""
    s, c = np.sin(th), np.cos(th)
    S = r * r + a * a * c * c
    gtt = -(1.0 - 2.0 * r / S)
    gtp = -2.0 * a * r * s * s / S
    gpp = (r * r + a * a + 2.0 * a * a * r * s * s / S) * s * s
    # gtt ut^2 + 2 gtp ut up + gpp up^2 + 1 = 0, root with ut > 0
    A_, B_, C_ = gtt, 2.0 * gtp * uphi, gpp * uphi * uphi + 1.0
    disc = np.sqrt(B_ * B_ - 4.0 * A_ * C_)
    return (-B_ - disc) / (2.0 * A_)


def _fluid_frame(prims, g, gi):
    """u^mu, u_mu, b^mu, b_mu and b^2 from primitives, by the relations the
    reference's sampler uses (raptor_harm_model.c:235-259)."""
    V = prims[U1:U3 + 1]
    Bp = prims[B1:B3 + 1]
    gs = g[..., 1:, 1:]
    VdotV = np.einsum("...ij,i...,j...->...", gs, V, V)
    gam = np.sqrt(1.0 + np.abs(VdotV))
    alpha = 1.0 / np.sqrt(-gi[..., 0, 0])
    ucon = np.empty((4,) + V.shape[1:])
    ucon[0] = gam / alpha
    for i in range(1, 4):
        ucon[i] = V[i - 1] - gam * alpha * gi[..., 0, i]
    ucov = np.einsum("...ij,j...->i...", g, ucon)
    bcon = np.empty_like(ucon)
    bcon[0] = np.einsum("i...,i...->...", ucov[1:], Bp)
    for i in range(1, 4):
        bcon[i] = (Bp[i - 1] + ucon[i] * bcon[0]) / ucon[0]
    bcov = np.einsum("...ij,j...->i...", g, bcon)
    bsq = np.einsum("i...,i...->...", bcon, bcov)
    return ucon, ucov, bcon, bcov, bsq


def fishbone_moncrief(n1, n2, n3=1, *, a=0.9375, hslope=0.3, Rout=40.0, gam=13.0 / 9.0,
                      rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, noise=0.02, seed=12345,
                      m_amp=0.1, rho_floor=1e-4, u_floor=1e-6):
    """Build a seeded torus dump.  n3 > 1 adds the m = 1..4 modulation."""
    grid = make_grid(n1, n2, n3, a=a, hslope=hslope, Rout=Rout, gam=gam)
    rng = np.random.default_rng(seed)
    i = np.arange(n1)
    j = np.arange(n2)
    X1 = grid.startx1 + (i + 0.5) * grid.dx1
    X2 = grid.startx2 + (j + 0.5) * grid.dx2
    x1, x2 = np.meshgrid(X1, X2, indexing="ij")
    r = np.exp(x1)
    th = _theta(x2, hslope)
    l = _kepler_l(rmax, a)
    rho, u, uphi, inside = _fm_state(r, th, x2, grid, l, rin, kappa)
    ut = np.where(inside, _bl_ut(r, th, uphi, a), 1.0)
    rhomax = rho.max()
    rho = rho / rhomax
    u = u / rhomax
    g, gi, gdet = _mks_metric(r, th, x2, a, hslope)
    # u^mu in MKS equals BL (u^t, 0, 0, u^phi) because u^r = 0; to primitives
    prim2 = np.zeros((NPRIM, n1, n2))
    prim2[U1] = np.where(inside, -ut * gi[..., 0, 1] / gi[..., 0, 0], 0.0)
    prim2[U2] = np.where(inside, -ut * gi[..., 0, 2] / gi[..., 0, 0], 0.0)
    prim2[U3] = np.where(inside, uphi - ut * gi[..., 0, 3] / gi[..., 0, 0], 0.0)
    prim2[KRHO] = np.maximum(rho, rho_floor * r ** -1.5)
    prim2[UU] = np.maximum(u, u_floor * r ** -2.5)
    # poloidal field from A_phi on cell corners
    X1c = grid.startx1 + np.arange(n1 + 1) * grid.dx1
    X2c = grid.startx2 + np.arange(n2 + 1) * grid.dx2
    x1c, x2c = np.meshgrid(X1c, X2c, indexing="ij")
    rc, thc = np.exp(x1c), _theta(x2c, hslope)
    rhoc, _, _, _ = _fm_state(rc, thc, x2c, grid, l, rin, kappa)
    Aphi = np.maximum(rhoc / rhomax - 0.2, 0.0)
    prim2[B1] = -(Aphi[:-1, 1:] - Aphi[:-1, :-1] + Aphi[1:, 1:] - Aphi[1:, :-1]) / (2.0 * grid.dx2 * gdet)
    prim2[B2] = (Aphi[1:, :-1] - Aphi[:-1, :-1] + Aphi[1:, 1:] - Aphi[:-1, 1:]) / (2.0 * grid.dx1 * gdet)
    prim2[B3] = 0.0
    _, _, _, _, bsq = _fluid_frame(prim2, g, gi)
    bsq_max = bsq.max()
    if bsq_max > 0:
        beta_act = (gam - 1.0) * prim2[UU].max() / (0.5 * bsq_max)
        prim2[B1:B3 + 1] *= np.sqrt(beta_act / beta)
    prims = np.repeat(prim2[..., None], n3, axis=3)
    if n3 > 1:
        phi = grid.startx3 + (np.arange(n3) + 0.5) * grid.dx3
        w = rng.uniform(0.2, 1.0, 4)
        w /= w.sum()
        psi = rng.uniform(0.0, 2.0 * np.pi, 4)
        mod = 1.0 + m_amp * sum(w[m - 1] * np.cos(m * phi + psi[m - 1]) for m in range(1, 5))
        prims[KRHO] *= mod[None, None, :]
        prims[UU] *= mod[None, None, :]
    if noise > 0:
        prims[KRHO] *= np.exp(noise * rng.standard_normal(prims[KRHO].shape))
        prims[UU] *= np.exp(noise * rng.standard_normal(prims[UU].shape))
    prims = prims.astype(np.float32).astype(np.float64)
    ucon, ucov, bcon, bcov, _ = _fluid_frame(prims[..., 0], g, gi)
    extras = dict(ucon=ucon, ucov=ucov, bcon=bcon, bcov=bcov, gdet=gdet, x1=x1, x2=x2, r=r, th=th)
    return Dump(grid=grid, prims=np.ascontiguousarray(prims), extras=extras)
