"""HARM dump I/O.

Two formats:

* The 2D HARM ASCII dump the reference reads (raptor_harm_model.c:90-176):
  a 26-token header
      t n1 n2 startx1 startx2 dx1 dx2 tf nstep a gam cour DTd DTl DTi DTr
      dump_cnt image_cnt rdump_cnt dt lim failed Rin Rout hslope R0
  then n1*n2 records, record k = i*n2 + j (X2 index fastest), 34 numbers each:
      x1 x2 r h | rho u U1 U2 U3 B1 B2 B3 | divb | Ucon[4] | Ucov[4] |
      Bcon[4] | Bcov[4] | vmin vmax vmin vmax | gdet
  Only the 8 primitives reach the image; the rest feed the reference's
  accretion-rate print (raptor_harm_model.c:171-182).

* A binary primitive file for any N3 (the reference has no 3D format): a
  fixed little-endian header (``PRIM_HEADER``) and float32 primitives in the
  reference's index order p[v][i][j][k] (structure of arrays).  The C host
  driver reads the same layout (raptor_amd/host/raptor_io.c).
"""
from __future__ import annotations

import struct
import numpy as np

from .torus import Dump, Grid, NPRIM

PRIM_MAGIC = b"RPTRPRIM"
# magic, version, n1, n2, n3, pad, then 12 doubles:
# startx1 startx2 startx3 dx1 dx2 dx3 a gam hslope R0 Rin Rout
PRIM_HEADER = struct.Struct("<8siiiii12d")


def write_harm_ascii(path, dump: Dump, t=0.0):
    g = dump.grid
    if g.n3 != 1:
        raise ValueError("the HARM ASCII format is 2D (n3 must be 1)")
    ex = dump.extras or {}
    n1, n2 = g.n1, g.n2
    hdr = [t, n1, n2, g.startx1, g.startx2, g.dx1, g.dx2, 0.0, 0, g.a, g.gam, 0.9,
           0.5, 0.5, 1.0, 10, 0, 0, 0, 1e-3, 0, 0, g.Rin, g.Rout, g.hslope, g.R0]
    p = dump.prims[..., 0]
    cols = np.zeros((n1 * n2, 34))
    if "x1" in ex:
        cols[:, 0] = ex["x1"].ravel()
        cols[:, 1] = ex["x2"].ravel()
        cols[:, 2] = ex["r"].ravel()
        cols[:, 3] = ex["th"].ravel()
    for v in range(NPRIM):
        cols[:, 4 + v] = p[v].ravel()
    for name, off in (("ucon", 13), ("ucov", 17), ("bcon", 21), ("bcov", 25)):
        if name in ex:
            cols[:, off:off + 4] = np.moveaxis(ex[name], 0, -1).reshape(-1, 4)
    if "gdet" in ex:
        cols[:, 33] = ex["gdet"].ravel()
    with open(path, "w") as fh:
        fh.write(" ".join(_fmt_hdr(v) for v in hdr) + "\n")
        # %.17g round-trips every double; the primitives are float32 values, so an
        # FP64 reader and an FP32 reader see exactly the same numbers
        np.savetxt(fh, cols, fmt="%.17g")


def _fmt_hdr(v):
    if isinstance(v, (int, np.integer)):
        return "%d" % v
    return "%.17g" % v


def read_harm_ascii(path) -> Dump:
    """Read a 2D HARM ASCII dump the way raptor_harm_model.c:90-169 does."""
    with open(path) as fh:
        toks = fh.readline().split()
        if len(toks) < 26:
            raise ValueError("HARM header needs 26 tokens, got %d" % len(toks))
        body = np.loadtxt(fh)
    t = float(toks[0])
    n1, n2 = int(toks[1]), int(toks[2])
    startx1, startx2, dx1, dx2 = map(float, toks[3:7])
    a, gam = float(toks[9]), float(toks[10])
    Rin, Rout, hslope, R0 = map(float, toks[22:26])
    body = body.reshape(-1, 34)
    if body.shape[0] != n1 * n2:
        raise ValueError("dump has %d records, header says %d x %d" % (body.shape[0], n1, n2))
    prims = np.empty((NPRIM, n1, n2, 1))
    for v in range(NPRIM):
        prims[v, :, :, 0] = body[:, 4 + v].reshape(n1, n2)
    grid = Grid(n1=n1, n2=n2, n3=1, startx1=startx1, startx2=startx2, startx3=0.0, dx1=dx1,
                dx2=dx2, dx3=2.0 * np.pi, a=a, gam=gam, hslope=hslope, R0=R0, Rin=Rin, Rout=Rout)
    return Dump(grid=grid, prims=prims, t=t)


def write_prims_bin(path, dump: Dump):
    g = dump.grid
    with open(path, "wb") as fh:
        fh.write(PRIM_HEADER.pack(PRIM_MAGIC, 1, g.n1, g.n2, g.n3, 0, g.startx1, g.startx2,
                                  g.startx3, g.dx1, g.dx2, g.dx3, g.a, g.gam, g.hslope, g.R0,
                                  g.Rin, g.Rout))
        fh.write(np.ascontiguousarray(dump.prims, dtype="<f4").tobytes())


def read_prims_bin(path) -> Dump:
    with open(path, "rb") as fh:
        raw = fh.read(PRIM_HEADER.size)
        f = PRIM_HEADER.unpack(raw)
        if f[0] != PRIM_MAGIC or f[1] != 1:
            raise ValueError("not a primitive file (bad magic/version)")
        n1, n2, n3 = f[2], f[3], f[4]
        (startx1, startx2, startx3, dx1, dx2, dx3, a, gam, hslope, R0, Rin, Rout) = f[6:]
        data = np.frombuffer(fh.read(), dtype="<f4")
    if data.size != NPRIM * n1 * n2 * n3:
        raise ValueError("primitive file truncated")
    grid = Grid(n1=n1, n2=n2, n3=n3, startx1=startx1, startx2=startx2, startx3=startx3, dx1=dx1,
                dx2=dx2, dx3=dx3, a=a, gam=gam, hslope=hslope, R0=R0, Rin=Rin, Rout=Rout)
    return Dump(grid=grid, prims=data.astype(np.float64).reshape(NPRIM, n1, n2, n3))
